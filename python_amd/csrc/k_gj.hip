// k_gj.hip -- batched blocked Gauss-Jordan inverse on the matrix cores (FP64 MFMA), for the
// dense macro-atom matrices that invert_matrix receives (source/matrix_cpu.c:223-257, called per
// cell at source/macro_accelerate.c:315 on N = I - Q, order nlevels_macro + 1).
//
// The row-major input A, read as a column-major matrix, is M = A^T; the inverse of M written back
// in the same place, read row-major again, is A^-1.  M is inverted in place by the blocked
// Gauss-Jordan sweep over 64 x 64 blocks, every cell of a size class in lock step.  A step pivots
// on the w diagonal blocks S (w = 4, or fewer at the end of an order that is not a multiple of
// 256), O = every other block:
//   1. M[S,S] <- D = M[S,S]^-1        the same sweep over S's own blocks in steps of w/2, down to
//                                     k_gj_diag (one 64 x 64 block, in registers)
//   2. M[S,O] <- D M[S,O]              k_gj_tile<2, w>
//   3. M[O,O] <- M[O,O] - M[O,S] M[S,O] k_gj_tile<4, w>  (M[O,S] still the old values)
//   4. M[O,S] <- -M[O,S] D             k_gj_tile<5, w>
// After the last step M holds M^-1.  Step 3 reads and writes every other block once per step, so
// its HBM traffic per flop falls as 1/w.  Every product is a 64 x 64 output block with K = 64 w on
// v_mfma_f64_16x16x4f64 (four waves, a 32 x 32 quarter each, operands staged through LDS in slabs
// of 32 k; lane l holds A[l % 16][l / 16] and B[l / 16][l % 16] of a 16 x 4 x 16 step and element
// e of its result at row l / 16 + 4 e, column l % 16).  The sweep does not pivot: it runs on cells
// whose A is row diagonally dominant
// (|a_ii| >= sum_j!=i |a_ij|, which N = I - Q with substochastic Q is).  M is then column
// diagonally dominant, every Schur complement stays so, and partial pivoting would never swap a
// row; k_gj_prepare hands every other cell to the GSL-order pivoting kernel (k_invert_f64).  An
// exact zero pivot (a zero column of M: A singular) gives code 1 (GSL_EDOM), as the reference.
// The inverse is not in GSL's operation order: parity is tolerance-based (relative Frobenius
// error against the FP64 oracle).
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

#include <algorithm>
#include <cstdlib>

#define GJ_NB 64                  // block order
#ifndef GJ_KC
#define GJ_KC 32                  // K rows staged in LDS per slab
#endif
#ifndef GJ_OCC
#define GJ_OCC 4                  // tile workgroups per CU (launch bounds)
#endif
#define GJ_NE (GJ_KC / 4)         // elements of each operand slab per thread
#define GJ_PA (GJ_NB + 4)         // LDS pitch of the A slab (k-major: [k][row])
#define GJ_PB (GJ_KC + 2)         // LDS pitch of the B slab (column-major: [col][k])

typedef double gj_d4 __attribute__ ((ext_vector_type (4)));

// ---- prepare: copy A into the output, row dominance test, hand-off list ----
__global__ __launch_bounds__ (256) void k_gj_prepare (GjArgs g)
{
  const int cell = g.list[blockIdx.x];
  const int n = g.nsz;
  const double *a = g.A + g.offA[cell];
  double *m = g.M + g.offA[cell];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int s_bad;
  if (threadIdx.x == 0)
    s_bad = 0;
  __syncthreads ();
  int bad = 0;
  // eight rows per pass (n is a multiple of 64, so of 32 rows per wave), so that eight loads are
  // in flight per column step and the eight row reductions interleave
  constexpr int PR = 8;
  for (int r0 = wv * PR; r0 < n; r0 += 4 * PR)
  {
    double off[PR], dg[PR];
#pragma unroll
    for (int q = 0; q < PR; q++)
      off[q] = dg[q] = 0.0;
    for (int c = lane; c < n; c += 64)
    {
      double v[PR];
#pragma unroll
      for (int q = 0; q < PR; q++)
        v[q] = a[(size_t) (r0 + q) * n + c];
#pragma unroll
      for (int q = 0; q < PR; q++)
      {
        m[(size_t) (r0 + q) * n + c] = v[q];
        if (c == r0 + q)
          dg[q] = fabs (v[q]);
        else
          off[q] += fabs (v[q]);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int q = 0; q < PR; q++)
      {
        off[q] += __shfl_xor (off[q], o);
        dg[q] += __shfl_xor (dg[q], o);
      }
#pragma unroll
    for (int q = 0; q < PR; q++)
      bad |= !(dg[q] >= off[q]);  // NaN counts as not dominant
  }
  if (lane == 0 && bad)
    atomicOr (&s_bad, 1);
  __syncthreads ();
  if (threadIdx.x == 0)
  {
    g.codes[cell] = 0;
    g.flag[cell] = s_bad ? 1 : 0;
    if (s_bad)
    {
      const int pos = atomicAdd (g.fb, 1);
      g.fb[1 + pos] = cell;
    }
  }
}

// ---- step 1: D = M[S,S]^-1 by the unblocked sweep (no pivoting) ----
// Thread (column j = lane, rows r0 .. r0+15 with r0 = 16 wave) keeps its 16 elements in registers.
// Step k publishes column k (its owners) and row k (the wave holding that row) to LDS slots of the
// step's parity, so one barrier per step suffices: a thread can only overwrite a slot two steps
// later, after every thread has passed the barrier of the step in between.
__global__ __launch_bounds__ (256) void k_gj_diag (GjArgs g)
{
  const int cell = g.list[blockIdx.x];
  if (g.flag[cell])
    return;
  const int n = g.nsz;
  const int k0 = g.sb * GJ_NB;
  double *m = g.M + g.offA[cell];
  __shared__ double s_col[2][GJ_NB];       // column k (by row), by step parity
  __shared__ double s_row[2][GJ_NB];       // row k (by column)
  const int tid = threadIdx.x;
  const int j = tid & 63;                  // this thread's column
  const int wv = tid >> 6;
  const int r0 = wv * 16;                  // and its 16 rows
  double b[16];
  {
    const double *src = m + (size_t) (k0 + j) * n + k0 + r0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      b[i] = src[i];
  }
#pragma unroll
  for (int k = 0; k < GJ_NB; k++)
  {
    const int buf = k & 1;
    if (j == k)
    {
#pragma unroll
      for (int i = 0; i < 16; i++)
        s_col[buf][r0 + i] = b[i];
    }
    if (wv == k / 16)
      s_row[buf][j] = b[k % 16];
    __syncthreads ();
    const double piv = s_col[buf][k];
    if (piv == 0.0)
    {
      // an exact zero pivot: M has a zero column (the block's Schur complement of a column
      // diagonally dominant matrix is zero only there); every thread sees the same value
      if (tid == 0)
      {
        g.flag[cell] = 2;
        g.codes[cell] = 1;
      }
      return;
    }
    const double inv = 1.0 / piv;
    double f[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
      f[i] = s_col[buf][r0 + i];           // column k, this thread's rows
    if (j == k)
    {
#pragma unroll
      for (int i = 0; i < 16; i++)
        b[i] = (r0 + i == k) ? inv : -f[i] * inv;
    }
    else
    {
      const double uj = s_row[buf][j] * inv;   // row k, column j
#pragma unroll
      for (int i = 0; i < 16; i++)
        b[i] = (r0 + i == k) ? uj : fma (-f[i], uj, b[i]);
    }
  }
  double *dst = m + (size_t) (k0 + j) * n + k0 + r0;
#pragma unroll
  for (int i = 0; i < 16; i++)
    dst[i] = b[i];
}

// Workgroups are dispatched to the eight XCDs round robin; number them so that consecutive tiles
// (the same cell's, which share the step's row and column panels) run on the same XCD and its L2.
__device__ __forceinline__ int xcd_item (int bid, int G)
{
  const int x = bid & 7, i = bid >> 3;
  const int q0 = G >> 3, rem = G & 7;
  return x < rem ? x * (q0 + 1) + i : rem * (q0 + 1) + (x - rem) * q0 + i;
}

// ---- steps 2-4: block products over the sweep's block range ----
// A step pivots on the W = g.sw diagonal blocks S = sb .. sb+W-1 (one or two; with two, M[S,S]
// has already been inverted in place by a two-step sweep of its 2 x 2 blocks), over the blocks
// O = [r0, r0 + nr) minus S.  Block rows and columns of O are enumerated by
// o -> r0 + o (+ W past S).  One workgroup forms RB x CB output blocks with K = 64 W:
//   MODE 2: C(S, tj) = D C(S, tj)                (RB = W, CB = 1; one workgroup per tj in O)
//   MODE 4: C(ti, tj) -= M(ti, S) M(S, tj)       (RB = CB = 1; one per (ti, tj) in O x O)
//   MODE 5: C(ti, S) = -C(ti, S) D               (RB = 1, CB = W; one per ti in O)
// MODES 2 and 5 read the blocks they overwrite, so one workgroup (four waves per block) computes
// all W of them before it writes any (W = g.sw is a template parameter).  Operands pass through
// LDS in slabs of GJ_KC k-rows, the next slab's global loads in flight while the current one runs
// on the matrix cores; each wave forms a 32 x 32 quarter of one output block.
__device__ __forceinline__ int gj_oblock (int o, int r0, int sb, int w)
{
  const int b = r0 + o;
  return b >= sb ? b + w : b;
}

template <int MODE, int W>
__global__ __launch_bounds__ (256 * (MODE == 4 ? 1 : W), (MODE == 4 || W == 1) ? GJ_OCC : (W == 2 ? 2 : 1)) void k_gj_tile (GjArgs g)
{
  constexpr int RB = MODE == 2 ? W : 1, CB = MODE == 5 ? W : 1;
  constexpr int NT = 256 * RB * CB;               // four waves per output block
  constexpr int NRA = GJ_KC * GJ_NB * RB / NT;    // A elements per thread and slab
  constexpr int NRB = GJ_KC * GJ_NB * CB / NT;    // B elements per thread and slab
  constexpr int PA = GJ_NB * RB + 4;              // LDS pitch of the A slab ([k][row])
  constexpr int w = W;                            // g.sw
  const int no = g.nr - w;                        // blocks of O
  const int tpc = MODE == 4 ? no * no : no;
  const int item = xcd_item (blockIdx.x, gridDim.x);
  const int ci = item / tpc, t = item - ci * tpc;
  const int cell = g.list[ci];
  if (g.flag[cell])
    return;
  const int n = g.nsz, sb = g.sb;
  int ai, bj;                                     // first block row of A, first block column of B
  if (MODE == 4)
  {
    const int oi = t / no;
    ai = gj_oblock (oi, g.r0, sb, w);
    bj = gj_oblock (t - oi * no, g.r0, sb, w);
  }
  else if (MODE == 2)
  {
    ai = sb;
    bj = gj_oblock (t, g.r0, sb, w);
  }
  else
  {
    ai = gj_oblock (t, g.r0, sb, w);
    bj = sb;
  }
  constexpr int K = GJ_NB * W;
  const double sgn = MODE == 2 ? 1.0 : -1.0;
  double *m = g.M + g.offA[cell];
  __shared__ double As[GJ_KC][PA];                // [k][row] of the left operand, sign applied
  __shared__ double Bs[GJ_NB * CB][GJ_PB];        // [col][k] of the right operand
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int pb = wv >> 2;                         // this wave's output block
  const int pr = MODE == 2 ? pb : 0, pc = MODE == 5 ? pb : 0;
  const int rw = ((wv & 3) >> 1) * 32, cw = (wv & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  // A(k, row) = abase[k n + row] (rows of the RB blocks from ai on), B(col, k) = bbase[col n + k]
  // (columns of the CB blocks from bj on); this thread's slab elements are idx = tid + NT e:
  // A (k = idx / (64 RB), row = idx % (64 RB)), B (col = idx / GJ_KC, k = idx % GJ_KC)
  const double *abase = m + (size_t) (sb * GJ_NB) * n + ai * GJ_NB;
  const double *bbase = m + (size_t) (bj * GJ_NB) * n + sb * GJ_NB;
  double ra[NRA], rb[NRB];
#pragma unroll
  for (int e = 0; e < NRA; e++)
  {
    const int idx = tid + NT * e;
    ra[e] = abase[(size_t) (idx / (GJ_NB * RB)) * n + (idx % (GJ_NB * RB))];
  }
#pragma unroll
  for (int e = 0; e < NRB; e++)
  {
    const int idx = tid + NT * e;
    rb[e] = bbase[(size_t) (idx / GJ_KC) * n + (idx % GJ_KC)];
  }

  gj_d4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      if (MODE == 4)
      {
        // v_mfma_f64_16x16x4f64 keeps element e of lane l at row l/16 + 4 e, column l % 16
        const double *src = m + (size_t) (bj * GJ_NB + cw + 16 * bq + lr) * n + ai * GJ_NB + rw + 16 * bi + lk;
        acc[bi][bq] = gj_d4{ src[0], src[4], src[8], src[12] };
      }
      else
        acc[bi][bq] = gj_d4{ 0.0, 0.0, 0.0, 0.0 };
    }

#pragma unroll
  for (int kh = 0; kh < K; kh += GJ_KC)
  {
#pragma unroll
    for (int e = 0; e < NRA; e++)
    {
      const int idx = tid + NT * e;
      As[idx / (GJ_NB * RB)][idx % (GJ_NB * RB)] = sgn * ra[e];
    }
#pragma unroll
    for (int e = 0; e < NRB; e++)
    {
      const int idx = tid + NT * e;
      Bs[idx / GJ_KC][idx % GJ_KC] = rb[e];
    }
    __syncthreads ();
    if (kh + GJ_KC < K)
    {
      // the next slab's loads, in flight during this slab's products
#pragma unroll
      for (int e = 0; e < NRA; e++)
      {
        const int idx = tid + NT * e;
        ra[e] = abase[(size_t) (kh + GJ_KC + idx / (GJ_NB * RB)) * n + (idx % (GJ_NB * RB))];
      }
#pragma unroll
      for (int e = 0; e < NRB; e++)
      {
        const int idx = tid + NT * e;
        rb[e] = bbase[(size_t) (idx / GJ_KC) * n + kh + GJ_KC + (idx % GJ_KC)];
      }
    }
#pragma unroll
    for (int k = 0; k < GJ_KC; k += 4)
    {
      double av[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        av[bi] = As[k + lk][GJ_NB * pr + rw + 16 * bi + lr];
#pragma unroll
      for (int bq = 0; bq < 2; bq++)
        bv[bq] = Bs[GJ_NB * pc + cw + 16 * bq + lr][k + lk];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bq = 0; bq < 2; bq++)
          acc[bi][bq] = __builtin_amdgcn_mfma_f64_16x16x4f64 (av[bi], bv[bq], acc[bi][bq], 0, 0, 0);
    }
    __syncthreads ();
  }
  // every read of the output blocks (MODES 2 and 5 read the blocks they overwrite) is done
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      double *dst = m + (size_t) ((bj + pc) * GJ_NB + cw + 16 * bq + lr) * n + (ai + pr) * GJ_NB + rw + 16 * bi + lk;
#pragma unroll
      for (int e = 0; e < 4; e++)
        dst[4 * e] = acc[bi][bq][e];
    }
}

__global__ void k_gj_reset (int *fb)
{
  if (threadIdx.x == 0)
    fb[0] = 0;
}

// HIPLA_INVERT_SWEEP=0 in the environment sends every cell to the GSL-order pivoting kernel
bool gj_eligible (int n)
{
  static const bool on = [] {
    const char *e = getenv ("HIPLA_INVERT_SWEEP");
    return !(e && e[0] == '0');
  }();
  return on && n >= 2 * GJ_NB && n % GJ_NB == 0;
}

hipError_t gj_reset (int *fb, hipStream_t st)
{
  hipLaunchKernelGGL (k_gj_reset, dim3 (1), dim3 (64), 0, st, fb);
  return hipGetLastError ();
}

// The products of one sweep step over the blocks [r0, r0 + nr), pivoting on blocks sb .. sb+w-1
// whose diagonal block is already inverted in place.
template <int W>
static void gj_products_w (GjArgs g, int sb, int r0, int nr, hipStream_t st)
{
  const int no = nr - W;
  if (no <= 0)
    return;
  g.sb = sb;
  g.sw = W;
  g.r0 = r0;
  g.nr = nr;
  hipLaunchKernelGGL ((k_gj_tile<2, W>), dim3 (g.ncells * no), dim3 (256 * W), 0, st, g);
  hipLaunchKernelGGL ((k_gj_tile<4, W>), dim3 (g.ncells * no * no), dim3 (256), 0, st, g);
  hipLaunchKernelGGL ((k_gj_tile<5, W>), dim3 (g.ncells * no), dim3 (256 * W), 0, st, g);
}

static void gj_products (const GjArgs &g, int sb, int w, int r0, int nr, hipStream_t st)
{
  if (w == 1)
    gj_products_w<1> (g, sb, r0, nr, st);
  else if (w == 2)
    gj_products_w<2> (g, sb, r0, nr, st);
  else
    gj_products_w<4> (g, sb, r0, nr, st);
}

// M[S,S] <- M[S,S]^-1 for the w blocks S = s .. s+w-1 (w = 1, 2 or 4): one block in LDS, else
// the sweep of S's own blocks in half-width steps
static void gj_invert_diag (GjArgs g, int s, int w, hipStream_t st)
{
  if (w == 1)
  {
    g.sb = s;
    hipLaunchKernelGGL (k_gj_diag, dim3 (g.ncells), dim3 (256), 0, st, g);
    return;
  }
  for (int q = s; q < s + w; q += w / 2)
  {
    gj_invert_diag (g, q, w / 2, st);
    gj_products (g, q, w / 2, s, w, st);
  }
}

// Steps of up to 256 columns (four blocks) cut the trailing update's HBM traffic per flop: each
// step reads and writes every other block once.  HIPLA_GJ_STEP=64 or 128 in the environment
// caps the step width.
static int gj_max_width ()
{
  static const int wmax = [] {
    const char *e = getenv ("HIPLA_GJ_STEP");
    const int v = e ? std::atoi (e) : 256;
    return v <= 64 ? 1 : v <= 128 ? 2 : 4;
  }();
  return wmax;
}

// every launch of one size class: prepare, then the sweep.  A step of w > 1 blocks first inverts
// its diagonal block in place by a sweep of that block's own blocks, then forms the products of
// the whole step with K = 64 w.
hipError_t gj_enqueue_class (const GjArgs &g0, hipStream_t st)
{
  if (g0.ncells <= 0)
    return hipSuccess;
  GjArgs g = g0;
  const int nb = g.nsz / GJ_NB;
  g.sb = 0;
  g.sw = 1;
  g.r0 = 0;
  g.nr = nb;
  hipLaunchKernelGGL (k_gj_prepare, dim3 (g.ncells), dim3 (256), 0, st, g);
  hipError_t e = hipGetLastError ();
  for (int s = 0; s < nb && e == hipSuccess;)
  {
    int w = gj_max_width ();
    while (w > nb - s)
      w /= 2;
    gj_invert_diag (g, s, w, st);
    gj_products (g, s, w, 0, nb, st);
    e = hipGetLastError ();
    s += w;
  }
  return e;
}
