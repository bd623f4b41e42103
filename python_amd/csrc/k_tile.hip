// k_tile.hip -- register-tiled FP64 LU factorization in GSL's operation order (gfx950), for
// orders above one wave's reach (64 < n <= 175).
//
// One workgroup of NW waves per matrix, [A | b] held in registers: lane l of wave w holds rows
// l + 64 r (r < R) and columns w*CPW .. w*CPW + CPW - 1 (column n is b).  Rows never move:
// each keeps its GSL position (pos) in a register, so GSL's pivot choice -- first strict maximum
// of |a| in position order (cblas/source_iamax_r.h:30-35) -- and its row swaps are reproduced on
// positions.  Step k:
//   1. after one barrier every wave reads the pivot row and position of step k and the
//      multipliers l_i = a_ik / a_pk (reciprocal multiply when |a_pk| >= DBL_MIN,
//      linalg/lu.c:162-176) that the owner of column k published, and takes the pivot row's
//      values of its own columns from the lane that holds that row (through a wave-private LDS
//      slot);
//   2. the wave owning column k+1 updates that column first, finds the pivot of step k+1 with a
//      DPP reduction over its lanes, forms the multipliers and publishes them (one step ahead,
//      in the LDS slot of the other parity);
//   3. every wave updates the rest of its block, a_ij <- a_ij - l_i u_j with the product rounded
//      first (no FMA contraction).
// The pivot search of step k+1 thus overlaps the other waves' updates of step k.  Every element
// receives its updates in ascending k, the order of GSL's recursive LU (linalg/lu.c:204-288), so
// the factors equal GSL's.  The factors (GSL row order) and the forward-substituted b go to the
// factor store; k_bsub_f64 back-substitutes and checks the residual.
#include <float.h>

#include "hipla_dev.h"
#include "hipla_kernels.h"

// Pivot of step k on register column kc of the calling wave: GSL's choice among the rows at
// positions >= k, multipliers into a[.][kc] (the pivot row keeps its pivot), published in the
// slot `buf` of s_l (by physical row) and s_pv (pivot row, position, zero-pivot flag).
template <int R, int CPW, bool KEEP>
__device__ __forceinline__ void tile_pivot (double (&a)[R][CPW], const int (&pos)[R], int kc, int k, int n, int lane,
                                            double *s_l, int *s_pv, int *s_pq)
{
  long long key = -1;
  int kpos = 0x7fffffff, kslot = 0;
#pragma unroll
  for (int r = 0; r < R; r++)
    if (pos[r] >= k && pos[r] < n)
    {
      const long long q = pivot_key (a[r][kc]);
      if (q > key || (q == key && pos[r] < kpos))
      {
        key = q;
        kpos = pos[r];
        kslot = r;
      }
    }
  const int wl = wave_pivot_lane (key >= 0, key, kpos);
  const int src = wl >= 0 ? wl : 0;
  const int p = __builtin_amdgcn_readlane (lane + 64 * kslot, src);
  const int pq = __builtin_amdgcn_readlane (kpos, src);
  double col[R];
#pragma unroll
  for (int r = 0; r < R; r++)
    col[r] = a[r][kc];
  const double piv = rl_f64 (pick_slot<R> (col, kslot), src);
  const bool recip = fabs (piv) >= DBL_MIN;
  const double sc = 1.0 / piv;
#pragma unroll
  for (int r = 0; r < R; r++)
  {
    const int row = lane + 64 * r;
    if (pos[r] >= k && pos[r] < n && row != p)
    {
      double l;
      if (recip)
        l = a[r][kc] * sc;
      else
        l = a[r][kc] / piv;
      a[r][kc] = l;
      s_l[row] = l;
    }
  }
  if (lane == 0)
  {
    s_pv[0] = p;
    s_pv[1] = pq;
    s_pv[2] = piv == 0.0;
    if (KEEP)
      s_pq[k] = pq;
  }
}

template <int R, int CPW, int NW, bool KEEP>
__global__ __launch_bounds__ (64 * NW, 4) void k_tile_f64 (SolveArgs s)
{
  constexpr int CPWP = (CPW + 1) & ~1;       // even pitch: 16-byte rows
  __shared__ double s_l[2][64 * R];          // multipliers of a step, by physical row
  __shared__ int s_pv[2][4];                 // pivot row, pivot position, zero-pivot flag
  __shared__ __align__ (16) double s_u[NW][CPWP]; // per-wave: the pivot row's values of its columns
  __shared__ int s_pq[KEEP ? 64 * R : 1];    // KEEP: GSL's ipiv, written out after the loop
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
  const int c0 = wv * CPW;                   // first column of this wave's block
  const int neff = list_len (s);

  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];

    double a[R][CPW];
    int pos[R];
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int row = lane + 64 * r;
      pos[r] = row < n ? row : 0x7fffffff;   // rows beyond n never take part
#pragma unroll
      for (int c = 0; c < CPW; c++)
      {
        const int col = c0 + c;
        a[r][c] = (row < n && col < n) ? A[(size_t) row * n + col] : (row < n && col == n) ? b[row] : 0.0;
      }
    }

    // pivot of step 0 (column 0 belongs to wave 0)
    if (wv == 0)
      tile_pivot<R, CPW, KEEP> (a, pos, 0, 0, n, lane, s_l[0], s_pv[0], s_pq);

    int singular = 0;
    for (int kb = 0; kb < NW; kb++)
    {
      // kc must be a compile-time index into the register block: the loop is unrolled fully
#pragma clang loop unroll(full)
      for (int kc = 0; kc < CPW; kc++)
      {
        const int k = kb * CPW + kc;
        if (k < n)
        {
          const int buf = k & 1;
          __syncthreads ();
          // the wave that owns column k+1 carries the step-to-step chain (its column, then the
          // next pivot): its instructions go first on its SIMD
          const int wn = (kc + 1 < CPW) ? kb : kb + 1;  // owner of column k+1
          if (wv == wn)
            __builtin_amdgcn_s_setprio (2);
          const int p = __builtin_amdgcn_readfirstlane (s_pv[buf][0]);
          const int pq = __builtin_amdgcn_readfirstlane (s_pv[buf][1]);
          singular |= s_pv[buf][2];

          // ---- positions: GSL swaps the rows at positions k and pq ----
#pragma unroll
          for (int r = 0; r < R; r++)
          {
            const int row = lane + 64 * r;
            if (row == p)
              pos[r] = k;
            else if (pos[r] == k)
              pos[r] = pq;
          }
          if (wv >= kb)                      // blocks left of column k have nothing to update
          {
            // ---- the pivot row's values of this wave's columns, from the lane holding it ----
            const int ps = p >> 6;           // uniform: register slot of the pivot row
            if (lane == (p & 63))
            {
#pragma unroll
              for (int r = 0; r < R; r++)
                if (r == ps)
                {
#pragma unroll
                  for (int c = 0; c + 1 < CPW; c += 2)
                    *reinterpret_cast<double2 *> (&s_u[wv][c]) = make_double2 (a[r][c], a[r][c + 1]);
                  if (CPW & 1)
                    s_u[wv][CPW - 1] = a[r][CPW - 1];
                }
            }
            // The whole row and this lane's multipliers at once.  The pivot lane's writes precede
            // these reads in LDS order, but to the compiler they are another thread's stores: the
            // memory clobber keeps it from forwarding this lane's own (stale) view of s_u.  (s_l of
            // a row that is not updated is not used.)
            __asm__ volatile ("" ::: "memory");
            double u[CPW], lv[R];
#pragma unroll
            for (int c = 0; c + 1 < CPW; c += 2)
            {
              const double2 t = *reinterpret_cast<const double2 *> (&s_u[wv][c]);
              u[c] = t.x;
              u[c + 1] = t.y;
            }
            if (CPW & 1)
              u[CPW - 1] = s_u[wv][CPW - 1];
#pragma unroll
            for (int r = 0; r < R; r++)
              lv[r] = s_l[buf][lane + 64 * r];
            // in the owner's block the columns up to kc hold multipliers: u = +0 leaves them
            // unchanged (up to the sign of a zero, and inf*0 for a multiplier formed from a zero
            // pivot, whose cell is singular).  Columns past n hold zeros in every row and are
            // never read back.
            if (wv == kb)
            {
#pragma unroll
              for (int c = 0; c <= kc; c++)  // kc is a constant once the kc loop is unrolled
                u[c] = 0.0;
            }

            // ---- step k+1 looks ahead: its column first, then its pivot ----
            const int kn = (kc + 1 < CPW) ? kc + 1 : 0;   // column k+1 in its owner's block
            if (wv == wn && k + 1 < n)
            {
#pragma unroll
              for (int r = 0; r < R; r++)
                if (pos[r] > k && pos[r] < n)
                  a[r][kn] = a[r][kn] - lv[r] * u[kn];
              tile_pivot<R, CPW, KEEP> (a, pos, kn, k + 1, n, lane, s_l[buf ^ 1], s_pv[buf ^ 1], s_pq);
              u[kn] = 0.0;                   // column k+1 holds its multipliers now
            }
            if (wv == wn)
              __builtin_amdgcn_s_setprio (0);

            // ---- update: rows below the pivot (exec-masked per register slot, so finished rows
            //      are not touched) ----
#pragma unroll
            for (int r = 0; r < R; r++)
              if (pos[r] > k && pos[r] < n)
              {
#pragma unroll
                for (int c = 0; c < CPW; c++)
                  a[r][c] = a[r][c] - lv[r] * u[c];
              }
            __asm__ volatile ("" ::: "memory");   // this step's reads before the next step's writes
          }
        }
      }
    }

    // ---- factors (GSL row order: row at position q) and forward-substituted b out ----
    {
      double *F = s.F + s.offA[cell];
      double *Y = s.Y + s.offV[cell];
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        const int row = lane + 64 * r;
        if (row < n)
        {
          double *Fq = F + (size_t) pos[r] * n;
#pragma unroll
          for (int c = 0; c < CPW; c++)
          {
            const int col = c0 + c;
            if (col < n)
              Fq[col] = a[r][c];
            else if (col == n)
              Y[pos[r]] = a[r][c];
          }
        }
      }
    }
    if (KEEP && s.piv)
      for (int t = threadIdx.x; t < n; t += 64 * NW)
        s.piv[s.offV[cell] + t] = s_pq[t];
    if (threadIdx.x == 0)
      s.codes[cell] = singular ? 4 : 0;
    __syncthreads ();
  }
}

// Variants: n <= 103 (8 waves x 13 columns), n <= 128 (8 x 17), n <= 175 (16 waves x 11, three
// rows per lane).
int tile_f64_limit ()
{
  return 175;
}

hipError_t launch_tile_f64 (const SolveArgs &a, int nmax, bool keep, int grid, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  if (nmax > tile_f64_limit () || !a.F || !a.Y)
    return hipErrorInvalidValue;
  grid = (grid > 0 && grid < a.nlist) ? grid : a.nlist;
#define TILE_LAUNCH(R, CPW, NW)                                                                    \
  do                                                                                               \
  {                                                                                                \
    if (keep)                                                                                      \
      hipLaunchKernelGGL ((k_tile_f64<R, CPW, NW, true>), dim3 (grid), dim3 (64 * NW), 0, st, a);  \
    else                                                                                           \
      hipLaunchKernelGGL ((k_tile_f64<R, CPW, NW, false>), dim3 (grid), dim3 (64 * NW), 0, st, a); \
  } while (0)
  if (nmax <= 103)
    TILE_LAUNCH (2, 13, 8);
  else if (nmax <= 128)
    TILE_LAUNCH (2, 17, 8);
  else
    TILE_LAUNCH (3, 11, 16);
#undef TILE_LAUNCH
  return hipGetLastError ();
}

void tile_kernel_symbol (int nmax, bool keep, char *buf, size_t len)
{
  const char *k = keep ? "true" : "false";
  if (nmax <= 103)
    snprintf (buf, len, "k_tile_f64<2, 13, 8, %s>", k);
  else if (nmax <= 128)
    snprintf (buf, len, "k_tile_f64<2, 17, 8, %s>", k);
  else
    snprintf (buf, len, "k_tile_f64<3, 11, 16, %s>", k);
}
