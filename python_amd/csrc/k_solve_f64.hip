// k_solve_f64.hip -- FP64 partial-pivot LU + solve in GSL's operation order (gfx950).
//
// The reference solves A x = b with gsl_linalg_LU_decomp + gsl_linalg_LU_solve and checks the
// residual (source/matrix_cpu.c:96-204).  GSL's recursive LU (GSL 2.6 linalg/lu.c:204-288)
// gives every element a(pos, j) its updates a - l*u for k = 0, 1, ... in ascending order with
// the product rounded first, so a right-looking elimination with the same pivot choice (first
// strict max of |a| by current row position, cblas/source_iamax_r.h:30-35), the same multiplier
// rule (reciprocal multiply when |pivot| >= DBL_MIN, linalg/lu.c:162-176) and no FMA contraction
// reproduces its factors.  The forward solve is folded into the elimination as an extra column;
// the back substitution keeps dtrsv's ascending-j row dot and final divide
// (cblas/source_trsv_r.h:35-52); the residual check keeps dgemv's ascending-j dot
// (cblas/source_gemv_r.h:52-66).  Compiled with -ffp-contract=off.
//
// GSL's dgemm skips exactly-zero multipliers while its dger / dtrsm do not; that only decides
// signs of zero results and whether inf*0 makes a NaN, so these kernels always apply the update
// (values equal GSL's; -0.0 vs +0.0 and inf*0 in overflowing inputs are the documented
// differences).
//
// Shapes, chosen per cell by the host plan:
//   k_wave_f64<NC>     n <= NC <= 64: one wave per matrix, lane = row, row in registers; a pivot
//                      lane keeps its U row in registers (no row swaps: positions tracked per
//                      lane, so GSL's tie rule uses positions, not lanes); factor, solve and
//                      residual check in one kernel.
//   k_tile_f64         64 < n <= 175 (k_tile.hip): one workgroup per matrix, [A|b] in registers.
//   k_blocked_f64<NB>  larger n: right-looking blocked LU, NB-column panel factored in LDS, the
//                      panel's row permutation, U12 and the A22 update applied in place in the
//                      factor store, persistent over the cell list.
// The last two write the factors and the forward-substituted b; k_bsub_f64 (one wave per cell,
// many cells per CU, so the serial dtrsv chains of different cells overlap) back-substitutes
// and checks the residual.
#include <stdio.h>

#include <algorithm>

#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

static size_t hipla_device_max_smem ()
{
  return hipla_device_info ().max_smem;
}

// A value every lane of the wave holds (an LDS broadcast read), moved to a scalar register.
__device__ __forceinline__ int uni (int v)
{
  return __builtin_amdgcn_readfirstlane (v);
}

// Row `row` of an n x n row-major matrix (row wave-uniform: a scalar pointer).
__device__ __forceinline__ double *wrow (double *W, int n, int row)
{
  return W + (size_t) (unsigned) row * (unsigned) n;
}
__device__ __forceinline__ const double *wrow (const double *W, int n, int row)
{
  return W + (size_t) (unsigned) row * (unsigned) n;
}

// Eight consecutive products p[g .. g+7] (g even: 16-byte aligned reads).
struct Oct
{
  double2 a, b, c, d;
};

__device__ __forceinline__ Oct load_oct (const double *p)
{
  Oct o;
  o.a = *reinterpret_cast<const double2 *> (p);
  o.b = *reinterpret_cast<const double2 *> (p + 2);
  o.c = *reinterpret_cast<const double2 *> (p + 4);
  o.d = *reinterpret_cast<const double2 *> (p + 6);
  return o;
}

__device__ __forceinline__ double sub_oct (double t, const Oct &o)
{
  t = t - o.a.x;
  t = t - o.a.y;
  t = t - o.b.x;
  t = t - o.b.y;
  t = t - o.c.x;
  t = t - o.c.y;
  t = t - o.d.x;
  t = t - o.d.y;
  return t;
}

// Eight subtractions of a group after one wait for all four of its reads (the empty asm takes
// the group as input, so the compiler waits once instead of before each read's first use).
__device__ __forceinline__ double sub_oct_w (double t, const Oct &o)
{
  __asm__ volatile ("" ::"v"(o.a.x), "v"(o.a.y), "v"(o.b.x), "v"(o.b.y), "v"(o.c.x), "v"(o.c.y), "v"(o.d.x),
                    "v"(o.d.y));
  return sub_oct (t, o);
}

// tmp - p[j] - p[j+1] - ... - p[n-1], subtracted in ascending order (a dependent chain).  p is
// in LDS, 16-byte aligned, and holds +0.0 at every index below j (down to j rounded to even) and
// from n up to n rounded up to a multiple of eight: the chain runs over whole groups of eight,
// and subtracting +0.0 leaves every value unchanged (IEEE: t - (+0) == t, -0 and NaN included),
// so the padding changes nothing.  Three groups are in flight in three register sets that are
// each reloaded right after their subtractions while later groups remain (no register copies, so
// no wait on a just-issued read); the last one to five groups issue at most one read past the
// row.  chain_load issues the first three reads, so that a caller can start them early (p must
// be readable up to j + 24); chain_run subtracts.
struct Chain
{
  Oct c, d, e;
  int g;
};

__device__ __forceinline__ Chain chain_load (const double *p, int j)
{
  Chain h;
  h.g = j & ~1;
  h.c = load_oct (p + h.g);
  h.d = load_oct (p + h.g + 8);
  h.e = load_oct (p + h.g + 16);
  return h;
}

__device__ __forceinline__ double chain_run (double tmp, const double *p, Chain h, int n)
{
  int g = h.g;
  if (g >= n)
    return tmp;
  const int ng = (n - g + 7) >> 3;   // groups of eight to subtract
  int t = 0;
  for (; t + 6 <= ng; t += 3, g += 24)
  {
    // the scheduling barriers keep each reload right after the subtractions that free its
    // registers (left alone, the scheduler sinks all twelve reads to the end of the iteration,
    // and the next iteration starts by waiting a full LDS round trip)
    tmp = sub_oct_w (tmp, h.c);
    __builtin_amdgcn_sched_barrier (0);
    h.c = load_oct (p + g + 24);
    __builtin_amdgcn_sched_barrier (0);
    tmp = sub_oct_w (tmp, h.d);
    __builtin_amdgcn_sched_barrier (0);
    h.d = load_oct (p + g + 32);
    __builtin_amdgcn_sched_barrier (0);
    tmp = sub_oct_w (tmp, h.e);
    __builtin_amdgcn_sched_barrier (0);
    h.e = load_oct (p + g + 40);
    __builtin_amdgcn_sched_barrier (0);
  }
  // one to five groups left: groups t .. t+2 are in the registers
  const int r = ng - t;
  if (r > 3)
  {
    tmp = sub_oct_w (tmp, h.c);
    __builtin_amdgcn_sched_barrier (0);
    h.c = load_oct (p + g + 24);
    __builtin_amdgcn_sched_barrier (0);
    tmp = sub_oct_w (tmp, h.d);
    __builtin_amdgcn_sched_barrier (0);
    h.d = load_oct (p + g + 32);     // past the row when r == 4 (read, never subtracted)
    __builtin_amdgcn_sched_barrier (0);
    tmp = sub_oct_w (tmp, h.e);
    tmp = sub_oct_w (tmp, h.c);
    if (r > 4)
      tmp = sub_oct_w (tmp, h.d);
  }
  else
  {
    tmp = sub_oct_w (tmp, h.c);
    if (r > 1)
      tmp = sub_oct_w (tmp, h.d);
    if (r > 2)
      tmp = sub_oct_w (tmp, h.e);
  }
  return tmp;
}

// Back substitution U x = y in dtrsv order, run by one wave.  Row i of U is at U + i*ldu
// (columns i..n-1 at their own column index) and y_i at y[i*ys].  Row i subtracts U_ij x_j in
// ascending j (the chain of dependent FP64 subtractions of GSL's dtrsv,
// cblas/source_trsv_r.h:35-52) and divides by U_ii.  Its first term U_i,i+1 x_{i+1} needs the x
// just computed, so it is formed in registers; the products of its later terms only need older
// x, so the lanes form them in parallel one row ahead (during row i+1) and stage them in one of
// two LDS buffers, where the chain reads them 8 at a time; the head of row i's chain is read
// before row i+1's division, which covers that LDS round trip.  A wave issues about one
// instruction per four cycles, and an instruction between two subtractions of the chain delays
// it, so a row keeps its own work small: x lives in LDS (one store per row, two reads per lane),
// the products are written unmasked and the two places where that leaves a wrong value are
// overwritten with +0 (column i of row i-1's buffer, whose x is not known yet, and columns n ..
// n+63, past the matrix), and U comes through D slots (the row's columns, y, U_ii and U_i,i+1),
// each refilled as soon as its row is done.  xs holds the x buffer and the two product buffers,
// XB >= max(64 MAXM, n + 64) doubles each; on return the x buffer holds x.
template <int MAXM>
__device__ __forceinline__ void wave_backsub (const double *U, size_t ldu, const double *y, size_t ys, int n, int lane,
                                              double *xs, int XB)
{
  constexpr int D = 6;
  double *xb = xs, *bufs0 = xs + XB, *bufs1 = xs + 2 * XB;
  // x and both product buffers start at +0 (x_j for j >= n and the padding stay so)
  for (int k = lane; k < 3 * XB; k += 64)
    xs[k] = 0.0;
  // slot sl serves rows n-1-sl, n-1-sl-D, ...; rows n-1 .. n-D+1 are read here, row n-D by the
  // first row's refill.  A lane past column n-1 reads column n-1 instead (its product lands in
  // the padding, which is cleared after the products are written).  Row n-1 has no first term:
  // its U_i,i+1 slot holds +0 and x_n is +0, so y - (+0 * +0) == y.
  int off[MAXM];
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    off[m] = min (lane + 64 * m, n - 1);
  double cu[D][MAXM], cy[D], cd[D], cn[D];
#pragma unroll
  for (int sl = 0; sl < D - 1; sl++)
  {
    const int r = uni (max (n - 1 - sl, 0));
    const double *Ur = U + (size_t) r * ldu;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      cu[sl][m] = Ur[off[m]];
    cy[sl] = y[(size_t) r * ys];
    cd[sl] = Ur[r];
    cn[sl] = sl == 0 ? 0.0 : Ur[min (r + 1, n - 1)];
    __builtin_amdgcn_sched_barrier (0);
  }
  double xnew = 0.0;             // x_{i+1}, the newest solution value (wave-uniform)
  __asm__ volatile ("" ::: "memory");
  Chain h = chain_load (((n - 1) & 1) ? bufs1 : bufs0, n + 1);   // padding only
  for (int i0 = n - 1; i0 >= 0; i0 -= D)
  {
#pragma unroll
    for (int sl = 0; sl < D; sl++)
    {
      const int i = i0 - sl;
      if (i >= 0)
      {
        const double *pb = (i & 1) ? bufs1 : bufs0;
        double *bw = ((i - 1) & 1) ? bufs1 : bufs0;
        double tmp = cy[sl] - cn[sl] * xnew;
        // x_{i+1} into the x buffer (x_n: +0 again) and this lane's x back for row i-1's products
        xb[i + 1] = xnew;
        __asm__ volatile ("" ::: "memory");
        double xv[MAXM];
#pragma unroll
        for (int m = 0; m < MAXM; m++)
          xv[m] = xb[lane + 64 * m];
        // refill row i+1's slot with row i+1-D (row 0 again past the top)
        {
          const int sr = (sl + D - 1) % D;
          const int rr = uni (max (i + 1 - D, 0));
          const double *Ur = U + (size_t) rr * ldu;
#pragma unroll
          for (int m = 0; m < MAXM; m++)
            cu[sr][m] = Ur[off[m]];
          cy[sr] = y[(size_t) rr * ys];
          cd[sr] = Ur[rr];
          cn[sr] = Ur[min (rr + 1, n - 1)];
        }
        __builtin_amdgcn_sched_barrier (0);
        tmp = chain_run (tmp, pb, h, n);
        __builtin_amdgcn_sched_barrier (0);
        // row i-1's products of columns i+1 .. n-1; column i (x_i not known yet) and columns n
        // and up are +0.  Row 0 writes a buffer no row reads.
#pragma unroll
        for (int m = 0; m < MAXM; m++)
          bw[lane + 64 * m] = cu[(sl + 1) % D][m] * xv[m];
        bw[n + lane] = 0.0;
        bw[i] = 0.0;
        // row i-1's head: the chain reads other lanes' products, which to the compiler are
        // other threads' stores
        __asm__ volatile ("" ::: "memory");
        h = chain_load (bw, i + 1);
        __builtin_amdgcn_sched_barrier (0);
        xnew = tmp / cd[sl];
      }
    }
  }
  xb[0] = xnew;
  __asm__ volatile ("" ::: "memory");
}

// ============================================================================================
// k_wave_f64: one wave per matrix
// ============================================================================================
// Lane r holds row r of [A | b] in registers; columns are shifted left as they are eliminated, so
// a[0] is always the current column and a[t] the column t to its right.  Registers past a row's
// live columns hold +0 (loaded so and kept so by the shifted update), and so does the broadcast
// row past its live entries: the update and the back substitution then run over whole groups of
// eight columns under wave-uniform guards, the padded terms being exact no-ops (x - (+0) == x,
// +0 - l * (+0) == +0).  The lane whose row is chosen as pivot at step k keeps U(k, k..n-1) and
// y_k in its registers from then on; the other lanes read the row from a small LDS broadcast
// buffer.  A is loaded (and re-read for the residual check) with coalesced flat loads through a
// staging tile of SR rows.  LDS per wave: the tile, the broadcast row, x with its padding (plus
// the multipliers and pivots when the factors are kept).
template <int NC>
struct WaveF64Layout
{
  static constexpr int SR = NC < 16 ? NC : 16;   // rows per staging pass of A
  static constexpr int SP = NC | 1;              // staging pitch (odd: no bank conflicts)
  static constexpr int BR = NC + 2;              // broadcast row: [t] = U(k, k+t), [NC] = y_k
  static constexpr int XS = NC + 8;              // x, then +0 up to the next group of eight
  static constexpr int base (bool keep)
  {
    return ((SR * SP + 1) & ~1) + BR + XS + (keep ? NC / 2 : 0);
  }
};

// Rows r0 .. r0+nr-1 of the row-major n x n matrix A into the staging tile (pitch SP): the rows
// are contiguous in A, so lane l loads elements l, l+64, ... of that run (coalesced, all loads in
// flight together) and scatters them to (row, column) in LDS.  Columns n .. (n+7)&~7 of each
// staged row are set to +0.
template <int SR, int SP, int NC>
__device__ __forceinline__ void wave_stage_rows (const double *A, int n, int r0, int nr, int lane, double *stg)
{
  constexpr int TMAX = (SR * NC + 63) / 64;
  const int cnt = nr * n;
  const double *src = A + (size_t) r0 * n;
  const float rn = 1.0f / (float) n;
  double v[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; t++)
  {
    const int f = lane + 64 * t;
    v[t] = f < cnt ? src[f] : 0.0;
  }
#pragma unroll
  for (int t = 0; t < TMAX; t++)
  {
    int f = lane + 64 * t;
    // recomputed in every pass (the same in each): kept live across passes, the scatter
    // addresses would be spilled to scratch memory
    __asm__ volatile ("" : "+v"(f));
    if (f < cnt)
    {
      // f / n: (f + 0.5) / n is at least 0.5 / n from an integer, far above the float error
      const int rr = (int) (((float) f + 0.5f) * rn);
      stg[rr * SP + (f - rr * n)] = v[t];
    }
  }
  const int n8 = min ((n + 7) & ~7, SP);
  if (lane < nr)
    for (int c = n; c < n8; c++)
      stg[lane * SP + c] = 0.0;
}

template <int NC, int WPB, bool KEEP>
__global__ __launch_bounds__ (64 * WPB, NC == 64 ? 3 : 4) void k_wave_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  using Lay = WaveF64Layout<NC>;
  constexpr int SR = Lay::SR, SP = Lay::SP;
  constexpr int PER_WAVE = (Lay::base (KEEP) + 1) & ~1;
  const int wv = uni (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int neff = list_len (s);
  double *stg = smem + wv * PER_WAVE;
  double *brow = stg + ((SR * SP + 1) & ~1);
  double *xs = brow + Lay::BR;
  int *pvs = reinterpret_cast<int *> (xs + Lay::XS);  // KEEP: GSL's ipiv
  for (int item = blockIdx.x * WPB + wv; item < neff; item += gridDim.x * WPB)
  {
    wave_lds_fence ();
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    if (n <= 0 || n > NC)
    {
      if (lane == 0)
        s.codes[cell] = 5;
      continue;
    }

    // ---- row `lane` of A into registers (+0 past column n) ----
    const bool in = lane < n;
    double a[NC];
#pragma unroll
    for (int j = 0; j < NC; j++)
      a[j] = 0.0;
    for (int r0 = 0; r0 < n; r0 += SR)
    {
      const int nr = min (SR, n - r0);
      wave_stage_rows<SR, SP, NC> (A, n, r0, nr, lane, stg);
      wave_lds_fence ();
      if (lane >= r0 && lane < r0 + nr)
      {
#pragma unroll
        for (int j = 0; j < NC; j++)
          a[j] = j < n ? stg[(lane - r0) * SP + j] : 0.0;
      }
      wave_lds_fence ();
    }
    const double b0 = in ? b[lane] : 0.0;
    double bb = b0;

    // ---- elimination, forward solve folded in ----
    bool alive = in;
    int pos = lane;
    int singular = 0;
    for (int k = 0; k < n; k++)
    {
      const int q = max (wave_pivot_lane (alive, pivot_key (a[0]), pos), 0);
      const int ppos = __builtin_amdgcn_readlane (pos, q);
      const double piv = rl_f64 (a[0], q);
      const int live = n - k;                          // live columns a[0 .. live-1]
      if (lane == q)
      {
#pragma unroll
        for (int g = 0; g < NC; g += 8)
          if (g < live)
          {
#pragma unroll
            for (int e = 0; e < 8; e += 2)
              *reinterpret_cast<double2 *> (brow + g + e) = make_double2 (a[g + e], a[g + e + 1]);
          }
        brow[NC] = bb;
      }
      if (KEEP && lane == 0)
        pvs[k] = ppos;
      if (alive && pos == k)
        pos = ppos;
      if (lane == q)
      {
        alive = false;        // from here on this lane holds U(k, k..) and y_k
        pos = k;
      }
      singular |= (piv == 0.0);
      __asm__ volatile ("" ::: "memory");    // broadcast row written before it is read (LDS order)
      if (alive)
      {
        const double l = gsl_multiplier (a[0], piv);
        // KEEP: multipliers go to row `lane` (the physical row) of the factor output; the rows
        // move to their GSL positions after the last step
        if (KEEP && s.LU)
          s.LU[s.offA[cell] + (size_t) lane * n + k] = l;
        bb = bb - l * brow[NC];
#pragma unroll
        for (int g = 0; g < NC; g += 8)
          if (g < live)
          {
            double u8[8];
#pragma unroll
            for (int e = 0; e < 8; e += 2)
            {
              const double2 t = *reinterpret_cast<const double2 *> (brow + g + e);
              u8[e] = t.x;
              u8[e + 1] = t.y;
            }
#pragma unroll
            for (int e = 0; e < 8; e++)
              if (g + e >= 1)
                a[g + e - 1] = a[g + e] - l * u8[e];
          }
        a[NC - 1] = 0.0;
      }
      __asm__ volatile ("" ::: "memory");    // reads of this step before the next step's writes
    }
    wave_lds_fence ();

    if (KEEP && s.LU)
    {
      // row `lane` became pivot at step pos: LU row pos = [L multipliers | U row].  Its
      // multipliers sit in row `lane`; they move to row pos eight columns at a time, each
      // chunk read by every lane before any lane writes it (the stores wait on the loads they
      // store), then the lane writes its U row behind them.
      double *lu = s.LU + s.offA[cell];
      __builtin_amdgcn_fence (__ATOMIC_ACQ_REL, "workgroup");   // this wave's multiplier stores
      for (int c0 = 0; c0 < n; c0 += 8)
      {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; e++)
          v[e] = (in && c0 + e < pos) ? lu[(size_t) lane * n + c0 + e] : 0.0;
        __builtin_amdgcn_fence (__ATOMIC_ACQ_REL, "workgroup");
#pragma unroll
        for (int e = 0; e < 8; e++)
          if (in && c0 + e < pos)
            lu[(size_t) pos * n + c0 + e] = v[e];
      }
      if (in)
      {
        double *out = lu + (size_t) pos * n;
#pragma unroll
        for (int t = 0; t < NC; t++)
          if (t < n - pos)
            out[pos + t] = a[t];
      }
    }
    if (KEEP && s.piv)
      for (int t = lane; t < n; t += 64)
        s.piv[s.offV[cell] + t] = pvs[t];
    if (singular)
    {
      if (lane == 0)
        s.codes[cell] = 4;
      continue;
    }

    // ---- back substitution in dtrsv order: the lane holding U row i subtracts its products in
    //      ascending column order from x in LDS (+0 from column n on), then divides ----
    if (lane < 8)
      xs[n + lane] = 0.0;
    for (int i = n - 1; i >= 0; i--)
    {
      __asm__ volatile ("" ::: "memory");    // x_{i+1} (another lane) written before it is read
      if (in && pos == i)
      {
        double tmp = bb;
#pragma unroll
        for (int g = 0; g < NC; g += 8)
          if (g < n - i)
          {
            double x8[8];
#pragma unroll
            for (int e = 0; e < 8; e++)
              x8[e] = xs[i + g + e];
#pragma unroll
            for (int e = 0; e < 8; e++)
              if (g + e >= 1)
              {
                const double pr = a[g + e] * x8[e];
                tmp = tmp - pr;
              }
          }
        xs[i] = tmp / a[0];
      }
    }
    wave_lds_fence ();

    // ---- residual check t = A x (gsl_blas_dgemv order: 0 + sum in ascending j) and codes 2 / 3
    //      of the first failing row; A is re-read through the staging tile ----
    int fail = 0;
    for (int r0 = 0; r0 < n; r0 += SR)
    {
      const int nr = min (SR, n - r0);
      wave_stage_rows<SR, SP, NC> (A, n, r0, nr, lane, stg);
      wave_lds_fence ();
      if (lane >= r0 && lane < r0 + nr)
      {
        const double *Ai = stg + (lane - r0) * SP;
        double t = 0.0;
#pragma unroll
        for (int g = 0; g < NC; g += 8)
          if (g < n)
          {
#pragma unroll
            for (int e = 0; e < 8; e++)
            {
              const double pr = Ai[g + e] * xs[g + e];   // +0 * +0 past column n
              t = t + pr;
            }
          }
        fail = residual_row_code (0.0 + t, b0);
      }
      wave_lds_fence ();
    }
    const unsigned long long fb = __ballot (fail != 0);
    int code = 0;
    if (fb)
      code = __builtin_amdgcn_readlane (fail, __ffsll ((long long) fb) - 1);
    if (in)
      s.X[s.offV[cell] + lane] = xs[lane];
    if (lane == 0)
      s.codes[cell] = code;
  }
}

// ============================================================================================
// k_blocked_f64: right-looking blocked LU, NB-column panel in LDS, trailing matrix in a
// per-workgroup HBM workspace (MALL / L2 resident between panels)
// ============================================================================================
//
// Per panel k0 .. k1 = k0 + NB:
//   1. the panel (all rows from k0 down, NB columns and b) is factored in LDS.  Rows are not
//      moved: pos[r] / phys[q] track which panel row sits at which position, so one column costs
//      two barriers (pivot candidates, then elimination against the pivot row, which nobody
//      writes in that step).  GSL's pivot rule and tie break use the positions.
//   2. the panel is written back in position order; the net row permutation of the panel (at
//      most 2 NB moved rows) is applied to every other column as one gather / scatter per column,
//      and U12 = L11^-1 A12 is formed per column (ascending kk, dtrsm's order) into W and LDS.
//   3. A22 -= L21 U12 in 8 TR x 32 tiles per wave, TR x 4 elements per lane; each element
//      takes its NB updates in ascending kk with the product rounded first, the order of GSL's
//      recursion.
// Then the back substitution runs from LDS-staged chunks of U and the workgroup forms the
// residual check.
// A22 values of a blocked trailing-update tile: rows i0 + 8 r and columns c0 + 8 c (clamped to
// n-1; r < TR, c < TC).  In the first panel a row the panel's permutation left in place is still only in A;
// every other row is in the factor store W.
template <int TR, int TC>
__device__ __forceinline__ void blk_load_a22 (double (&d)[TR][TC], const double *A, const double *W, const int *phys,
                                              int k0, int n, int i0, int c0)
{
#pragma unroll
  for (int r = 0; r < TR; r++)
  {
    const int q = min (i0 + 8 * r, n - 1);
    const bool fromA = k0 == 0 && phys[q - k0] == q - k0;
    const double *row = (fromA ? A : W) + (size_t) q * n;
#pragma unroll
    for (int c = 0; c < TC; c++)
      d[r][c] = row[(unsigned) min (c0 + 8 * c, n - 1)];
  }
}

template <int BS, int NB, bool US, bool KEEP>
__global__ __launch_bounds__ (BS, BS == 256 ? 2 : 4) void k_blocked_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  constexpr int PP = NB + 2;           // panel row pitch (even: 16-byte aligned rows); column NB = b
  constexpr int NW = BS / 64;
  constexpr int PR = US ? 1 : 2;       // panel rows per thread (US: n <= BS, one; else two)
  __shared__ long long c_key[2][NW];   // per-wave pivot candidates, by step parity
  __shared__ int c_pos[2][NW], c_row[2][NW];
  __shared__ double c_val[2][NW][NB + 2];
  __shared__ int s_ipiv[NB];
  __shared__ int mv_dst[NB], mv_src[NB];
  __shared__ int s_nmv, s_fail;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane (tid >> 6);   // wave-uniform

  const int neff = list_len (s);
  if (!list_in_window (s, neff))
    return;
  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    double *W = s.F + s.offA[cell];     // factored in place in the factor store
    double *wb = s.Y + s.offV[cell];
    double *P = smem;
    double *Us = P + (size_t) n * PP;
    int *pos = (int *) (Us + (US ? (size_t) NB * n : 0));   // position of panel row r (relative to k0)
    int *phys = pos + n;               // panel row at position q

    if (tid == 0)
      s_fail = 0x7fffffff;
    __syncthreads ();

    int singular = 0;
    for (int k0 = 0; k0 < n; k0 += NB)
    {
      const int nbk = min (NB, n - k0);
      const int rows = n - k0;
      const int k1 = k0 + nbk;
      // the first panel reads A and b themselves (no copy into the factor store): everything it
      // does not write lands in the factor store with the first trailing update
      const double *S = k0 == 0 ? A : W;
      const double *Sb = k0 == 0 ? b : wb;
      // The panel and the U12 phase are chains of short dependent steps, one barrier apart;
      // the trailing update is throughput work.  Two workgroups share a CU, so while one runs
      // its trailing update the other's panel waves take the issue slots first.
      __builtin_amdgcn_s_setprio (2);

      // ---- 1. panel factorization in registers: thread t holds panel rows t + BS pr ----
      // Each step publishes, per wave, the best candidate (GSL key, position) together with
      // that row's values, so after one barrier every thread can eliminate against the winner.
      double v[PR][NB + 1];            // panel columns k0 .. k0+NB-1, then b
      int mp[PR];                      // position of each held row (relative to k0)
#pragma unroll
      for (int pr = 0; pr < PR; pr++)
      {
        const int r = tid + BS * pr;
        mp[pr] = r < rows ? r : 0x7fffffff;
        const double *src = S + (size_t) (k0 + min (r, rows - 1)) * n + k0;
#pragma unroll
        for (int c = 0; c < NB; c++)
          v[pr][c] = (r < rows && c < nbk) ? src[c] : 0.0;
        v[pr][NB] = r < rows ? Sb[k0 + r] : 0.0;
      }
      if (tid == 0)
        s_nmv = 0;
#pragma clang loop unroll(full)
      for (int kk = 0; kk < NB; kk++)
      {
        if (kk < nbk)
        {
          const int buf = kk & 1;
          long long key = -1;
          int bp = 0x7fffffff, bs = 0;
#pragma unroll
          for (int pr = 0; pr < PR; pr++)
            if (mp[pr] >= kk && mp[pr] < rows)
            {
              const long long q = pivot_key (v[pr][kk]);
              if (q > key || (q == key && mp[pr] < bp))
              {
                key = q;
                bp = mp[pr];
                bs = pr;
              }
            }
          const int wl = wave_pivot_lane (key >= 0, key, bp);
          if (lane == (wl >= 0 ? wl : 0))
          {
            c_key[buf][wv] = wl >= 0 ? key : -1;
            c_pos[buf][wv] = wl >= 0 ? bp : 0x7fffffff;
            c_row[buf][wv] = tid + BS * bs;
#pragma unroll
            for (int c = kk; c <= NB; c++)
            {
              double col[PR];
#pragma unroll
              for (int pr = 0; pr < PR; pr++)
                col[pr] = v[pr][c];
              c_val[buf][wv][c] = pick_slot<PR> (col, bs);
            }
          }
          __syncthreads ();
          int pq, p;
          combine_pivot_candidates (c_key[buf], c_pos[buf], c_row[buf], NW, lane, pq, p);
          const double *u = c_val[buf][(p & (BS - 1)) >> 6];
          const double piv = u[kk];
          // the pivot row's remaining columns, read together (one LDS wait, not one per column);
          // columns from nbk on are zero padding in every row and are never written back
          double uu[NB + 1];
#pragma unroll
          for (int c = kk + 1; c <= NB; c++)
            uu[c] = u[c];
          singular |= (piv == 0.0);
          const bool recip = fabs (piv) >= DBL_MIN;
#pragma unroll
          for (int pr = 0; pr < PR; pr++)
          {
            const int r = tid + BS * pr;
            // every row still below the pivot, the row at position kk (moving to pq) included
            if (r != p && mp[pr] >= kk && mp[pr] < rows)
            {
              // GSL's multiplier: reciprocal multiply, or a division for a tiny pivot
              const double t = (recip ? 1.0 : v[pr][kk]) / piv;
              const double l = recip ? v[pr][kk] * t : t;
              v[pr][kk] = l;
#pragma unroll
              for (int c = kk + 1; c < NB; c++)
                v[pr][c] = v[pr][c] - l * uu[c];
              v[pr][NB] = v[pr][NB] - l * uu[NB];
            }
            // GSL swaps the rows at positions kk and pq
            if (r == p)
              mp[pr] = kk;
            else if (mp[pr] == kk)
              mp[pr] = pq;
          }
          if (tid == 0)
            s_ipiv[kk] = pq;
        }
      }
      // factored panel rows to LDS and the position maps, for the U12 / A22 phases
#pragma unroll
      for (int pr = 0; pr < PR; pr++)
      {
        const int r = tid + BS * pr;
        if (r < rows)
        {
          double *dst = P + (size_t) r * PP;
#pragma unroll
          for (int c = 0; c <= NB; c++)
            dst[c] = v[pr][c];
          pos[r] = mp[pr];
          phys[mp[pr]] = r;
        }
      }
      __syncthreads ();

      // ---- 2a. panel back in position order; rows moved below the panel ----
      for (int q = tid; q < rows; q += BS)
      {
        const int r = phys[q];
        const double *src = P + (size_t) r * PP;
        double *dst = W + (size_t) (k0 + q) * n + k0;
        for (int c = 0; c < nbk; c++)
          dst[c] = src[c];
        wb[k0 + q] = src[NB];
        if (q >= nbk && r != q)
        {
          const int m = atomicAdd (&s_nmv, 1);
          mv_dst[m] = q;
          mv_src[m] = r;
        }
      }
      if (KEEP && s.piv && tid < nbk)
        s.piv[s.offV[cell] + k0 + tid] = k0 + s_ipiv[tid];
      __syncthreads ();

      // ---- 2b. other columns: row permutation, then U12 ----
      // Row indices are wave-uniform (LDS broadcasts moved to scalar registers): each access is
      // a scalar row base plus the lane's unsigned column offset.
      const int nmv = uni (s_nmv);
      const int last = nbk - 1;
      for (int c = tid; c < n; c += BS)
      {
        if (c >= k0 && c < k1)
          continue;
        const unsigned cu = (unsigned) c;
        double u[NB];
        // all loads first (indices clamped so every one is in bounds, conditions wave-uniform)
#pragma unroll
        for (int kk = 0; kk < NB; kk++)
          u[kk] = wrow (S, n, k0 + uni (phys[min (kk, last)]))[cu];
        if (US)
        {
          // park the permuted rows in this column's U12 slots (free until written below) so the
          // displaced rows' registers are not live at the same time
#pragma unroll
          for (int kk = 0; kk < NB; kk++)
            Us[kk * n + c] = u[kk];
          __asm__ volatile ("" ::: "memory");
        }
        if (nmv > 0)
        {
          double v[NB];
#pragma unroll
          for (int m = 0; m < NB; m++)
            v[m] = wrow (S, n, k0 + uni (mv_src[min (m, nmv - 1)]))[cu];
#pragma unroll
          for (int m = 0; m < NB; m++)
            if (m < nmv)
              wrow (W, n, k0 + uni (mv_dst[m]))[cu] = v[m];
        }
        if (US)
        {
          __asm__ volatile ("" ::: "memory");
#pragma unroll
          for (int kk = 0; kk < NB; kk++)
            u[kk] = Us[kk * n + c];
        }
        if (c >= k1)
        {
          // a column right of the panel exists only when the panel is full (nbk == NB)
#pragma unroll
          for (int r = 1; r < NB; r++)
          {
            const double *Lr = P + (size_t) uni (phys[r]) * PP;
#pragma unroll
            for (int kk = 0; kk < r; kk++)
              u[r] = u[r] - Lr[kk] * u[kk];
          }
#pragma unroll
          for (int kk = 0; kk < NB; kk++)
          {
            wrow (W, n, k0 + kk)[cu] = u[kk];
            if (US)
              Us[kk * n + c] = u[kk];
          }
        }
        else
        {
#pragma unroll
          for (int kk = 0; kk < NB; kk++)
            if (kk < nbk && uni (phys[min (kk, last)]) != kk)
              wrow (W, n, k0 + kk)[cu] = u[kk];
        }
      }
      __syncthreads ();

      // ---- 3. A22 -= L21 U12 (only when the panel is full) ----
      // Wave tiles of 8 TR x 32 elements: lane (rg, cg) = (lane >> 3, lane & 7) updates the
      // TR x 4 elements at rows i0 + 8 r and columns c0 + 8 c (i0 = tile row + rg, c0 = tile
      // column + cg).  Per update step a lane reads TR multipliers and 4 U12 values for 4 TR
      // updates (TR = 8 with 256 threads, whose waves have the registers; 4 with 512): a
      // quarter to a third of the LDS bytes per update of a tile with one column per lane,
      // whose multipliers are whole-wave broadcasts (the LDS, not the FP64 pipes, bounded that
      // one).  Each element takes its NB updates in ascending kk with the product rounded
      // first.  The next tile's A22 values are read before the current tile is updated; rows
      // and columns past n are clamped to n-1 for the reads and not stored.
      __builtin_amdgcn_s_setprio (0);
      const int nr = n - k1;
      if (nr > 0)
      {
        constexpr int TR = 4, TC = 4;
        const int nrt = (nr + 8 * TR - 1) / (8 * TR);
        const int nct = (nr + 31) >> 5;
        const int nit = nrt * nct;
        const int rg = lane >> 3, cg = lane & 7;
        double an[TR][TC];
        int it = wv;
        if (it < nit)
          blk_load_a22<TR, TC> (an, A, W, phys, k0, n, k1 + (it / nct) * 8 * TR + rg, k1 + (it % nct) * 32 + cg);
        for (; it < nit; it += NW)
        {
          const int i0 = k1 + (it / nct) * 8 * TR + rg;
          const int c0 = k1 + (it % nct) * 32 + cg;
          double a[TR][TC];
#pragma unroll
          for (int r = 0; r < TR; r++)
#pragma unroll
            for (int c = 0; c < TC; c++)
              a[r][c] = an[r][c];
          const int nx = it + NW;
          if (nx < nit)
            blk_load_a22<TR, TC> (an, A, W, phys, k0, n, k1 + (nx / nct) * 8 * TR + rg, k1 + (nx % nct) * 32 + cg);
          int pl[TR];                  // LDS offsets of the tile rows' panel rows
          unsigned cc[TC];             // the lane's columns (clamped)
#pragma unroll
          for (int r = 0; r < TR; r++)
            pl[r] = phys[min (i0 + 8 * r, n - 1) - k0] * PP;
#pragma unroll
          for (int c = 0; c < TC; c++)
            cc[c] = (unsigned) min (c0 + 8 * c, n - 1);
#pragma unroll
          for (int kk = 0; kk < NB; kk++)
          {
            double l[TR], u[TC];
#pragma unroll
            for (int r = 0; r < TR; r++)
              l[r] = P[pl[r] + kk];
#pragma unroll
            for (int c = 0; c < TC; c++)
              u[c] = US ? Us[kk * n + cc[c]] : wrow (W, n, k0 + kk)[cc[c]];
            if constexpr (BS == 256)
            {
              // every product formed before the first subtraction (these waves have the
              // registers): the subtractions then never wait on a multiply just issued
              double pr[TR][TC];
#pragma unroll
              for (int r = 0; r < TR; r++)
#pragma unroll
                for (int c = 0; c < TC; c++)
                {
                  pr[r][c] = l[r] * u[c];
                  __asm__ volatile ("" : "+v"(pr[r][c]));
                }
#pragma unroll
              for (int r = 0; r < TR; r++)
#pragma unroll
                for (int c = 0; c < TC; c++)
                  a[r][c] = a[r][c] - pr[r][c];
            }
            else
            {
#pragma unroll
              for (int r = 0; r < TR; r++)
#pragma unroll
                for (int c = 0; c < TC; c++)
                  a[r][c] = a[r][c] - l[r] * u[c];
            }
            // one step's operands live at a time (hoisting every step's LDS reads to the top
            // of the unrolled loop would need NB times the registers)
            __builtin_amdgcn_sched_barrier (0);
          }
          // the updates are done here, for every element: left alone, the compiler sinks each
          // element's update chain into its conditional store and reads every step's operands
          // first, far more than the registers hold
#pragma unroll
          for (int r = 0; r < TR; r++)
#pragma unroll
            for (int c = 0; c < TC; c++)
              __asm__ volatile ("" : "+v"(a[r][c]));
#pragma unroll
          for (int r = 0; r < TR; r++)
          {
            const int i = i0 + 8 * r;
            if (i < n)
            {
#pragma unroll
              for (int c = 0; c < TC; c++)
                if (c0 + 8 * c < n)
                  W[(size_t) i * n + (unsigned) (c0 + 8 * c)] = a[r][c];
            }
          }
        }
      }
      __syncthreads ();
    }

    if (tid == 0)
      s.codes[cell] = singular ? 4 : 0;
    __syncthreads ();
  }
}

// ============================================================================================
// k_bsub_f64: back substitution and residual check, one wave per cell
// ============================================================================================
// Reads the factor store and y the workgroup factor kernels wrote (GSL row order), runs the
// dtrsv chain (wave_backsub) and the dgemv residual check, and writes x and the cell's code.
// Cells the factor kernel found singular (code 4) are skipped: x is not written.  Many cells'
// chains share each SIMD, so the serial chains of large matrices overlap instead of idling a
// whole workgroup.
template <int MAXM>
__global__ __launch_bounds__ (64, MAXM <= 4 ? 2 : 1) void k_bsub_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  constexpr int WPB = 1;                 // one wave per workgroup: the waves spread over every CU
  constexpr int XB = 64 * MAXM + 64;     // x, and each product buffer (n + 63 readable)
  constexpr int XP = 3 * XB;             // per wave: the x buffer and two product buffers
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
  double *xs = smem + wv * XP;
  const int neff = list_len (s);
  for (int item = blockIdx.x * WPB + wv; item < neff; item += gridDim.x * WPB)
  {
    const int cell = s.list[item];
    if (s.codes[cell] == 4)
      continue;
    const int n = s.n[cell];
    wave_backsub<MAXM> (s.F + s.offA[cell], (size_t) n, s.Y + s.offV[cell], 1, n, lane, xs, XB);

#pragma unroll
    for (int m = 0; m < MAXM; m++)
      if (lane + 64 * m < n)
        s.X[s.offV[cell] + lane + 64 * m] = xs[lane + 64 * m];

    // residual check t = A x (gsl_blas_dgemv order: t_r = sum over ascending j of A_rj x_j) and
    // codes 2 / 3 of the first failing row.  Lane l runs the chains of rows l + 64 m side by
    // side, eight columns at a time, with that block of every row read at once: one wave keeps
    // 16 KB of A in flight (MAXM = 4) instead of 4 KB, and the residual is bound by memory
    // latency.  Lanes past row n-1 read row n-1 and drop the result.
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    const double *Ar[MAXM];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      Ar[m] = A + (size_t) min (lane + 64 * m, n - 1) * n;
    double t[MAXM];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      t[m] = 0.0;
    int j0 = 0;
    for (; j0 + 8 <= n; j0 += 8)
    {
      double a[MAXM][8];
#pragma unroll
      for (int m = 0; m < MAXM; m++)
#pragma unroll
        for (int e = 0; e < 8; e++)
          a[m][e] = Ar[m][j0 + e];
#pragma unroll
      for (int e = 0; e < 8; e++)
      {
        const double xj = xs[j0 + e];
#pragma unroll
        for (int m = 0; m < MAXM; m++)
        {
          const double pr = a[m][e] * xj;
          t[m] = t[m] + pr;
        }
      }
    }
    for (; j0 < n; j0++)
    {
      const double xj = xs[j0];
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const double pr = Ar[m][j0] * xj;
        t[m] = t[m] + pr;
      }
    }
    int first = 0x7fffffff;
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int r = lane + 64 * m;
      if (r < n)
      {
        const int f = residual_row_code (0.0 + t[m], b[r]);
        if (f)
          first = min (first, (r << 2) | f);
      }
    }
    first = wave_min_i32_dpp (first);
    if (lane == 0)
      s.codes[cell] = (first == 0x7fffffff) ? 0 : (first & 3);
  }
}

// ============================================================================================
// launchers
// ============================================================================================
static hipError_t launch_bsub_f64 (const SolveArgs &a, int nmax, hipStream_t st);

size_t wave_f64_smem (int nc, int wpb, bool keep)
{
  int per = 0;
  switch (nc)
  {
  case 16:
    per = WaveF64Layout<16>::base (keep);
    break;
  case 32:
    per = WaveF64Layout<32>::base (keep);
    break;
  default:
    per = WaveF64Layout<64>::base (keep);
    break;
  }
  return (size_t) wpb * ((per + 1) & ~1) * sizeof (double);
}


int wgr_f64_limit ()
{
  return tile_f64_limit ();
}

// Blocked kernel variants: NB = 16 with U12 staged in LDS up to n = 512, NB = 8 reading U12 from
// the factor store up to n = 1024 (LDS: panel n x (NB+2), U12 NB x n when staged, two int maps
// of n).
static void blocked_variant (int nmax, int &nb, bool &us)
{
  nb = nmax <= 512 ? 16 : 8;
  us = nmax <= 512;
}

size_t blocked_f64_smem (int nmax)
{
  int nb;
  bool us;
  blocked_variant (nmax, nb, us);
  return (size_t) nmax * (nb + 2) * sizeof (double) + (us ? (size_t) nb * nmax * sizeof (double) : 0) +
         2 * (size_t) nmax * sizeof (int);
}

// Workgroups of the blocked kernel one CU holds at once (LDS and registers), for the persistent
// grid: two (about 74 KB of LDS and 4 x 128 registers per SIMD each at n = 256).
int blocked_f64_wgs_per_cu (int)
{
  return 2;
}

template <int NC, int WPB>
static hipError_t launch_wave_nc (const SolveArgs &a, bool keep, int grid, hipStream_t st)
{
  const int full = (a.nlist + WPB - 1) / WPB;
  grid = (grid > 0 && grid < full) ? grid : full;
  const size_t sm = wave_f64_smem (NC, WPB, keep);
  if (keep)
    hipLaunchKernelGGL ((k_wave_f64<NC, WPB, true>), dim3 (grid), dim3 (64 * WPB), sm, st, a);
  else
    hipLaunchKernelGGL ((k_wave_f64<NC, WPB, false>), dim3 (grid), dim3 (64 * WPB), sm, st, a);
  return hipGetLastError ();
}

hipError_t launch_wave_f64 (const SolveArgs &a, int nc, bool keep, int grid, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  switch (nc)
  {
  case 16:
    return launch_wave_nc<16, 4> (a, keep, grid, st);
  case 32:
    return launch_wave_nc<32, 2> (a, keep, grid, st);
  case 64:
    return launch_wave_nc<64, 2> (a, keep, grid, st);
  default:
    return hipErrorInvalidValue;
  }
}

hipError_t launch_wgr_f64 (const SolveArgs &a, int nmax, bool keep, int grid, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  hipError_t e = launch_tile_f64 (a, nmax, keep, grid, st);
  if (e != hipSuccess)
    return e;
  return launch_bsub_f64 (a, nmax, st);
}

template <int MAXM>
static void launch_bsub_m (const SolveArgs &a, hipStream_t st)
{
  const int grid = std::max (1, std::min (a.nlist, 8 * hipla_device_info ().num_cu));
  const size_t sm = 3 * (64 * MAXM + 64) * sizeof (double);
  hipLaunchKernelGGL ((k_bsub_f64<MAXM>), dim3 (grid), dim3 (64), sm, st, a);
}

// back substitution + residual check of the cells a workgroup factor kernel just factored
static hipError_t launch_bsub_f64 (const SolveArgs &a, int nmax, hipStream_t st)
{
  if (nmax <= 128)
    launch_bsub_m<2> (a, st);
  else if (nmax <= 192)
    launch_bsub_m<3> (a, st);
  else if (nmax <= 256)
    launch_bsub_m<4> (a, st);
  else if (nmax <= 512)
    launch_bsub_m<8> (a, st);
  else
    launch_bsub_m<BLK_MAXM> (a, st);
  return hipGetLastError ();
}

template <int BS, int NB, bool US>
static void launch_blocked_m (const SolveArgs &a, int grid, size_t sm, bool keep, hipStream_t st)
{
  if (keep)
    hipLaunchKernelGGL ((k_blocked_f64<BS, NB, US, true>), dim3 (grid), dim3 (BS), sm, st, a);
  else
    hipLaunchKernelGGL ((k_blocked_f64<BS, NB, US, false>), dim3 (grid), dim3 (BS), sm, st, a);
}

hipError_t launch_bsub_cells (const SolveArgs &a, int nmax, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  return launch_bsub_f64 (a, nmax, st);
}

hipError_t launch_blocked_f64 (const SolveArgs &a, int nmax, int grid, bool keep, hipStream_t st, bool with_bsub)
{
  if (a.nlist <= 0)
    return hipSuccess;
  if (nmax > 64 * BLK_MAXM || !a.F || !a.Y)
    return hipErrorInvalidValue;
  const size_t sm = blocked_f64_smem (nmax);
  if (nmax <= 256)
  {
    // up to one cell per CU the per-cell latency decides (one-wave panel with look-ahead); with
    // more cells per CU, throughput (two four-wave workgroups per CU sharing the panel chains);
    // a device-side list length gets both launches, each working only in its own range
    const int lim = lookahead_f64_max_cells ();
    if (a.dcount)
    {
      SolveArgs lo = a, hi = a;
      lo.cnt_hi = lim;
      hi.cnt_lo = lim + 1;
      hipError_t e = launch_lookahead_f64 (lo, std::min (grid, lim), keep, st);
      if (e != hipSuccess)
        return e;
      launch_blocked_m<256, 16, true> (hi, grid, sm, keep, st);
    }
    else if (a.nlist <= lim)
    {
      hipError_t e = launch_lookahead_f64 (a, std::min (grid, a.nlist), keep, st);
      if (e != hipSuccess)
        return e;
    }
    else
      launch_blocked_m<256, 16, true> (a, grid, sm, keep, st);
  }
  else if (nmax <= 512)
    launch_blocked_m<BLK_BS, 16, true> (a, grid, sm, keep, st);
  else
    launch_blocked_m<BLK_BS, 8, false> (a, grid, sm, keep, st);
  hipError_t e = hipGetLastError ();
  if (e != hipSuccess || !with_bsub)
    return e;
  return launch_bsub_f64 (a, nmax, st);
}

// The order 175 < n <= 256 route for a device-side list its caller knows to hold at most one cell
// per CU (lookahead_f64_max_cells): the look-ahead factor kernel alone, then the back
// substitution, with no launch of the blocked kernel standing by for longer lists.  Correct for
// any list length (the kernel loops over its list), only slower than the blocked kernel past that
// bound.
hipError_t launch_lookahead_route_f64 (const SolveArgs &a, int nmax, bool keep, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  if (nmax > LA_ROUTE_MAXN || !a.F || !a.Y)
    return hipErrorInvalidValue;
  SolveArgs w = a;
  w.cnt_lo = 0;
  w.cnt_hi = 0x7fffffff;
  hipError_t e = launch_lookahead_f64 (w, std::min (lookahead_f64_max_cells (), a.nlist), keep, st);
  if (e != hipSuccess)
    return e;
  return launch_bsub_f64 (w, nmax, st);
}

// Name of the FP64 kernel instantiation a group of largest order nmax runs (as profilers print
// it), for the plan description.
void fp64_kernel_symbol (int nmax, bool keep, char *buf, size_t len, int count)
{
  const char *k = keep ? "true" : "false";
  const bool la = count >= 0 && count <= lookahead_f64_max_cells ();
  if (nmax <= 64)
  {
    const int nc = nmax <= 16 ? 16 : nmax <= 32 ? 32 : 64;
    snprintf (buf, len, "k_wave_f64<%d, %d, %s>", nc, nc == 16 ? 4 : 2, k);
  }
  else if (nmax <= wgr_f64_limit ())
  {
    char t[64];
    tile_kernel_symbol (nmax, keep, t, sizeof t);
    snprintf (buf, len, "%s + k_bsub_f64<%d>", t, nmax <= 128 ? 2 : 3);
  }
  else if (nmax <= 256 && la)
    snprintf (buf, len, "k_lookahead_f64<%s> + k_bsub_f64<%d>", k, nmax <= 128 ? 2 : nmax <= 192 ? 3 : 4);
  else if (nmax <= 512)
    snprintf (buf, len, "k_blocked_f64<%d, 16, true, %s> + k_bsub_f64<%d>", nmax <= 256 ? 256 : BLK_BS, k,
              nmax <= 128 ? 2 : nmax <= 192 ? 3 : nmax <= 256 ? 4 : 8);
  else
    snprintf (buf, len, "k_blocked_f64<%d, 8, false, %s> + k_bsub_f64<%d>", BLK_BS, k, BLK_MAXM);
}

// Best effort: AMD's runtime accepts dynamic LDS up to the device limit without the attribute;
// where the attribute is honoured it is set to the limit minus the kernel's static LDS.
template <typename K>
static void set_smem_attr (K kern, size_t max_smem)
{
  hipFuncAttributes fa;
  if (hipFuncGetAttributes (&fa, (const void *) kern) != hipSuccess)
    return;
  const size_t dyn = max_smem > fa.sharedSizeBytes ? max_smem - fa.sharedSizeBytes : 0;
  (void) hipFuncSetAttribute ((const void *) kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int) dyn);
  (void) hipGetLastError ();
}

hipError_t setup_solve_f64_attrs (size_t max_smem)
{
  set_smem_attr (k_wave_f64<16, 4, false>, max_smem);
  set_smem_attr (k_wave_f64<32, 4, false>, max_smem);
  set_smem_attr (k_wave_f64<64, 2, false>, max_smem);
  set_smem_attr (k_wave_f64<16, 4, true>, max_smem);
  set_smem_attr (k_wave_f64<32, 4, true>, max_smem);
  set_smem_attr (k_wave_f64<64, 2, true>, max_smem);
  set_smem_attr (k_blocked_f64<256, 16, true, false>, max_smem);
  set_smem_attr (k_blocked_f64<256, 16, true, true>, max_smem);
  set_smem_attr (k_blocked_f64<BLK_BS, 16, true, false>, max_smem);
  set_smem_attr (k_blocked_f64<BLK_BS, 16, true, true>, max_smem);
  set_smem_attr (k_blocked_f64<BLK_BS, 8, false, false>, max_smem);
  set_smem_attr (k_blocked_f64<BLK_BS, 8, false, true>, max_smem);
  set_smem_attr (k_bsub_f64<2>, max_smem);
  set_smem_attr (k_bsub_f64<3>, max_smem);
  set_smem_attr (k_bsub_f64<4>, max_smem);
  set_smem_attr (k_bsub_f64<8>, max_smem);
  set_smem_attr (k_bsub_f64<BLK_MAXM>, max_smem);
  return setup_lookahead_attrs (max_smem);
}

// Kernel shapes of one FP64 kernel family (equal values: every variant keeps GSL's order), for
// the autotuner.  Family f = 0, 1, 2 (n <= 16, 32, 64): one wave per matrix with WPB waves per
// workgroup, variant 0 the default WPB, 1 the other; family 3 (64 < n <= wgr_f64_limit ()):
// variant 0 the register-tiled kernel, 1 the blocked kernel; family 4: the blocked kernel with
// 256 (variant 0) or 512 (variant 1, n <= 256 only) threads.
int fp64_family_of (int nmax)
{
  return nmax <= 16 ? 0 : nmax <= 32 ? 1 : nmax <= 64 ? 2 : nmax <= wgr_f64_limit () ? 3 : 4;
}

int fp64_variants (int family, int nmax)
{
  return family < 4 ? 2 : (nmax <= 256 ? 2 : 1);
}

hipError_t launch_fp64_variant (const SolveArgs &a, int nmax, int variant, bool keep, int grid, hipStream_t st,
                                bool with_bsub)
{
  if (a.nlist <= 0)
    return hipSuccess;
  const int fam = fp64_family_of (nmax);
  if (variant <= 0 || variant >= fp64_variants (fam, nmax))
  {
    if (fam < 3)
      return launch_wave_f64 (a, fam == 0 ? 16 : fam == 1 ? 32 : 64, keep, grid, st);
    if (fam == 3)
    {
      hipError_t e = launch_tile_f64 (a, nmax, keep, grid, st);
      return (e != hipSuccess || !with_bsub) ? e : launch_bsub_f64 (a, nmax, st);
    }
    return launch_blocked_f64 (a, nmax, grid > 0 ? grid : std::min (a.nlist, blocked_f64_wgs_per_cu (nmax) * hipla_device_info ().num_cu), keep, st, with_bsub);
  }
  switch (fam)
  {
  case 0:
    return launch_wave_nc<16, 2> (a, keep, grid, st);
  case 1:
    return launch_wave_nc<32, 4> (a, keep, grid, st);
  case 2:
    return launch_wave_nc<64, 1> (a, keep, grid, st);
  case 3:
    return launch_blocked_f64 (a, nmax, grid > 0 ? grid : std::min (a.nlist, blocked_f64_wgs_per_cu (nmax) * hipla_device_info ().num_cu), keep, st, with_bsub);
  default:
  {
    if (!a.F || !a.Y)
      return hipErrorInvalidValue;
    const int g2 = grid > 0 ? grid : std::min (a.nlist, blocked_f64_wgs_per_cu (nmax) * hipla_device_info ().num_cu);
    launch_blocked_m<BLK_BS, 16, true> (a, g2, blocked_f64_smem (nmax), keep, st);
    hipError_t e = hipGetLastError ();
    return (e != hipSuccess || !with_bsub) ? e : launch_bsub_f64 (a, nmax, st);
  }
  }
}

// Name of the kernel instantiation(s) a variant runs (as profilers print it)
void fp64_variant_symbol (int nmax, int variant, bool keep, char *buf, size_t len, int count)
{
  const int fam = fp64_family_of (nmax);
  if (variant <= 0 || variant >= fp64_variants (fam, nmax))
  {
    fp64_kernel_symbol (nmax, keep, buf, len, count);
    return;
  }
  const char *k = keep ? "true" : "false";
  const int m = nmax <= 128 ? 2 : nmax <= 192 ? 3 : nmax <= 256 ? 4 : 8;
  if (fam < 3)
    snprintf (buf, len, "k_wave_f64<%d, %d, %s>", fam == 0 ? 16 : fam == 1 ? 32 : 64, fam == 0 ? 2 : fam == 1 ? 4 : 1, k);
  else if (fam == 3 && count >= 0 && count <= lookahead_f64_max_cells ())
    snprintf (buf, len, "k_lookahead_f64<%s> + k_bsub_f64<%d>", k, m);
  else if (fam == 3)
    snprintf (buf, len, "k_blocked_f64<256, 16, true, %s> + k_bsub_f64<%d>", k, m);
  else
    snprintf (buf, len, "k_blocked_f64<%d, 16, true, %s> + k_bsub_f64<%d>", BLK_BS, k, m);
}
