// k_mixed.hip -- equilibrated FP32 LU + FP64 residual refinement + FP64 fallback (gfx950).
//
// Same contract as solve_matrix (source/matrix_cpu.c:96-204: solve A x = b, residual-check
// codes 0/2/3, 4 when singular) over a ragged batch, computed as:
//   1. equilibrate: A_s = R A C with R, C exact powers of two (row max, then column max, in
//      [1/2, 1)); FP64 -> FP32 rounding is the only inexact step of the scaling;
//   2. factor P A_s = L U in FP32 with partial pivoting, the matrix resident in registers;
//   3. x_0 = C U^-1 L^-1 P R b, then up to max_refine steps of
//        r = b - A x  (FP64, the original A),  d = C U^-1 L^-1 P R r (FP32),  x += d (FP64)
//      stopping when ||d||/||x|| <= tol or the normwise backward error
//      ||r|| / (||A|| ||x|| + ||b||) reaches the FP64 level;
//   4. a cell that stalls (backward error not halving), runs out of steps or meets an exact zero
//      FP32 pivot is appended to a device list and re-solved by the FP64 GSL-order kernels.
// Residual-check codes come from t = A x of the final x (FP64, fused multiply-add).
//
// Two shapes:
//   k_wave_mixed<NC>   n <= NC <= 64: one wave per matrix, lane = row, the active trailing row in
//                      registers shifted left each step; U rows and L multipliers go to LDS.
//   k_blk_mixed<R,NW>  64 < n <= 256: one workgroup of NW waves per matrix; wave w keeps columns
//                      [16w, 16w+16) of every row in registers (lane l, slot r -> row l + 64 r);
//                      rows are never swapped (pivot order is tracked per row), one barrier per
//                      elimination step, multipliers double-buffered in LDS.
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"
#include "hipla_tune.h"

#include <algorithm>
#include <utility>

struct MixedArgs
{
  SolveArgs s;
  hipla_cell_stats *stats;
  int *fb;         // 5 lists of (1 + fbcap) ints: [count, cells...] per FP64 route
  int fbcap;
  int lds_limit;   // FP64 route thresholds: <=16, <=32, <=64 wave; <=lds_limit LDS; else blocked
  int max_refine;
  double tol;
  int no_fallback;
};

__device__ __forceinline__ int fp64_route (int n, int lds_limit)
{
  return n <= 16 ? 0 : n <= 32 ? 1 : n <= 64 ? 2 : n <= lds_limit ? 3 : 4;
}

__device__ __forceinline__ int exp_of (double v)
{
  int e = 0;
  if (v > 0.0 && v <= DBL_MAX)
    (void) frexp (v, &e);
  return e;
}

__device__ __forceinline__ int exp_of_f (float v)
{
  int e = 0;
  if (v > 0.0f && v <= FLT_MAX)
    (void) frexpf (v, &e);
  return e;
}

// Smallest FP32 pivot (equilibrated scale, entries <= 1) accepted: below it the FP32 factors are
// too far from A for refinement to be trusted and the cell goes to the FP64 kernels directly.
#define MIXED_PIVOT_GATE 7.62939453125e-06f   /* 2^-17 */

// Refinement decision after correction k (relc = ||d_k|| / ||x||): 0 continue, 1 converged,
// 2 hand back to FP64.  Convergence needs a correction below tol after at least one refinement
// step; a correction that does not halve, or running out of steps, hands the cell back.
__device__ __forceinline__ int refine_decision (int it, double relc, double prev_relc, const MixedArgs &m)
{
  if (it >= 1 && relc <= m.tol)
    return 1;
  if (it >= 1 && relc > 0.5 * prev_relc)
    return 2;
  if (it >= m.max_refine)
    return 2;
  return 0;
}

__device__ __forceinline__ void finish_cell (const MixedArgs &m, int cell, int n, int decision, int code, int steps,
                                             double relc, double eta, bool leader)
{
  if (!leader)
    return;
  if (m.stats)
  {
    m.stats[cell].refine_steps = steps;
    m.stats[cell].fallback = (decision == 2 && !m.no_fallback) ? 1 : 0;
    m.stats[cell].rel_correction = relc;
    m.stats[cell].backward_err = eta;
  }
  if (decision == 2 && !m.no_fallback)
  {
    const int r = fp64_route (n, m.lds_limit);
    int *lst = m.fb + r * (m.fbcap + 1);
    const int slot = atomicAdd (lst, 1);
    lst[1 + slot] = cell;
  }
  else
    m.s.codes[cell] = code;
}

// ============================================================================================
// k_wave_mixed<NC>: one wave per matrix
// ============================================================================================
template <int NC>
struct WaveMixLayout
{
  static constexpr int LDU = NC + 2;   // U row stride (odd stride - 1 for the column reads)
  static constexpr int LDL = NC + 1;   // L row stride (odd)
  static constexpr int STAGE = NC * (NC | 1) * 2;      // floats: FP64 staging n x (n|1)
  static constexpr int UL = NC * LDU + NC * LDL;       // floats: U + L
  static constexpr int MAIN = STAGE > UL ? STAGE : UL; // union of the two
  // + prow[NC] int, cexp[NC] int, rexp[NC] int, ys[NC] float, xs[NC] double, rs[NC] double
  static constexpr int FLOATS = MAIN + 4 * NC + 2 * NC + 2 * NC;
  static constexpr size_t BYTES = (size_t) FLOATS * 4;
};

template <int NC>
__global__ __launch_bounds__ (64) void k_wave_mixed (MixedArgs m)
{
  using Lay = WaveMixLayout<NC>;
  extern __shared__ float wsm[];
  float *Us = wsm;
  float *Ls = wsm + NC * Lay::LDU;
  double *stage = reinterpret_cast<double *> (wsm);
  int *prow = reinterpret_cast<int *> (wsm + Lay::MAIN);
  int *cexp = prow + NC;
  int *rexp = cexp + NC;
  float *ys = reinterpret_cast<float *> (rexp + NC);
  double *xs = reinterpret_cast<double *> (ys + NC);
  double *rs = xs + NC;
  const int lane = threadIdx.x;
  const SolveArgs &s = m.s;

  for (int item = blockIdx.x; item < s.nlist; item += gridDim.x)
  {
    wave_lds_fence ();
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    const bool in = lane < n;

    // ---- coalesced load through LDS, row exponents ----
    const int pitch = n | 1;
    for (int r = 0; r < n; r++)
      if (lane < n)
        stage[r * pitch + lane] = A[(size_t) r * n + lane];
    wave_lds_fence ();
    double rmax = 0.0, rsum = 0.0;
    if (in)
      for (int j = 0; j < n; j++)
      {
        const double v = fabs (stage[lane * pitch + j]);
        rmax = fmax (rmax, v);
        rsum += v;
      }
    const int E = exp_of (rmax);
    if (in)
      rexp[lane] = E;
    const double anorm = wave_max_f64 (rsum);
    const double bv = in ? b[lane] : 0.0;
    const double bnorm = wave_max_f64 (fabs (bv));
    wave_lds_fence ();
    // column exponents (lane = column): max_i |a_ij| 2^-E_i
    {
      float cmax = 0.0f;
      if (in)
        for (int i = 0; i < n; i++)
          cmax = fmaxf (cmax, (float) fabs (ldexp (stage[i * pitch + lane], -rexp[i])));
      if (in)
        cexp[lane] = exp_of_f (cmax);
    }
    wave_lds_fence ();
    float a[NC];
#pragma unroll
    for (int j = 0; j < NC; j++)
      a[j] = (in && j < n) ? (float) ldexp (stage[lane * pitch + j], -E - cexp[j]) : 0.0f;
    wave_lds_fence ();   // staging is dead from here; U / L overwrite it

    // ---- FP32 LU, shift scheme ----
    bool alive = in;
    int pstep = -1;
    int zero_pivot = 0;
    for (int k = 0; k < n; k++)
    {
      int key = alive ? pivot_key32 (a[0]) : -1;
      int pr = alive ? lane : 0x7fffffff;
      wave_argmax_keypos32 (key, pr);
      if (key <= 0)
      {
        zero_pivot = 1;
        break;
      }
      float *Uk = Us + k * Lay::LDU;
      if (lane == pr)
      {
#pragma unroll
        for (int t = 0; t < NC; t++)
          if (t < n - k)
            Uk[t] = a[t];
        prow[k] = pr;
        alive = false;
        pstep = k;
      }
      wave_lds_fence ();
      if (alive)
      {
        const float rcp = 1.0f / Uk[0];
        const float l = a[0] * rcp;
        Ls[lane * Lay::LDL + k] = l;
#pragma unroll
        for (int t = 1; t < NC; t++)
          if (t < n - k)
            a[t - 1] = fmaf (-l, Uk[t], a[t]);
      }
    }
    wave_lds_fence ();

    // pivot gate: the smallest |U_kk| of the equilibrated FP32 factors
    float pmin = 1.0f;
    if (!zero_pivot)
      for (int k = lane; k < n; k += 64)
        pmin = fminf (pmin, fabsf (Us[k * Lay::LDU]));
    pmin = -wave_max_f32 (-pmin);

    int decision = 2, code = 0, steps = 0;
    double relc = 1.0, eta = 1.0;
    if (!zero_pivot && pmin >= MIXED_PIVOT_GATE)
    {
      double xj = 0.0;          // lane j holds x_j
      double prev_relc = 1e300;
      double rv = bv;           // lane i holds r_i (starts as b: x = 0)
      for (int it = 0;; it++)
      {
        if (it > 0)
        {
          // r = b - A x in FP64 (lane = row), codes of the current x
          double t = 0.0;
          if (in)
          {
            const double *Ai = A + (size_t) lane * n;
            for (int j = 0; j < n; j++)
              t = fma (Ai[j], xs[j], t);
          }
          rv = bv - t;
          const double xnorm = wave_max_f64 (fabs (xj));
          eta = wave_max_f64 (fabs (rv)) / (anorm * xnorm + bnorm);
          const int f = in ? residual_row_code (t, bv) : 0;
          const unsigned long long fbits = __ballot (f != 0);
          code = fbits ? __shfl (f, __ffsll ((unsigned long long) fbits) - 1) : 0;
        }
        // rhs = R r normalised to [1/2, 1) in FP32
        const double rsc = in ? ldexp (rv, -E) : 0.0;
        const int S = exp_of (wave_max_f64 (fabs (rsc)));
        float v = (float) ldexp (rsc, -S);
        // forward: L y = P rhs (lane = row)
        for (int k = 0; k < n; k++)
        {
          const int p = prow[k];
          const float yk = __int_as_float (__builtin_amdgcn_readlane (__float_as_int (v), p));
          if (lane == 0)
            ys[k] = yk;
          if (pstep > k)
            v = fmaf (-Ls[lane * Lay::LDL + k], yk, v);
        }
        wave_lds_fence ();
        // backward: U z = y (lane = step)
        float z = in ? ys[lane] : 0.0f;
        for (int k = n - 1; k >= 0; k--)
        {
          const float zk = __int_as_float (__builtin_amdgcn_readlane (__float_as_int (z), k));
          const float xk = zk / Us[k * Lay::LDU];
          if (lane == k)
            z = xk;
          else if (lane < k)
            z = fmaf (-Us[lane * Lay::LDU + (k - lane)], xk, z);
        }
        // lane j now holds the scaled correction for unknown j
        const double dj = in ? ldexp ((double) z, S - cexp[lane]) : 0.0;
        xj += dj;
        relc = wave_max_f64 (fabs (dj)) / fmax (wave_max_f64 (fabs (xj)), 1e-300);
        if (in)
          xs[lane] = xj;
        wave_lds_fence ();
        const int d = refine_decision (it, relc, prev_relc, m);
        prev_relc = relc;
        if (d)
        {
          decision = d;
          steps = it;
          break;
        }
      }
      if (decision != 2 || m.no_fallback)
        if (in)
          s.X[s.offV[cell] + lane] = xj;
    }
    finish_cell (m, cell, n, decision, code, steps, relc, eta, lane == 0);
  }
}

// ============================================================================================
// k_blk_mixed<R, NW>: column-block-per-wave workgroup kernel, 64 < n <= 256
// ============================================================================================
constexpr int CB = 16;

template <int R>
__device__ __forceinline__ float bcast_row (const float (&v)[R], int row)
{
  const int slot = row >> 6;
  float t = v[0];
#pragma unroll
  for (int rr = 1; rr < R; rr++)
    if (slot == rr)
      t = v[rr];
  return __int_as_float (__builtin_amdgcn_readlane (__float_as_int (t), row & 63));
}

template <int R>
struct BlkState
{
  float a[R][CB];   // a[rr][c]: row lane + 64 rr, column 16 w + c; L multipliers replace A in place
  int alive[R];     // 1 while the row is not yet chosen as pivot (an int: keeps it out of lane masks)
  int pstep[R];     // elimination step at which the row was chosen (-1: not yet / no row)
};

constexpr int TP = CB + 1;   // LDS panel row pitch (odd: conflict-free column reads)

template <int R>
__device__ __forceinline__ void tile_to_panel (const BlkState<R> &st, float *T, int lane)
{
#pragma unroll
  for (int rr = 0; rr < R; rr++)
#pragma unroll
    for (int c = 0; c < CB; c++)
      T[(lane + 64 * rr) * TP + c] = st.a[rr][c];
}

template <int R>
__device__ __forceinline__ void panel_to_tile (BlkState<R> &st, const float *T, int lane)
{
#pragma unroll
  for (int rr = 0; rr < R; rr++)
#pragma unroll
    for (int c = 0; c < CB; c++)
      st.a[rr][c] = T[(lane + 64 * rr) * TP + c];
}

template <int R, int NW>
__global__ __launch_bounds__ (64 * NW) void k_blk_mixed (MixedArgs m)
{
  constexpr int NMAX = 64 * R;
  __shared__ float s_T[NMAX * TP];      // the owner wave's column block during its 16 steps
  __shared__ int s_rexp[NMAX];
  __shared__ int s_cexp[NMAX];
  __shared__ double s_rsum[NMAX];
  __shared__ float s_mult[2][NMAX];
  __shared__ int s_prow[NMAX];
  __shared__ float s_pval[NMAX];
  __shared__ float s_prs[NW][CB];
  __shared__ double s_x[NMAX];
  __shared__ double s_r[NMAX];
  __shared__ float s_v[NMAX];
  __shared__ float s_y[NMAX];
  __shared__ double s_red[3][NW];
  __shared__ int s_flag;
  __shared__ int s_fail;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const SolveArgs &s = m.s;
  const int col0 = w * CB;

  for (int item = blockIdx.x; item < s.nlist; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    const int nwu = (n + CB - 1) / CB;
    for (int i = tid; i < NMAX; i += 64 * NW)
    {
      s_rexp[i] = -100000;
      s_rsum[i] = 0.0;
      s_cexp[i] = 0;           // column max bits first, exponent after
    }
    if (tid == 0)
      s_flag = 0;
    __syncthreads ();

    // ---- load my column block (one 128-byte line per lane per row), partial row exponents ----
    BlkState<R> st;
    int epart[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++)
    {
      const int row = lane + 64 * rr;
      double vals[CB];
      double pm = 0.0, ps = 0.0;
#pragma unroll
      for (int c = 0; c < CB; c++)
      {
        const int col = col0 + c;
        vals[c] = (row < n && col < n) ? A[(size_t) row * n + col] : 0.0;
        pm = fmax (pm, fabs (vals[c]));
        ps += fabs (vals[c]);
      }
      epart[rr] = exp_of (pm);
#pragma unroll
      for (int c = 0; c < CB; c++)
        st.a[rr][c] = (float) ldexp (vals[c], -epart[rr]);
      if (row < n && col0 < n)
      {
        if (pm > 0.0)
          atomicMax (&s_rexp[row], epart[rr]);
        atomicAdd (&s_rsum[row], ps);
      }
      __builtin_amdgcn_sched_barrier (0);
    }
    __syncthreads ();
    // fix up to the full-row exponent; column maxima through LDS integer atomics
#pragma unroll
    for (int rr = 0; rr < R; rr++)
    {
      const int row = lane + 64 * rr;
      const int E = (row < n && s_rexp[row] > -100000) ? s_rexp[row] : 0;
#pragma unroll
      for (int c = 0; c < CB; c++)
        st.a[rr][c] = ldexpf (st.a[rr][c], epart[rr] - E);
    }
#pragma unroll
    for (int c = 0; c < CB; c++)
    {
      float cm = 0.0f;
#pragma unroll
      for (int rr = 0; rr < R; rr++)
        cm = fmaxf (cm, fabsf (st.a[rr][c]));
      if (col0 + c < n && cm > 0.0f)
        atomicMax (&s_cexp[col0 + c], __float_as_int (cm));
    }
    __syncthreads ();
#pragma unroll
    for (int c = 0; c < CB; c++)
    {
      const int F = (col0 + c < n) ? exp_of_f (__int_as_float (s_cexp[col0 + c])) : 0;
#pragma unroll
      for (int rr = 0; rr < R; rr++)
        st.a[rr][c] = ldexpf (st.a[rr][c], -F);
    }
    __syncthreads ();
    if (lane == 0)
      for (int c = 0; c < CB; c++)
        if (col0 + c < n)
          s_cexp[col0 + c] = exp_of_f (__int_as_float (s_cexp[col0 + c]));
    // norms, b, zero solution
    {
      double am = 0.0, bm = 0.0;
      for (int i = tid; i < n; i += 64 * NW)
      {
        am = fmax (am, s_rsum[i]);
        const double bv = b[i];
        s_r[i] = bv;
        bm = fmax (bm, fabs (bv));
        s_x[i] = 0.0;
      }
      am = wave_max_f64 (am);
      bm = wave_max_f64 (bm);
      if (lane == 0)
      {
        s_red[0][w] = am;
        s_red[1][w] = bm;
      }
    }
#pragma unroll
    for (int rr = 0; rr < R; rr++)
    {
      st.alive[rr] = ((lane + 64 * rr) < n) ? 1 : 0;
      st.pstep[rr] = -1;
    }
    __syncthreads ();
    double anorm = 0.0, bnorm = 0.0;
#pragma unroll
    for (int q = 0; q < NW; q++)
    {
      anorm = fmax (anorm, s_red[0][q]);
      bnorm = fmax (bnorm, s_red[1][q]);
    }
    __syncthreads ();

    // ---- FP32 LU: step k = 16 blk + cc; the owner wave blk works on its block in LDS ----
    int zero_pivot = 0;
    for (int blk = 0; blk < nwu && !zero_pivot; blk++)
    {
      if (w == blk)
      {
        tile_to_panel<R> (st, s_T, lane);
        wave_lds_fence ();
      }
      for (int cc = 0; cc < CB; cc++)
      {
        const int k = blk * CB + cc;
        if (k >= n)
          break;
        const int buf = k & 1;
        if (w == blk)
        {
          int key = -1, pr = 0x7fffffff;
#pragma unroll
          for (int rr = 0; rr < R; rr++)
          {
            const int q = st.alive[rr] ? pivot_key32 (s_T[(lane + 64 * rr) * TP + cc]) : -1;
            const bool better = q > key;
            key = better ? q : key;
            pr = better ? lane + 64 * rr : pr;
          }
          wave_argmax_keypos32 (key, pr);
          if (key <= 0)
          {
            if (lane == 0)
              s_flag = 1;
          }
          else
          {
            const float pv = s_T[pr * TP + cc];
            const float rcp = 1.0f / pv;
#pragma unroll
            for (int rr = 0; rr < R; rr++)
            {
              const int row = lane + 64 * rr;
              const bool me = row == pr;
              st.pstep[rr] = me ? k : st.pstep[rr];
              st.alive[rr] = me ? 0 : st.alive[rr];
              const float l = st.alive[rr] ? s_T[row * TP + cc] * rcp : 0.0f;
              const float old = s_T[row * TP + cc];
              s_T[row * TP + cc] = st.alive[rr] ? l : old;
              s_mult[buf][row] = l;
            }
            if (lane == 0)
            {
              s_prow[k] = pr;
              s_pval[k] = pv;
            }
          }
        }
        __syncthreads ();
        if (s_flag)
        {
          zero_pivot = 1;
          break;
        }
        const int pr = s_prow[k];
        if (w == blk)
        {
          // trailing columns of the panel: T[row][c] -= l_row * T[pr][c], c > cc
#pragma unroll
          for (int rr = 0; rr < R; rr++)
          {
            const int row = lane + 64 * rr;
            const float l = s_mult[buf][row];
            for (int c = cc + 1; c < CB; c++)
              s_T[row * TP + c] = fmaf (-l, s_T[pr * TP + c], s_T[row * TP + c]);
          }
          wave_lds_fence ();
        }
        else
        {
#pragma unroll
          for (int rr = 0; rr < R; rr++)
          {
            const bool me = (lane + 64 * rr == pr);
            st.pstep[rr] = me ? k : st.pstep[rr];
            st.alive[rr] = me ? 0 : st.alive[rr];
          }
          if (w > blk && col0 < n)
          {
            const int prs = pr >> 6;
            if (lane == (pr & 63))
            {
#pragma unroll
              for (int rr = 0; rr < R; rr++)
                if (rr == prs)
                {
#pragma unroll
                  for (int c = 0; c < CB; c++)
                    s_prs[w][c] = st.a[rr][c];
                }
            }
            wave_lds_fence ();
            float u[CB];
#pragma unroll
            for (int c = 0; c < CB; c++)
              u[c] = s_prs[w][c];
#pragma unroll
            for (int rr = 0; rr < R; rr++)
            {
              const float l = s_mult[buf][lane + 64 * rr];
#pragma unroll
              for (int c = 0; c < CB; c++)
                st.a[rr][c] = fmaf (-l, u[c], st.a[rr][c]);
            }
            wave_lds_fence ();
          }
        }
      }
      if (w == blk)
        panel_to_tile<R> (st, s_T, lane);
      __syncthreads ();
    }

    // ---- pivot gate, then the refinement loop ----
    float pmin = 1.0f;
    if (!zero_pivot)
      for (int k = tid; k < n; k += 64 * NW)
        pmin = fminf (pmin, fabsf (s_pval[k]));
    pmin = -wave_max_f32 (-pmin);
    if (lane == 0)
      s_red[2][w] = pmin;
    __syncthreads ();
#pragma unroll
    for (int q = 0; q < NW; q++)
      pmin = fminf (pmin, (float) s_red[2][q]);
    __syncthreads ();
    int decision = 2, code = 0, steps = 0;
    double relc = 1.0, eta = 1.0, prev_relc = 1e300;
    if (!zero_pivot && pmin >= MIXED_PIVOT_GATE)
    {
      for (int it = 0;; it++)
      {
        if (it > 0)
        {
          // residual: waves over rows, lanes over columns (coalesced), deterministic sums
          if (tid == 0)
            s_fail = 0x7fffffff;
          __syncthreads ();
          double rmax = 0.0;
          for (int row = w; row < n; row += NW)
          {
            const double *Ar = A + (size_t) row * n;
            double t = 0.0;
            for (int j = lane; j < n; j += 64)
              t = fma (Ar[j], s_x[j], t);
            t = wave_sum_f64 (t);
            const double bv = b[row];
            const double r = bv - t;
            if (lane == 0)
            {
              s_r[row] = r;
              const int f = residual_row_code (t, bv);
              if (f)
                atomicMin (&s_fail, (row << 2) | f);
            }
            rmax = fmax (rmax, fabs (r));
          }
          double xm = 0.0;
          for (int j = tid; j < n; j += 64 * NW)
            xm = fmax (xm, fabs (s_x[j]));
          xm = wave_max_f64 (xm);
          if (lane == 0)
          {
            s_red[0][w] = rmax;
            s_red[1][w] = xm;
          }
          __syncthreads ();
          double rn = 0.0, xn = 0.0;
#pragma unroll
          for (int q = 0; q < NW; q++)
          {
            rn = fmax (rn, s_red[0][q]);
            xn = fmax (xn, s_red[1][q]);
          }
          eta = rn / (anorm * xn + bnorm);
          code = (s_fail == 0x7fffffff) ? 0 : (s_fail & 3);
          __syncthreads ();
        }
        // rhs = R r, normalised: S from the max of |2^-E_i r_i|
        {
          double sm = 0.0;
          for (int i = tid; i < n; i += 64 * NW)
          {
            const int Ei = (s_rexp[i] > -100000) ? s_rexp[i] : 0;
            sm = fmax (sm, fabs (ldexp (s_r[i], -Ei)));
          }
          sm = wave_max_f64 (sm);
          if (lane == 0)
            s_red[2][w] = sm;
        }
        __syncthreads ();
        double smax = 0.0;
#pragma unroll
        for (int q = 0; q < NW; q++)
          smax = fmax (smax, s_red[2][q]);
        const int S = exp_of (smax);
        for (int i = tid; i < n; i += 64 * NW)
        {
          const int Ei = (s_rexp[i] > -100000) ? s_rexp[i] : 0;
          s_v[i] = (float) ldexp (s_r[i], -Ei - S);
        }
        __syncthreads ();
        // forward: L y = P rhs, one column block (wave) at a time through the LDS panel
        for (int blk = 0; blk < nwu; blk++)
        {
          if (w == blk)
          {
            tile_to_panel<R> (st, s_T, lane);
            float v[R];
#pragma unroll
            for (int rr = 0; rr < R; rr++)
              v[rr] = (lane + 64 * rr < n) ? s_v[lane + 64 * rr] : 0.0f;
            wave_lds_fence ();
            for (int cc = 0; cc < CB; cc++)
            {
              const int k = blk * CB + cc;
              if (k >= n)
                break;
              const float yk = bcast_row<R> (v, s_prow[k]);
              if (lane == 0)
                s_y[k] = yk;
#pragma unroll
              for (int rr = 0; rr < R; rr++)
                v[rr] = fmaf ((st.pstep[rr] > k) ? -s_T[(lane + 64 * rr) * TP + cc] : 0.0f, yk, v[rr]);
            }
#pragma unroll
            for (int rr = 0; rr < R; rr++)
              if (lane + 64 * rr < n)
                s_v[lane + 64 * rr] = v[rr];
          }
          __syncthreads ();
        }
        // backward: U x = y with z_row = y_{pstep(row)}
        for (int blk = nwu - 1; blk >= 0; blk--)
        {
          if (w == blk)
          {
            tile_to_panel<R> (st, s_T, lane);
            float z[R];
#pragma unroll
            for (int rr = 0; rr < R; rr++)
              z[rr] = (blk == nwu - 1) ? (st.pstep[rr] >= 0 ? s_y[st.pstep[rr]] : 0.0f)
                                       : ((lane + 64 * rr < n) ? s_v[lane + 64 * rr] : 0.0f);
            wave_lds_fence ();
            for (int cc = CB - 1; cc >= 0; cc--)
            {
              const int k = blk * CB + cc;
              if (k >= n)
                continue;
              const float xk = bcast_row<R> (z, s_prow[k]) / s_pval[k];
              if (lane == 0)
                s_y[k] = xk;
#pragma unroll
              for (int rr = 0; rr < R; rr++)
                z[rr] = fmaf ((st.pstep[rr] >= 0 && st.pstep[rr] < k) ? -s_T[(lane + 64 * rr) * TP + cc] : 0.0f, xk,
                              z[rr]);
            }
#pragma unroll
            for (int rr = 0; rr < R; rr++)
              if (lane + 64 * rr < n)
                s_v[lane + 64 * rr] = z[rr];
          }
          __syncthreads ();
        }
        // x += C d (FP64)
        double dm = 0.0, xm = 0.0;
        for (int j = tid; j < n; j += 64 * NW)
        {
          const double d = ldexp ((double) s_y[j], S - s_cexp[j]);
          const double xv = s_x[j] + d;
          s_x[j] = xv;
          dm = fmax (dm, fabs (d));
          xm = fmax (xm, fabs (xv));
        }
        dm = wave_max_f64 (dm);
        xm = wave_max_f64 (xm);
        if (lane == 0)
        {
          s_red[0][w] = dm;
          s_red[1][w] = xm;
        }
        __syncthreads ();
        double dn = 0.0, xn = 0.0;
#pragma unroll
        for (int q = 0; q < NW; q++)
        {
          dn = fmax (dn, s_red[0][q]);
          xn = fmax (xn, s_red[1][q]);
        }
        relc = dn / fmax (xn, 1e-300);
        __syncthreads ();
        const int d = refine_decision (it, relc, prev_relc, m);
        prev_relc = relc;
        if (d)
        {
          decision = d;
          steps = it;
          break;
        }
      }
      if (decision != 2 || m.no_fallback)
        for (int j = tid; j < n; j += 64 * NW)
          s.X[s.offV[cell] + j] = s_x[j];
    }
    finish_cell (m, cell, n, decision, code, steps, relc, eta, tid == 0);
    __syncthreads ();
  }
}

// Stats of cells above 256 that the FP64 kernels solve directly: fallback = 2.
__global__ void k_mark_fp64_stats (const int *list, int count, hipla_cell_stats *st)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count)
  {
    const int c = list[i];
    st[c].refine_steps = 0;
    st[c].fallback = 2;
    st[c].rel_correction = 0.0;
    st[c].backward_err = 0.0;
  }
}

// Zero the fallback list counters (a kernel, so the sequence stays graph-capturable).
__global__ void k_fb_reset (int *fb, int fbcap)
{
  if (threadIdx.x < 5)
    fb[threadIdx.x * (fbcap + 1)] = 0;
}

// ============================================================================================
// routing, launch
// ============================================================================================
void route_mixed (int n, int &kind, int &param)
{
  if (n <= 16)
  {
    kind = KK_WAVE_MIX;
    param = 16;
  }
  else if (n <= 32)
  {
    kind = KK_WAVE_MIX;
    param = 32;
  }
  else if (n <= 64)
  {
    kind = KK_WAVE_MIX;
    param = 64;
  }
  else if (n <= 256)
  {
    kind = KK_LDS_MIX;            // register-resident column blocks; param = waves (16 columns each)
    param = (n + CB - 1) / CB;
  }
  else
  {
    kind = KK_BLOCKED_MIX;        // above 256 the FP64 kernels solve the cell directly
    param = 0;
  }
}

int64_t blocked_mixed_ws_floats (int nmax)
{
  return (int64_t) nmax * nmax;
}

template <int NC>
static void launch_wave_mixed_nc (const MixedArgs &m, hipStream_t st)
{
  hipLaunchKernelGGL ((k_wave_mixed<NC>), dim3 (m.s.nlist), dim3 (64), WaveMixLayout<NC>::BYTES, st, m);
}

template <int NW>
static void launch_blk_nw (const MixedArgs &m, hipStream_t st)
{
  constexpr int R = (NW * CB + 63) / 64;
  hipLaunchKernelGGL ((k_blk_mixed<R, NW>), dim3 (m.s.nlist), dim3 (64 * NW), 0, st, m);
}

static hipError_t launch_mixed_group (const MixedArgs &m, const Group &g, hipStream_t st)
{
  if (g.kind == KK_WAVE_MIX)
  {
    if (g.param == 16)
      launch_wave_mixed_nc<16> (m, st);
    else if (g.param == 32)
      launch_wave_mixed_nc<32> (m, st);
    else
      launch_wave_mixed_nc<64> (m, st);
    return hipGetLastError ();
  }
  switch (g.param)
  {
  case 5: launch_blk_nw<5> (m, st); break;
  case 6: launch_blk_nw<6> (m, st); break;
  case 7: launch_blk_nw<7> (m, st); break;
  case 8: launch_blk_nw<8> (m, st); break;
  case 9: launch_blk_nw<9> (m, st); break;
  case 10: launch_blk_nw<10> (m, st); break;
  case 11: launch_blk_nw<11> (m, st); break;
  case 12: launch_blk_nw<12> (m, st); break;
  case 13: launch_blk_nw<13> (m, st); break;
  case 14: launch_blk_nw<14> (m, st); break;
  case 15: launch_blk_nw<15> (m, st); break;
  case 16: launch_blk_nw<16> (m, st); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError ();
}

static SolveArgs group_args (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes, const Group &g)
{
  SolveArgs a;
  a.n = p->d_n;
  a.offA = p->d_offA;
  a.offV = p->d_offV;
  a.A = dA;
  a.B = dB;
  a.X = dX;
  a.codes = dcodes;
  a.list = p->d_list + g.first;
  a.nlist = g.count;
  a.LU = nullptr;
  a.piv = nullptr;
  a.dcount = nullptr;
  a.F = p->d_F;
  a.Y = p->d_Y;
  return a;
}

// FP64 GSL-order launch for a whole group (cells above 256, or a size class the tuner routes
// straight to FP64).
static int launch_group_direct_fp64 (hipla_plan p, const SolveArgs &a, const Group &g, hipla_cell_stats *dstats,
                                     hipStream_t s)
{
  HIPLA_CHECK (launch_fp64_variant (a, g.nmax, g.variant, false,
                                    g.nmax > wgr_f64_limit () ? std::max (1, std::min (g.count, g.grid)) : 0, s));
  if (dstats)
  {
    hipLaunchKernelGGL (k_mark_fp64_stats, dim3 ((g.count + 255) / 256), dim3 (256), 0, s, a.list, g.count, dstats);
    HIPLA_CHECK (hipGetLastError ());
  }
  return 0;
}

int enqueue_mixed_group (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes,
                         hipla_cell_stats *dstats, hipStream_t s)
{
  const Group &g = p->groups[gi];
  SolveArgs a = group_args (p, dA, dB, dX, dcodes, g);
  if (g.kind == KK_BLOCKED_MIX || g.direct)
    return launch_group_direct_fp64 (p, a, g, dstats, s);
  MixedArgs m;
  m.s = a;
  m.stats = dstats;
  m.fb = p->d_fallback;
  m.fbcap = p->ncells;
  m.lds_limit = wgr_f64_limit ();
  m.max_refine = p->max_refine;
  m.tol = p->tol;
  m.no_fallback = (p->flags & HIPLA_FLAG_NO_FALLBACK) ? 1 : 0;
  HIPLA_CHECK (launch_mixed_group (m, g, s));
  return 0;
}

int enqueue_mixed_reset (hipla_plan p, hipStream_t st)
{
  hipLaunchKernelGGL (k_fb_reset, dim3 (1), dim3 (64), 0, st, p->d_fallback, p->ncells);
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}

// FP64 re-solve of the cells handed back, one persistent launch per FP64 route
int enqueue_mixed_fallback (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes, hipStream_t st)
{
  if (p->flags & HIPLA_FLAG_NO_FALLBACK)
    return 0;
  const DeviceInfo &dev = hipla_device_info ();
  const int fbcap = p->ncells;
  const int nmax_all = *std::max_element (p->n.begin (), p->n.end ());
  for (int r = 0; r < 5; r++)
  {
    if (!p->fb_route_used[r])
      continue;
    SolveArgs a;
    a.n = p->d_n;
    a.offA = p->d_offA;
    a.offV = p->d_offV;
    a.A = dA;
    a.B = dB;
    a.X = dX;
    a.codes = dcodes;
    a.list = p->d_fallback + r * (fbcap + 1) + 1;
    a.nlist = fbcap;
    a.LU = nullptr;
    a.piv = nullptr;
    a.dcount = p->d_fallback + r * (fbcap + 1);
    a.F = p->d_F;
    a.Y = p->d_Y;
    if (r < 3)
      HIPLA_CHECK (launch_wave_f64 (a, r == 0 ? 16 : r == 1 ? 32 : 64, false, 2 * dev.num_cu, st));
    else if (r == 3)
      HIPLA_CHECK (launch_wgr_f64 (a, std::min (nmax_all, wgr_f64_limit ()), false, dev.num_cu, st));
    else
      HIPLA_CHECK (launch_blocked_f64 (a, nmax_all, blocked_f64_wgs_per_cu (nmax_all) * dev.num_cu, false, st));
  }
  return 0;
}

// FP64 kernel family of a group (wave 16 / 32 / 64, workgroup tile, blocked)
static int fp64_family (int nmax)
{
  return nmax <= 16 ? 0 : nmax <= 32 ? 1 : nmax <= 64 ? 2 : nmax <= wgr_f64_limit () ? 3 : 4;
}

int plan_enqueue_mixed (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipla_cell_stats *dstats, hipStream_t st)
{
  const int nb = (int) p->groups.size ();
  // no group on the FP32 path: no cell can be handed back, so no fallback list to clear or solve
  bool any_mixed = false;
  for (const Group &g : p->groups)
    any_mixed |= !(g.kind == KK_BLOCKED_MIX || g.direct);
  int rc = any_mixed ? enqueue_mixed_reset (p, st) : 0;
  if (rc)
    return rc;
  // Groups one after another on the caller's stream.  Adjacent groups that go straight to the
  // same FP64 kernel (their cells are contiguous in the plan's list) are solved by one launch, so
  // a persistent grid covers all of them at once.
  const DeviceInfo &dev = hipla_device_info ();
  if (!any_mixed)
  {
    // every group straight to FP64: merge.  When the last unit is a blocked one, it and its back
    // substitution run on the fork stream from the start while the smaller units run on the
    // caller's stream: the smaller kernels take the CU slots the blocked kernel's persistent
    // workgroups leave as they finish, and the back substitution (one latency-bound wave per
    // cell) overlaps them too.  The caller's stream waits for the fork at the end.
    std::vector<Group> units;
    for (int gi = 0; gi < nb;)
    {
      Group u = p->groups[gi];
      int gj = gi + 1;
      for (; gj < nb; gj++)
      {
        const Group &h = p->groups[gj];
        if (fp64_family (h.nmax) != fp64_family (u.nmax) || h.variant != u.variant || h.first != u.first + u.count)
          break;
        u.count += h.count;
        u.nmax = std::max (u.nmax, h.nmax);
      }
      u.grid = std::min (u.count, blocked_f64_wgs_per_cu (u.nmax) * dev.num_cu);
      units.push_back (u);
      gi = gj;
    }
    // the unit with the largest cells, when it runs the blocked kernel
    size_t big = 0;
    for (size_t i = 1; i < units.size (); i++)
      if (units[i].nmax > units[big].nmax)
        big = i;
    const bool fork = units.size () > 1 && units[big].nmax > wgr_f64_limit ();
    if (fork)
    {
      // to the back, the others keeping their order
      std::rotate (units.begin () + big, units.begin () + big + 1, units.end ());
      const Group &u = units.back ();
      SolveArgs a = group_args (p, dA, dB, dX, dcodes, u);
      HIPLA_CHECK (hipEventRecord (p->ev_fork, st));
      HIPLA_CHECK (hipStreamWaitEvent (p->fork_st, p->ev_fork, 0));
      HIPLA_CHECK (launch_fp64_variant (a, u.nmax, u.variant, false, std::max (1, u.grid), p->fork_st, true));
      if (dstats)
      {
        hipLaunchKernelGGL (k_mark_fp64_stats, dim3 ((u.count + 255) / 256), dim3 (256), 0, p->fork_st, a.list,
                            u.count, dstats);
        HIPLA_CHECK (hipGetLastError ());
      }
      HIPLA_CHECK (hipEventRecord (p->ev_join, p->fork_st));
      units.pop_back ();
    }
    for (const Group &u : units)
    {
      SolveArgs a = group_args (p, dA, dB, dX, dcodes, u);
      if ((rc = launch_group_direct_fp64 (p, a, u, dstats, st)))
        return rc;
    }
    if (fork)
      HIPLA_CHECK (hipStreamWaitEvent (st, p->ev_join, 0));
    return 0;
  }
  for (int gi = 0; gi < nb;)
  {
    const Group &g = p->groups[gi];
    if (!(g.kind == KK_BLOCKED_MIX || g.direct))
    {
      if ((rc = enqueue_mixed_group (p, gi, dA, dB, dX, dcodes, dstats, st)))
        return rc;
      gi++;
      continue;
    }
    Group u = g;
    int gj = gi + 1;
    for (; gj < nb; gj++)
    {
      const Group &h = p->groups[gj];
      if (!(h.kind == KK_BLOCKED_MIX || h.direct) || fp64_family (h.nmax) != fp64_family (u.nmax) ||
          h.variant != u.variant || h.first != u.first + u.count)
        break;
      u.count += h.count;
      u.nmax = std::max (u.nmax, h.nmax);
    }
    u.grid = std::min (u.count, blocked_f64_wgs_per_cu (u.nmax) * dev.num_cu);
    SolveArgs a = group_args (p, dA, dB, dX, dcodes, u);
    if ((rc = launch_group_direct_fp64 (p, a, u, dstats, st)))
      return rc;
    gi = gj;
  }
  return any_mixed ? enqueue_mixed_fallback (p, dA, dB, dX, dcodes, st) : 0;
}

hipError_t setup_mixed_attrs (size_t)
{
  return hipSuccess;
}
