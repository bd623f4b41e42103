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
// Three shapes, chosen per cell by the host plan:
//   k_wave_f64<NC>     n <= NC <= 64: one wave per matrix, lane = row, row in registers, U rows
//                      stashed in LDS as they are produced (no row swaps: positions tracked per
//                      lane, so GSL's tie rule uses positions, not lanes).
//   k_lds_f64          n <= ~141: one workgroup per matrix, [A|b] resident in LDS.
//   k_blocked_f64<NB>  larger n: right-looking blocked LU, NB-column panel factored in LDS,
//                      row swaps, U12 and the A22 update streamed through a per-workgroup
//                      HBM workspace (L2 resident), persistent over the cell list.
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

static size_t hipla_device_max_smem ()
{
  return hipla_device_info ().max_smem;
}

// Back substitution U x = y in dtrsv order, run by one wave.  Row i of U is at U + i*ldu
// (columns i..n-1 at their own column index), y_i at y[i*ys].  Lane l keeps x_{l + 64 m} in
// xr[m]; for row i every lane forms its products U_ij x_j in parallel and the wave subtracts them
// in ascending j through readlane (a chain of dependent FP64 subtractions, as GSL's).  Writes x to
// xs (LDS) when xs is not null and returns with xr filled.
template <int MAXM>
__device__ __forceinline__ void wave_backsub (const double *U, size_t ldu, const double *y, size_t ys, int n, int lane,
                                              double (&xr)[MAXM], double *xs)
{
#pragma unroll
  for (int m = 0; m < MAXM; m++)
    xr[m] = 0.0;
  for (int i = n - 1; i >= 0; i--)
  {
    const double *Ui = U + (size_t) i * ldu;
    double p[MAXM];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int j = lane + 64 * m;
      p[m] = (j > i && j < n) ? Ui[j] * xr[m] : 0.0;
    }
    double tmp = y[(size_t) i * ys];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
    {
      const int j0 = max (i + 1, 64 * m), j1 = min (n, 64 * m + 64);
      for (int j = j0; j < j1; j++)
        tmp = tmp - rl_f64 (p[m], j - 64 * m);
    }
    const double xi = tmp / Ui[i];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      if (lane + 64 * m == i)
        xr[m] = xi;
  }
  if (xs)
  {
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      if (lane + 64 * m < n)
        xs[lane + 64 * m] = xr[m];
  }
}

// ============================================================================================
// k_wave_f64: one wave per matrix
// ============================================================================================
template <int NC, int WPB, bool KEEP>
__global__ __launch_bounds__ (64 * WPB) void k_wave_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  constexpr int LDU = NC + 2;                        // U row k: [t] = U(k, k+t), [NC] = y_k
  constexpr int LDL = NC + 1;
  constexpr int PER_WAVE = NC * LDU + (KEEP ? NC * LDL : 0);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int neff = list_len (s);
  double *U = smem + wv * PER_WAVE;
  double *Ls = U + NC * LDU;                          // KEEP: Ls[row * LDL + k] = multiplier
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

    // coalesced load of A through LDS (odd pitch: at most 2-way bank conflicts on the row reads)
    const int pitch = n | 1;
    for (int r = 0; r < n; r++)
      for (int c = lane; c < n; c += 64)
        U[r * pitch + c] = A[(size_t) r * n + c];
    wave_lds_fence ();
    const bool in = lane < n;
    double a[NC];
#pragma unroll
    for (int j = 0; j < NC; j++)
      a[j] = (in && j < n) ? U[lane * pitch + j] : 0.0;
    double bb = in ? b[lane] : 0.0;
    wave_lds_fence ();

    bool alive = in;
    int pos = lane;
    int singular = 0;
    for (int k = 0; k < n; k++)
    {
      const int q = wave_pivot_lane (alive, pivot_key (a[0]), pos);
      const int ppos = __builtin_amdgcn_readlane (pos, q);
      double *Uk = U + k * LDU;
      if (lane == q)
      {
#pragma unroll
        for (int j = 0; j < NC; j += 2)
          if (j < n - k)
            *reinterpret_cast<double2 *> (Uk + j) = make_double2 (a[j], a[j + 1]);
        Uk[NC] = bb;
      }
      if (KEEP && lane == 0 && s.piv)
        s.piv[s.offV[cell] + k] = ppos;
      if (alive && pos == k)
        pos = ppos;
      if (lane == q)
      {
        alive = false;
        pos = k;
      }
      wave_lds_fence ();
      const double piv = Uk[0];
      singular |= (piv == 0.0);
      if (alive)
      {
        const double l = gsl_multiplier (a[0], piv);
        if (KEEP)
          Ls[lane * LDL + k] = l;
        bb = bb - l * Uk[NC];
#pragma unroll
        for (int j = 1; j < NC; j++)
          if (j < n - k)
            a[j - 1] = a[j] - l * Uk[j];
      }
    }
    wave_lds_fence ();

    if (KEEP && s.LU && in)
    {
      // row `lane` became pivot at step pos: LU row pos = [L multipliers | U row]
      double *out = s.LU + s.offA[cell] + (size_t) pos * n;
      for (int j = 0; j < pos; j++)
        out[j] = Ls[lane * LDL + j];
      const double *Ur = U + pos * LDU;
      for (int j = pos; j < n; j++)
        out[j] = Ur[j - pos];
    }
    if (singular)
    {
      if (lane == 0)
        s.codes[cell] = 4;
      continue;
    }

    // back substitution: U row i is U + i*LDU shifted by i, so hand wave_backsub a base that
    // makes row i's column index absolute
    double xr[1];
    {
      xr[0] = 0.0;
      for (int i = n - 1; i >= 0; i--)
      {
        const double *Ui = U + i * LDU;   // Ui[t] = U(i, i+t)
        const double p = (lane > i && lane < n) ? Ui[lane - i] * xr[0] : 0.0;
        double tmp = Ui[NC];
        for (int j = i + 1; j < n; j++)
          tmp = tmp - rl_f64 (p, j);
        const double xi = tmp / Ui[0];
        if (lane == i)
          xr[0] = xi;
      }
    }

    // residual check t = A x (gsl_blas_dgemv order) and codes 2 / 3 of the first failing row
    int fail = 0;
    if (in)
    {
      const double *Ai = A + (size_t) lane * n;
      double t = 0.0;
#pragma unroll 8
      for (int j = 0; j < n; j++)
      {
        const double p = rl_f64 (xr[0], j) * Ai[j];
        t = t + p;
      }
      const double tv = 0.0 + t;
      fail = residual_row_code (tv, b[lane]);
    }
    const unsigned long long fb = __ballot (fail != 0);
    int code = 0;
    if (fb)
      code = __builtin_amdgcn_readlane (fail, __ffsll ((long long) fb) - 1);
    if (in)
      s.X[s.offV[cell] + lane] = xr[0];
    if (lane == 0)
      s.codes[cell] = code;
  }
}

// ============================================================================================
// k_lds_f64: one workgroup per matrix, [A | b] resident in LDS
// ============================================================================================
template <int BS, bool KEEP>
__global__ __launch_bounds__ (BS) void k_lds_f64 (SolveArgs s)
{
  extern __shared__ double M[];
  __shared__ long long r_key[BS / 64];
  __shared__ int r_pos[BS / 64];
  __shared__ int s_fail;
  constexpr int NW = BS / 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int neff = list_len (s);
  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const int ld = (n + 1) | 1;
    double *xs = M + (size_t) n * ld;
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];

    for (int r = wv; r < n; r += NW)
      for (int c = lane; c < n; c += 64)
        M[r * ld + c] = A[(size_t) r * n + c];
    for (int r = tid; r < n; r += BS)
      M[r * ld + n] = b[r];
    if (tid == 0)
      s_fail = 0x7fffffff;
    __syncthreads ();

    int singular = 0;
    for (int k = 0; k < n; k++)
    {
      // pivot: rows k.. (one per thread for n <= BS), first strict max by row position
      long long key = -1;
      int pos = 0x7fffffff;
      for (int i = k + tid; i < n; i += BS)
      {
        const long long q = pivot_key (M[i * ld + k]);
        if (q > key)
        {
          key = q;
          pos = i;
        }
      }
      const int wl = wave_pivot_lane (key >= 0, key, pos);
      if (lane == 0)
      {
        r_key[wv] = wl >= 0 ? (long long) rl_u64 ((unsigned long long) key, wl) : -1;
        r_pos[wv] = wl >= 0 ? __builtin_amdgcn_readlane (pos, wl) : 0x7fffffff;
      }
      __syncthreads ();
      long long bk = r_key[0];
      int p = r_pos[0];
#pragma unroll
      for (int w = 1; w < NW; w++)
        if (r_key[w] > bk || (r_key[w] == bk && r_pos[w] < p))
        {
          bk = r_key[w];
          p = r_pos[w];
        }
      if (KEEP && tid == 0 && s.piv)
        s.piv[s.offV[cell] + k] = p;
      if (p != k)
        for (int c = tid; c <= n; c += BS)
        {
          const double t = M[k * ld + c];
          M[k * ld + c] = M[p * ld + c];
          M[p * ld + c] = t;
        }
      __syncthreads ();
      const double piv = M[k * ld + k];
      singular |= (piv == 0.0);
      const bool recip = fabs (piv) >= DBL_MIN;
      const double sc = 1.0 / piv;
      for (int i = k + 1 + wv; i < n; i += NW)
      {
        const double aik = M[i * ld + k];
        const double l = recip ? aik * sc : aik / piv;
        for (int j = k + 1 + lane; j <= n; j += 64)
          M[i * ld + j] = M[i * ld + j] - l * M[k * ld + j];
        if (lane == 0)
          M[i * ld + k] = l;
      }
      __syncthreads ();
    }

    if (KEEP && s.LU)
    {
      double *out = s.LU + s.offA[cell];
      for (int r = wv; r < n; r += NW)
        for (int c = lane; c < n; c += 64)
          out[(size_t) r * n + c] = M[r * ld + c];
    }
    if (singular)
    {
      if (tid == 0)
        s.codes[cell] = 4;
      __syncthreads ();
      continue;
    }

    if (wv == 0)
    {
      double xr[3];
      wave_backsub<3> (M, (size_t) ld, M + n, (size_t) ld, n, lane, xr, xs);
    }
    __syncthreads ();

    for (int i = tid; i < n; i += BS)
    {
      const double *Ai = A + (size_t) i * n;
      double t = 0.0;
#pragma unroll 8
      for (int j = 0; j < n; j++)
      {
        const double pr = xs[j] * Ai[j];
        t = t + pr;
      }
      const int f = residual_row_code (0.0 + t, b[i]);
      if (f)
        atomicMin (&s_fail, (i << 2) | f);
      s.X[s.offV[cell] + i] = xs[i];
    }
    __syncthreads ();
    if (tid == 0)
      s.codes[cell] = (s_fail == 0x7fffffff) ? 0 : (s_fail & 3);
    __syncthreads ();
  }
}

// U12 (trsm) and A22 (gemm) work of one column c >= k0 + nbk of the blocked LU, after the panel's
// row swaps.  Kept out of line so the L11 panel values are read per column instead of being
// hoisted into registers across the column loop.
template <int NB>
__device__ __forceinline__ void blocked_column (double *W, const double *P, int n, int k0, int nbk, int c)
{
  constexpr int PP = NB + 1;
  const int k1 = k0 + nbk;
  double u[NB];
#pragma unroll
  for (int kk = 0; kk < NB; kk++)
    u[kk] = (kk < nbk) ? W[(size_t) (k0 + kk) * n + c] : 0.0;
  // U12 = L11^-1 A12, row by row: u[r] receives its updates in ascending kk, the order of the
  // right-looking elimination (and of GSL's dtrsm, cblas/source_trsm_r.h)
#pragma unroll
  for (int r = 1; r < NB; r++)
#pragma unroll
    for (int kk = 0; kk < r; kk++)
      u[r] = u[r] - P[r * PP + kk] * u[kk];
#pragma unroll
  for (int kk = 0; kk < NB; kk++)
    if (kk < nbk)
      W[(size_t) (k0 + kk) * n + c] = u[kk];
  for (int i = k1; i < n; i++)
  {
    const double *Li = P + (i - k0) * PP;
    double a = W[(size_t) i * n + c];
#pragma unroll
    for (int kk = 0; kk < NB; kk++)
      a = a - Li[kk] * u[kk];
    W[(size_t) i * n + c] = a;
  }
}

// ============================================================================================
// k_blocked_f64: right-looking blocked LU with an NB-column panel in LDS
// ============================================================================================
template <int BS, int NB, int MAXM, bool KEEP>
__global__ __launch_bounds__ (BS, 4) void k_blocked_f64 (SolveArgs s, double *work, int64_t wstride)
{
  extern __shared__ double P[];                  // panel rows x (NB+1); column NB holds b
  constexpr int PP = NB + 1;
  constexpr int NW = BS / 64;
  __shared__ long long r_key[NW];
  __shared__ int r_pos[NW];
  __shared__ int s_ipiv[NB];
  __shared__ int s_fail;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  const int neff = list_len (s);
  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    double *W = (KEEP && s.LU) ? (s.LU + s.offA[cell]) : (work + (size_t) blockIdx.x * wstride);
    double *wb = (KEEP && s.LU) ? (work + (size_t) blockIdx.x * wstride) : (W + (size_t) n * n);
    double *xs = P + (size_t) n * PP;

    for (int64_t e = tid; e < (int64_t) n * n; e += BS)
      W[e] = A[e];
    for (int i = tid; i < n; i += BS)
      wb[i] = b[i];
    if (tid == 0)
      s_fail = 0x7fffffff;
    __syncthreads ();

    int singular = 0;
    for (int k0 = 0; k0 < n; k0 += NB)
    {
      const int nbk = min (NB, n - k0);
      const int rows = n - k0;
      for (int r = tid; r < rows; r += BS)
      {
        const double *src = W + (size_t) (k0 + r) * n + k0;
        double *dst = P + r * PP;
        for (int c = 0; c < nbk; c++)
          dst[c] = src[c];
        dst[NB] = wb[k0 + r];
      }
      __syncthreads ();

      // ---- panel factorization ----
      for (int kk = 0; kk < nbk; kk++)
      {
        long long key = -1;
        int pos = 0x7fffffff;
        for (int r = kk + tid; r < rows; r += BS)
        {
          const long long q = pivot_key (P[r * PP + kk]);
          if (q > key)
          {
            key = q;
            pos = r;
          }
        }
        const int wl = wave_pivot_lane (key >= 0, key, pos);
        if (lane == 0)
        {
          r_key[wv] = wl >= 0 ? (long long) rl_u64 ((unsigned long long) key, wl) : -1;
          r_pos[wv] = wl >= 0 ? __builtin_amdgcn_readlane (pos, wl) : 0x7fffffff;
        }
        __syncthreads ();
        long long bk = r_key[0];
        int p = r_pos[0];
#pragma unroll
        for (int w = 1; w < NW; w++)
          if (r_key[w] > bk || (r_key[w] == bk && r_pos[w] < p))
          {
            bk = r_key[w];
            p = r_pos[w];
          }
        if (tid == 0)
          s_ipiv[kk] = p;
        if (p != kk && tid < PP)
        {
          const double t = P[kk * PP + tid];
          P[kk * PP + tid] = P[p * PP + tid];
          P[p * PP + tid] = t;
        }
        __syncthreads ();
        const double piv = P[kk * PP + kk];
        singular |= (piv == 0.0);
        const bool recip = fabs (piv) >= DBL_MIN;
        const double sc = 1.0 / piv;
        for (int r = kk + 1 + tid; r < rows; r += BS)
        {
          double *Pr = P + r * PP;
          const double l = recip ? Pr[kk] * sc : Pr[kk] / piv;
          Pr[kk] = l;
          for (int c = kk + 1; c < nbk; c++)
            Pr[c] = Pr[c] - l * P[kk * PP + c];
          Pr[NB] = Pr[NB] - l * P[kk * PP + NB];
        }
        __syncthreads ();
      }

      // ---- write the panel back, pivots out ----
      for (int r = tid; r < rows; r += BS)
      {
        double *dst = W + (size_t) (k0 + r) * n + k0;
        const double *src = P + r * PP;
        for (int c = 0; c < nbk; c++)
          dst[c] = src[c];
        wb[k0 + r] = src[NB];
      }
      if (KEEP && s.piv && tid < nbk)
        s.piv[s.offV[cell] + k0 + tid] = k0 + s_ipiv[tid];

      // ---- per column outside the panel: row swaps, then U12 (trsm) and A22 (gemm) ----
      const int k1 = k0 + nbk;
      for (int c = tid; c < n; c += BS)
      {
        if (c >= k0 && c < k1)
          continue;
        for (int kk = 0; kk < nbk; kk++)
        {
          const int p = s_ipiv[kk];
          if (p != kk)
          {
            double *x0 = W + (size_t) (k0 + kk) * n + c;
            double *x1 = W + (size_t) (k0 + p) * n + c;
            const double t = *x0;
            *x0 = *x1;
            *x1 = t;
          }
        }
        if (c < k0)
          continue;
        blocked_column<NB> (W, P, n, k0, nbk, c);
      }
      __syncthreads ();
    }

    if (singular)
    {
      if (tid == 0)
        s.codes[cell] = 4;
      __syncthreads ();
      continue;
    }

    if (wv == 0)
    {
      double xr[MAXM];
      wave_backsub<MAXM> (W, (size_t) n, wb, 1, n, lane, xr, xs);
    }
    __syncthreads ();

    for (int i = tid; i < n; i += BS)
    {
      const double *Ai = A + (size_t) i * n;
      double t = 0.0;
#pragma unroll 8
      for (int j = 0; j < n; j++)
      {
        const double pr = xs[j] * Ai[j];
        t = t + pr;
      }
      const int f = residual_row_code (0.0 + t, b[i]);
      if (f)
        atomicMin (&s_fail, (i << 2) | f);
      s.X[s.offV[cell] + i] = xs[i];
    }
    __syncthreads ();
    if (tid == 0)
      s.codes[cell] = (s_fail == 0x7fffffff) ? 0 : (s_fail & 3);
    __syncthreads ();
  }
}

// ============================================================================================
// launchers
// ============================================================================================
size_t wave_f64_smem (int nc, int wpb, bool keep)
{
  return (size_t) wpb * ((size_t) nc * (nc + 2) + (keep ? (size_t) nc * (nc + 1) : 0)) * sizeof (double);
}

size_t lds_f64_smem (int n)
{
  const int ld = (n + 1) | 1;
  return ((size_t) n * ld + (size_t) n) * sizeof (double);
}

int lds_f64_limit ()
{
  static int cached = 0;
  if (!cached)
  {
    const size_t cap = hipla_device_max_smem () - 2048;
    int n = 1;
    while (lds_f64_smem (n + 1) <= cap)
      n++;
    cached = n;
  }
  return cached;
}

size_t blocked_f64_smem (int nmax)
{
  return ((size_t) nmax * (BLK_NB + 1) + (size_t) nmax) * sizeof (double);
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
    return launch_wave_nc<32, 4> (a, keep, grid, st);
  case 64:
    return launch_wave_nc<64, 2> (a, keep, grid, st);
  default:
    return hipErrorInvalidValue;
  }
}

hipError_t launch_lds_f64 (const SolveArgs &a, int nmax, bool keep, int grid, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  const size_t sm = lds_f64_smem (nmax);
  grid = (grid > 0 && grid < a.nlist) ? grid : a.nlist;
  if (keep)
    hipLaunchKernelGGL ((k_lds_f64<256, true>), dim3 (grid), dim3 (256), sm, st, a);
  else
    hipLaunchKernelGGL ((k_lds_f64<256, false>), dim3 (grid), dim3 (256), sm, st, a);
  return hipGetLastError ();
}

template <int MAXM>
static void launch_blocked_m (const SolveArgs &a, int grid, size_t sm, double *work, int64_t wstride, bool keep,
                              hipStream_t st)
{
  if (keep)
    hipLaunchKernelGGL ((k_blocked_f64<256, BLK_NB, MAXM, true>), dim3 (grid), dim3 (256), sm, st, a, work, wstride);
  else
    hipLaunchKernelGGL ((k_blocked_f64<256, BLK_NB, MAXM, false>), dim3 (grid), dim3 (256), sm, st, a, work, wstride);
}

hipError_t launch_blocked_f64 (const SolveArgs &a, int nmax, int grid, double *work, int64_t wstride, bool keep,
                               hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  if (nmax > 64 * BLK_MAXM)
    return hipErrorInvalidValue;
  const size_t sm = blocked_f64_smem (nmax);
  if (nmax <= 256)
    launch_blocked_m<4> (a, grid, sm, work, wstride, keep, st);
  else
    launch_blocked_m<BLK_MAXM> (a, grid, sm, work, wstride, keep, st);
  return hipGetLastError ();
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
  set_smem_attr (k_lds_f64<256, false>, max_smem);
  set_smem_attr (k_lds_f64<256, true>, max_smem);
  set_smem_attr (k_blocked_f64<256, BLK_NB, 4, false>, max_smem);
  set_smem_attr (k_blocked_f64<256, BLK_NB, 4, true>, max_smem);
  set_smem_attr (k_blocked_f64<256, BLK_NB, BLK_MAXM, false>, max_smem);
  set_smem_attr (k_blocked_f64<256, BLK_NB, BLK_MAXM, true>, max_smem);
  return hipSuccess;
}
