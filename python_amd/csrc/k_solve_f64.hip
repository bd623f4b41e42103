// k_solve_f64.hip -- FP64 partial-pivot LU + solve in GSL's operation order (gfx950).
//
// The reference solves A x = b with gsl_linalg_LU_decomp + gsl_linalg_LU_solve and checks the
// residual (source/matrix_cpu.c:96-204).  GSL's recursive LU (GSL 2.6 linalg/lu.c:204-288)
// gives every element a(pos, j) its updates a - l*u for k = 0, 1, ... in ascending order with
// the product rounded first, so a right-looking elimination with the same pivot choice
// (first strict max of |a| by current row position), the same multiplier rule (reciprocal
// multiply when |pivot| >= DBL_MIN, linalg/lu.c:162-176) and no FMA contraction reproduces its
// factors; the forward solve is folded into the elimination as an extra column and the back
// substitution keeps dtrsv's ascending-j row dot (cblas/source_trsv_r.h:35-52).  This file is
// compiled with -ffp-contract=off.
//
// Three shapes, chosen per cell by the host plan:
//   k_wave_f64<NC>     n <= NC <= 64: one wave per matrix, lane = row, row in registers,
//                      U rows stashed in LDS as they are produced (no row swaps: positions
//                      are tracked per lane).
//   k_lds_f64          n <= ~141: one workgroup per matrix, [A|b] resident in LDS.
//   k_blocked_f64<NB>  larger n: right-looking blocked LU, NB-column panel factored in LDS,
//                      row swaps, U12 and the A22 update streamed through a per-workgroup
//                      HBM workspace (L2 resident), persistent over the cell list.
#include "hipla_dev.h"
#include "hipla_kernels.h"

// ============================================================================================
// k_wave_f64: one wave per matrix
// ============================================================================================
template <int NC, int WPB, bool KEEP>
__global__ __launch_bounds__ (64 * WPB) void k_wave_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  constexpr int LDU = NC + 1;                        // U row k: [t] = U(k, k+t), [NC] = y_k
  constexpr int PER_WAVE = NC * LDU + 64 + (KEEP ? NC * NC : 0);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int item = blockIdx.x * WPB + wv;
  if (item >= s.nlist)
    return;
  const int cell = s.list[item];
  const int n = s.n[cell];
  double *U = smem + wv * PER_WAVE;
  double *xs = U + NC * LDU;
  double *Ls = xs + 64;                               // KEEP: Ls[row * NC + k] = multiplier
  const double *A = s.A + s.offA[cell];
  const double *b = s.B + s.offV[cell];
  if (n <= 0 || n > NC)
  {
    if (lane == 0)
      s.codes[cell] = 5;
    return;
  }

  // Coalesced load of A through LDS (odd pitch: at most 2-way bank conflicts on the row reads).
  const int pitch = n | 1;
  for (int r = 0; r < n; r++)
    for (int c = lane; c < n; c += 64)
      U[r * pitch + c] = A[(size_t) r * n + c];
  wave_lds_fence ();
  const bool inrange = lane < n;
  double a[NC];
#pragma unroll
  for (int j = 0; j < NC; j++)
    a[j] = (inrange && j < n) ? U[lane * pitch + j] : 0.0;
  double bb = inrange ? b[lane] : 0.0;
  wave_lds_fence ();

  bool alive = inrange;
  int pos = lane;
  int singular = 0;
  for (int k = 0; k < n; k++)
  {
    long long key = alive ? pivot_key (a[0]) : -1LL;
    int kpos = alive ? pos : 0x7fffffff;
    wave_argmax_keypos (key, kpos);
    const int ppos = kpos;
    const unsigned long long bal = __ballot (alive && pos == ppos);
    const int q = __ffsll ((unsigned long long) bal) - 1;
    double *Uk = U + k * LDU;
    if (lane == q)
    {
#pragma unroll
      for (int j = 0; j < NC; j++)
        if (j < n - k)
          Uk[j] = a[j];
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
        Ls[lane * NC + k] = l;
      const double ub = Uk[NC];
      bb = bb - l * ub;
      int lo, hi;
      skip_window (n, k, pos, lo, hi);
      const bool z = (l == 0.0);
#pragma unroll
      for (int j = 1; j < NC; j++)
      {
        if (j < n - k)
        {
          const double u = Uk[j];
          const double t = a[j] - l * u;
          const int col = k + j;
          a[j - 1] = (z && col >= lo && col < hi) ? a[j] : t;
        }
      }
    }
  }
  wave_lds_fence ();

  if (KEEP && s.LU)
  {
    // row `lane` became pivot at step pos: LU row pos = [L multipliers | U row]
    if (inrange)
    {
      double *out = s.LU + s.offA[cell] + (size_t) pos * n;
      for (int j = 0; j < pos; j++)
        out[j] = Ls[lane * NC + j];
      const double *Ur = U + pos * LDU;
      for (int j = pos; j < n; j++)
        out[j] = Ur[j - pos];
    }
  }

  if (singular)
  {
    if (lane == 0)
      s.codes[cell] = 4;
    return;
  }

  // Back substitution in dtrsv order (one lane; the chain of subtractions is sequential).
  if (lane == 0)
  {
    const double *Ul = U + (n - 1) * LDU;
    xs[n - 1] = Ul[NC] / Ul[0];
    for (int i = n - 2; i >= 0; i--)
    {
      const double *Ui = U + i * LDU;
      double tmp = Ui[NC];
      for (int j = i + 1; j < n; j++)
      {
        const double p = Ui[j - i] * xs[j];
        tmp = tmp - p;
      }
      xs[i] = tmp / Ui[0];
    }
  }
  wave_lds_fence ();

  // Residual check t = A x (gsl_blas_dgemv order) and codes 2 / 3 of the first failing row.
  int fail = 0;
  if (inrange)
  {
    const double *Ai = A + (size_t) lane * n;
    double t = 0.0;
    for (int j = 0; j < n; j++)
    {
      const double p = xs[j] * Ai[j];
      t = t + p;
    }
    const double tv = 0.0 + t;
    fail = residual_row_code (tv, b[lane]);
  }
  const unsigned long long fb = __ballot (fail != 0);
  int code = 0;
  if (fb)
    code = __shfl (fail, __ffsll ((unsigned long long) fb) - 1);
  if (inrange)
    s.X[s.offV[cell] + lane] = xs[lane];
  if (lane == 0)
    s.codes[cell] = code;
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
  const int cell = s.list[blockIdx.x];
  const int n = s.n[cell];
  const int ld = (n + 1) | 1;
  double *xs = M + (size_t) n * ld;
  double *prod = xs + n;
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
    wave_argmax_keypos (key, pos);
    if (lane == 0)
    {
      r_key[wv] = key;
      r_pos[wv] = pos;
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
    for (int i = k + 1 + wv; i < n; i += NW)
    {
      const double l = gsl_multiplier (M[i * ld + k], piv);
      int lo, hi;
      skip_window (n, k, i, lo, hi);
      const bool z = (l == 0.0);
      for (int j = k + 1 + lane; j <= n; j += 64)
      {
        const double u = M[k * ld + j];
        const double cur = M[i * ld + j];
        const double t = cur - l * u;
        M[i * ld + j] = (j < n && z && j >= lo && j < hi) ? cur : t;
      }
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
    return;
  }

  // Back substitution, dtrsv order: products in parallel, the subtraction chain in one lane.
  if (wv == 0)
  {
    if (lane == 0)
      xs[n - 1] = M[(n - 1) * ld + n] / M[(n - 1) * ld + (n - 1)];
    wave_lds_fence ();
    for (int i = n - 2; i >= 0; i--)
    {
      for (int j = i + 1 + lane; j < n; j += 64)
        prod[j] = M[i * ld + j] * xs[j];
      wave_lds_fence ();
      if (lane == 0)
      {
        double tmp = M[i * ld + n];
        for (int j = i + 1; j < n; j++)
          tmp = tmp - prod[j];
        xs[i] = tmp / M[i * ld + i];
      }
      wave_lds_fence ();
    }
  }
  __syncthreads ();

  for (int i = tid; i < n; i += BS)
  {
    const double *Ai = A + (size_t) i * n;
    double t = 0.0;
    for (int j = 0; j < n; j++)
    {
      const double p = xs[j] * Ai[j];
      t = t + p;
    }
    const int f = residual_row_code (0.0 + t, b[i]);
    if (f)
      atomicMin (&s_fail, (i << 2) | f);
    s.X[s.offV[cell] + i] = xs[i];
  }
  __syncthreads ();
  if (tid == 0)
    s.codes[cell] = (s_fail == 0x7fffffff) ? 0 : (s_fail & 3);
}

// ============================================================================================
// k_blocked_f64: right-looking blocked LU with an NB-column panel in LDS
// ============================================================================================
template <int BS, int NB, bool KEEP>
__global__ __launch_bounds__ (BS) void k_blocked_f64 (SolveArgs s, double *work, int64_t wstride)
{
  extern __shared__ double P[];                  // panel rows x (NB+1); column NB holds b
  constexpr int PP = NB + 1;
  constexpr int NW = BS / 64;
  __shared__ long long r_key[NW];
  __shared__ int r_pos[NW];
  __shared__ int s_ipiv[NB];
  __shared__ int s_fail;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  for (int item = blockIdx.x; item < s.nlist; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    double *W = (KEEP && s.LU) ? (s.LU + s.offA[cell]) : (work + (size_t) blockIdx.x * wstride);
    double *wb = (KEEP && s.LU) ? (work + (size_t) blockIdx.x * wstride) : (W + (size_t) n * n);
    double *xs = P + (size_t) n * PP;
    double *prod = xs + n;

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
        wave_argmax_keypos (key, pos);
        if (lane == 0)
        {
          r_key[wv] = key;
          r_pos[wv] = pos;
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
        for (int r = kk + 1 + tid; r < rows; r += BS)
        {
          double *Pr = P + r * PP;
          const double l = gsl_multiplier (Pr[kk], piv);
          Pr[kk] = l;
          const bool z = (l == 0.0);
          int lo = n, hi = n;
          if (z)
            skip_window (n, k0 + kk, k0 + r, lo, hi);
          for (int c = kk + 1; c < nbk; c++)
          {
            const double u = P[kk * PP + c];
            const double t = Pr[c] - l * u;
            const int col = k0 + c;
            if (!(z && col >= lo && col < hi))
              Pr[c] = t;
          }
          const double ub = P[kk * PP + NB];
          Pr[NB] = Pr[NB] - l * ub;
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
        double u[NB];
#pragma unroll
        for (int kk = 0; kk < NB; kk++)
          u[kk] = (kk < nbk) ? W[(size_t) (k0 + kk) * n + c] : 0.0;
#pragma unroll
        for (int kk = 0; kk < NB; kk++)
        {
          if (kk < nbk)
          {
#pragma unroll
            for (int r = kk + 1; r < NB; r++)
            {
              if (r < nbk)
              {
                const double l = P[r * PP + kk];
                const double t = u[r] - l * u[kk];
                bool skip = false;
                if (l == 0.0)
                {
                  int lo, hi;
                  skip_window (n, k0 + kk, k0 + r, lo, hi);
                  skip = c >= lo && c < hi;
                }
                if (!skip)
                  u[r] = t;
              }
            }
          }
        }
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
          {
            if (kk < nbk)
            {
              const double l = Li[kk];
              const double t = a - l * u[kk];
              bool skip = false;
              if (l == 0.0)
              {
                int lo, hi;
                skip_window (n, k0 + kk, i, lo, hi);
                skip = c >= lo && c < hi;
              }
              if (!skip)
                a = t;
            }
          }
          W[(size_t) i * n + c] = a;
        }
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

    // ---- back substitution, dtrsv order ----
    if (tid == 0)
      xs[n - 1] = wb[n - 1] / W[(size_t) (n - 1) * n + (n - 1)];
    __syncthreads ();
    for (int i = n - 2; i >= 0; i--)
    {
      const double *Ui = W + (size_t) i * n;
      for (int j = i + 1 + tid; j < n; j += BS)
        prod[j] = Ui[j] * xs[j];
      __syncthreads ();
      if (tid == 0)
      {
        double tmp = wb[i];
        for (int j = i + 1; j < n; j++)
          tmp = tmp - prod[j];
        xs[i] = tmp / Ui[i];
      }
      __syncthreads ();
    }

    for (int i = tid; i < n; i += BS)
    {
      const double *Ai = A + (size_t) i * n;
      double t = 0.0;
      for (int j = 0; j < n; j++)
      {
        const double p = xs[j] * Ai[j];
        t = t + p;
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
  return (size_t) wpb * ((size_t) nc * (nc + 1) + 64 + (keep ? (size_t) nc * nc : 0)) * sizeof (double);
}

size_t lds_f64_smem (int n)
{
  const int ld = (n + 1) | 1;
  return ((size_t) n * ld + 2 * (size_t) n) * sizeof (double);
}

size_t blocked_f64_smem (int nmax)
{
  return ((size_t) nmax * (BLK_NB + 1) + 2 * (size_t) nmax) * sizeof (double);
}

template <int NC, int WPB>
static hipError_t launch_wave_nc (const SolveArgs &a, bool keep, hipStream_t st)
{
  const int grid = (a.nlist + WPB - 1) / WPB;
  const size_t sm = wave_f64_smem (NC, WPB, keep);
  if (keep)
    hipLaunchKernelGGL ((k_wave_f64<NC, WPB, true>), dim3 (grid), dim3 (64 * WPB), sm, st, a);
  else
    hipLaunchKernelGGL ((k_wave_f64<NC, WPB, false>), dim3 (grid), dim3 (64 * WPB), sm, st, a);
  return hipGetLastError ();
}

hipError_t launch_wave_f64 (const SolveArgs &a, int nc, bool keep, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  switch (nc)
  {
  case 16:
    return launch_wave_nc<16, 4> (a, keep, st);
  case 32:
    return launch_wave_nc<32, 4> (a, keep, st);
  case 64:
    return launch_wave_nc<64, 2> (a, keep, st);
  default:
    return hipErrorInvalidValue;
  }
}

hipError_t launch_lds_f64 (const SolveArgs &a, int nmax, bool keep, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  const size_t sm = lds_f64_smem (nmax);
  if (keep)
    hipLaunchKernelGGL ((k_lds_f64<256, true>), dim3 (a.nlist), dim3 (256), sm, st, a);
  else
    hipLaunchKernelGGL ((k_lds_f64<256, false>), dim3 (a.nlist), dim3 (256), sm, st, a);
  return hipGetLastError ();
}

hipError_t launch_blocked_f64 (const SolveArgs &a, int nmax, int grid, double *work, int64_t wstride, bool keep,
                               hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  const size_t sm = blocked_f64_smem (nmax);
  if (keep)
    hipLaunchKernelGGL ((k_blocked_f64<256, BLK_NB, true>), dim3 (grid), dim3 (256), sm, st, a, work, wstride);
  else
    hipLaunchKernelGGL ((k_blocked_f64<256, BLK_NB, false>), dim3 (grid), dim3 (256), sm, st, a, work, wstride);
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
  set_smem_attr (k_blocked_f64<256, BLK_NB, false>, max_smem);
  set_smem_attr (k_blocked_f64<256, BLK_NB, true>, max_smem);
  return hipSuccess;
}
