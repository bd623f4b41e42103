// k_invert.hip -- batched inverse (invert_matrix) and the adjoint solve A^T lambda = g (gfx950).
//
// Inverse.  The reference inverts N = I - Q once per plasma cell with gsl_linalg_LU_decomp +
// gsl_linalg_LU_invert (source/matrix_cpu.c:223-257 -> GSL 2.6 linalg/lu.c:414-481), i.e.
// A^-1 = U^-1 L^-1 P.  Here the LU is the same GSL-order elimination as the solve kernels, run
// on the augmented [A | I]: the row swaps, U12 and trailing updates carry the n identity columns
// along, which leaves L^-1 P in them; a blocked backward sweep then solves U X = L^-1 P for all n
// columns.  The factors (and so the singular verdict, an exact zero pivot -> GSL_EDOM) are GSL's;
// the inverse itself uses fused multiply-adds in its own order (tolerance-based parity: 1e-8
// relative Frobenius against the FP64 oracle; the reference fixtures at 1e-6).
//
// Adjoint.  With P A = L U stored by a solve (HIPLA_FLAG_KEEP_FACTORS), A^T lambda = g is
// U^T y = g (forward), L^T z = y (backward, unit diagonal), lambda = P^T z, and the gradient of a
// loss L(x) through x = A^-1 b is dL/dA = -lambda x^T, dL/db = lambda.  One wave per cell, the
// rows of U and L read coalesced, the running vector in registers (lane j holds entry j + 64 m).
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

#include <algorithm>

// ============================================================================================
// inverse
// ============================================================================================
template <int NB>
__device__ __forceinline__ double &aug_at (double *W, double *V, int n, int i, int c)
{
  return c < n ? W[(size_t) i * n + c] : V[(size_t) i * n + (c - n)];
}

// Forward work of one column c (of [A | I], outside the panel) for panel [k0, k0+nbk): the
// panel's row swaps, U12 = L11^-1 A12 and the trailing update, all in the GSL order.
template <int NB>
__device__ __forceinline__ void inv_forward_column (double *W, double *V, const double *P, const int *ipiv, int n,
                                                    int k0, int nbk, int c)
{
  constexpr int PP = NB + 1;
  for (int kk = 0; kk < nbk; kk++)
  {
    const int p = ipiv[kk];
    if (p != kk)
    {
      double &x0 = aug_at<NB> (W, V, n, k0 + kk, c);
      double &x1 = aug_at<NB> (W, V, n, k0 + p, c);
      const double t = x0;
      x0 = x1;
      x1 = t;
    }
  }
  if (c < k0)
    return;
  const int k1 = k0 + nbk;
  double u[NB];
#pragma unroll
  for (int kk = 0; kk < NB; kk++)
    u[kk] = (kk < nbk) ? aug_at<NB> (W, V, n, k0 + kk, c) : 0.0;
#pragma unroll
  for (int r = 1; r < NB; r++)
#pragma unroll
    for (int kk = 0; kk < r; kk++)
      u[r] = u[r] - P[r * PP + kk] * u[kk];
#pragma unroll
  for (int kk = 0; kk < NB; kk++)
    if (kk < nbk)
      aug_at<NB> (W, V, n, k0 + kk, c) = u[kk];
  double *col = c < n ? W + c : V + (c - n);
  int i = k1;
  for (; i + 3 < n; i += 4)
  {
    double a0 = col[(size_t) i * n], a1 = col[(size_t) (i + 1) * n], a2 = col[(size_t) (i + 2) * n],
           a3 = col[(size_t) (i + 3) * n];
    const double *L0 = P + (i - k0) * PP;
#pragma unroll
    for (int kk = 0; kk < NB; kk++)
    {
      a0 = a0 - L0[kk] * u[kk];
      a1 = a1 - L0[PP + kk] * u[kk];
      a2 = a2 - L0[2 * PP + kk] * u[kk];
      a3 = a3 - L0[3 * PP + kk] * u[kk];
    }
    col[(size_t) i * n] = a0;
    col[(size_t) (i + 1) * n] = a1;
    col[(size_t) (i + 2) * n] = a2;
    col[(size_t) (i + 3) * n] = a3;
  }
  for (; i < n; i++)
  {
    const double *Li = P + (i - k0) * PP;
    double a = col[(size_t) i * n];
#pragma unroll
    for (int kk = 0; kk < NB; kk++)
      a = a - Li[kk] * u[kk];
    col[(size_t) i * n] = a;
  }
}

template <int BS, int NB>
__global__ __launch_bounds__ (BS, 4) void k_invert_f64 (SolveArgs s, double *work, int64_t wstride, double *Inv)
{
  extern __shared__ double P[];                  // panel rows x (NB+1) / diagonal block of U
  constexpr int PP = NB + 1;
  constexpr int NW = BS / 64;
  __shared__ long long r_key[NW];
  __shared__ int r_pos[NW];
  __shared__ int s_ipiv[NB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  const int neff = list_len (s);
  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    const double *A = s.A + s.offA[cell];
    double *W = work + (size_t) blockIdx.x * wstride;
    double *V = Inv + s.offA[cell];              // right-hand sides I, then the inverse

    for (int r = wv; r < n; r += NW)
      for (int c = lane; c < n; c += 64)
      {
        W[(size_t) r * n + c] = A[(size_t) r * n + c];
        V[(size_t) r * n + c] = (r == c) ? 1.0 : 0.0;
      }
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
      }
      __syncthreads ();
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
        if (p != kk && tid < NB)
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
        }
        __syncthreads ();
      }
      for (int r = tid; r < rows; r += BS)
      {
        double *dst = W + (size_t) (k0 + r) * n + k0;
        const double *src = P + r * PP;
        for (int c = 0; c < nbk; c++)
          dst[c] = src[c];
      }
      const int k1 = k0 + nbk;
      for (int c = tid; c < 2 * n; c += BS)
      {
        if (c >= k0 && c < k1)
          continue;
        __asm__ volatile ("" ::: "memory");
        inv_forward_column<NB> (W, V, P, s_ipiv, n, k0, nbk, c);
      }
      __syncthreads ();
    }

    if (singular)
    {
      if (tid == 0)
        s.codes[cell] = 1;   // GSL_EDOM, as gsl_linalg_LU_invx on an exact zero pivot
      __syncthreads ();
      continue;
    }

    // backward sweep U X = Y for all n right-hand side columns, NB rows at a time from the bottom
    for (int k1 = n; k1 > 0; k1 -= NB)
    {
      const int k0 = max (0, k1 - NB);
      const int nbk = k1 - k0;
      // diagonal block of U into LDS
      for (int e = tid; e < nbk * nbk; e += BS)
      {
        const int r = e / nbk, c = e - r * nbk;
        P[r * PP + c] = W[(size_t) (k0 + r) * n + k0 + c];
      }
      __syncthreads ();
      for (int c = tid; c < n; c += BS)
      {
        double x[NB];
#pragma unroll
        for (int r = 0; r < NB; r++)
          x[r] = (r < nbk) ? V[(size_t) (k0 + r) * n + c] : 0.0;
#pragma unroll
        for (int r = NB - 1; r >= 0; r--)
        {
          if (r < nbk)
          {
            double t = x[r];
#pragma unroll
            for (int j = r + 1; j < NB; j++)
              if (j < nbk)
                t = fma (-P[r * PP + j], x[j], t);
            x[r] = t / P[r * PP + r];
          }
        }
#pragma unroll
        for (int r = 0; r < NB; r++)
          if (r < nbk)
            V[(size_t) (k0 + r) * n + c] = x[r];
        int i = 0;
        for (; i + 3 < k0; i += 4)
        {
          double a0 = V[(size_t) i * n + c], a1 = V[(size_t) (i + 1) * n + c], a2 = V[(size_t) (i + 2) * n + c],
                 a3 = V[(size_t) (i + 3) * n + c];
          const double *U0 = W + (size_t) i * n + k0;
#pragma unroll
          for (int r = 0; r < NB; r++)
            if (r < nbk)
            {
              a0 = fma (-U0[r], x[r], a0);
              a1 = fma (-U0[n + r], x[r], a1);
              a2 = fma (-U0[2 * n + r], x[r], a2);
              a3 = fma (-U0[3 * n + r], x[r], a3);
            }
          V[(size_t) i * n + c] = a0;
          V[(size_t) (i + 1) * n + c] = a1;
          V[(size_t) (i + 2) * n + c] = a2;
          V[(size_t) (i + 3) * n + c] = a3;
        }
        for (; i < k0; i++)
        {
          const double *Ui = W + (size_t) i * n + k0;
          double a = V[(size_t) i * n + c];
#pragma unroll
          for (int r = 0; r < NB; r++)
            if (r < nbk)
              a = fma (-Ui[r], x[r], a);
          V[(size_t) i * n + c] = a;
        }
      }
      __syncthreads ();
    }
    if (tid == 0)
      s.codes[cell] = 0;
    __syncthreads ();
  }
}

size_t invert_f64_smem (int nmax)
{
  return (size_t) nmax * (INV_NB + 1) * sizeof (double);
}

hipError_t launch_invert_f64 (const SolveArgs &a, int nmax, int grid, double *work, int64_t wstride, double *Inv,
                              hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  hipLaunchKernelGGL ((k_invert_f64<256, INV_NB>), dim3 (grid), dim3 (256), invert_f64_smem (nmax), st, a, work,
                      wstride, Inv);
  return hipGetLastError ();
}

// Route the cells once: orders the MFMA Gauss-Jordan sweep takes (multiples of 64 from 128 on)
// in size classes, every other cell to the GSL-order pivoting kernel; the workspace of that
// kernel is sized for the largest order it may receive (cells of the sweep classes it is handed
// back included).
int plan_invert_prepare (hipla_plan p)
{
  const DeviceInfo &dev = hipla_device_info ();
  if (!p->inv_ready)
  {
    std::vector<int> order (p->ncells), rest;
    for (int c = 0; c < p->ncells; c++)
      order[c] = c;
    std::stable_sort (order.begin (), order.end (), [&] (int a, int b) { return p->n[a] > p->n[b]; });
    std::vector<int> gl;
    for (int i = 0; i < p->ncells;)
    {
      int j = i;
      while (j < p->ncells && p->n[order[j]] == p->n[order[i]])
        j++;
      const int nn = p->n[order[i]];
      if (gj_eligible (nn))
      {
        p->gj_n.push_back (nn);
        p->gj_first.push_back ((int) gl.size ());
        p->gj_count.push_back (j - i);
        for (int k = i; k < j; k++)
          gl.push_back (order[k]);
      }
      else
        for (int k = i; k < j; k++)
          rest.push_back (order[k]);
      i = j;
    }
    p->inv_nrest = (int) rest.size ();
    for (int c : rest)
      p->inv_rest_nmax = std::max (p->inv_rest_nmax, p->n[c]);
    if (!gl.empty ())
    {
      HIPLA_CHECK (hipla_dev_alloc (&p->d_gj_list, sizeof (int) * gl.size ()));
      HIPLA_CHECK (hipMemcpy (p->d_gj_list, gl.data (), sizeof (int) * gl.size (), hipMemcpyHostToDevice));
      HIPLA_CHECK (hipla_dev_alloc (&p->d_gj_flag, sizeof (int) * p->ncells));
      HIPLA_CHECK (hipla_dev_alloc (&p->d_gj_fb, sizeof (int) * (p->ncells + 1)));
    }
    if (!gl.empty () && p->mode == HIPLA_MODE_MIXED)
    {
      HIPLA_CHECK (hipla_dev_alloc (&p->d_gx_Mf, sizeof (float) * p->totA));
      HIPLA_CHECK (hipla_dev_alloc (&p->d_gx_Rf, sizeof (float) * p->totA));
    }
    if (!rest.empty ())
    {
      HIPLA_CHECK (hipla_dev_alloc (&p->d_inv_rest, sizeof (int) * rest.size ()));
      HIPLA_CHECK (hipMemcpy (p->d_inv_rest, rest.data (), sizeof (int) * rest.size (), hipMemcpyHostToDevice));
    }
    p->inv_ready = true;
  }
  // the pivoting kernel's largest order: every cell when some sweep class exists (any of its
  // cells may be handed back), else the largest of the rest
  const int nmax = p->gj_n.empty () ? p->inv_rest_nmax : *std::max_element (p->n.begin (), p->n.end ());
  if (nmax <= 0)
    return 0;
  if (invert_f64_smem (nmax) > dev.max_smem - 1024)
  {
    hipla_set_detail ("hipla_invert: n = %d exceeds the inverse kernel's panel capacity", nmax);
    return 5;
  }
  const int64_t need = (int64_t) nmax * nmax;
  const int grid = std::min (p->ncells, 2 * dev.num_cu);
  if (!p->d_inv_work || p->inv_wstride < need || p->inv_slots < grid)
  {
    if (p->d_inv_work)
      HIPLA_CHECK (hipla_dev_free (p->d_inv_work));
    p->d_inv_work = nullptr;
    p->inv_wstride = (need + 63) & ~(int64_t) 63;
    p->inv_slots = grid;
    HIPLA_CHECK (hipla_dev_alloc (&p->d_inv_work, sizeof (double) * p->inv_wstride * grid));
  }
  return 0;
}

int plan_enqueue_invert (hipla_plan p, const double *dA, double *dInv, int *dcodes, hipla_cell_stats *dstats,
                         hipStream_t st)
{
  (void) dstats;
  if (!p->inv_ready || (!p->d_inv_work && (p->inv_nrest > 0 || !p->gj_n.empty ())))
  {
    hipla_set_detail ("hipla_invert: workspace not prepared");
    return 5;
  }
  const int nmax_all = *std::max_element (p->n.begin (), p->n.end ());
  SolveArgs a;
  a.n = p->d_n;
  a.offA = p->d_offA;
  a.offV = p->d_offV;
  a.A = dA;
  a.B = nullptr;
  a.X = nullptr;
  a.codes = dcodes;
  a.LU = nullptr;
  a.piv = nullptr;
  a.dcount = nullptr;
  a.F = p->d_F;
  a.Y = p->d_Y;
  if (!p->gj_n.empty ())
  {
    HIPLA_CHECK (gj_reset (p->d_gj_fb, st));
    for (size_t k = 0; k < p->gj_n.size () && p->mode == HIPLA_MODE_MIXED; k++)
    {
      GjxArgs g;
      g.offA = p->d_offA;
      g.list = p->d_gj_list + p->gj_first[k];
      g.ncells = p->gj_count[k];
      g.nsz = p->gj_n[k];
      g.sb = 0;
      g.A = dA;
      g.Mf = p->d_gx_Mf;
      g.Rf = p->d_gx_Rf;
      g.X = dInv;
      g.codes = dcodes;
      g.flag = p->d_gj_flag;
      g.fb = p->d_gj_fb;
      HIPLA_CHECK (gjx_enqueue_class (g, GX_REFINE_STEPS, st));
    }
    // FP64 sweep: the class with the most work on the caller's stream, the others one after
    // another on the fork stream beside it (their launches hold fewer tiles than the chip has
    // slots); the caller's stream waits for them before the hand-back
    size_t kbig = 0;
    for (size_t k = 1; k < p->gj_n.size (); k++)
      if ((double) p->gj_n[k] * p->gj_n[k] * p->gj_n[k] * p->gj_count[k] >
          (double) p->gj_n[kbig] * p->gj_n[kbig] * p->gj_n[kbig] * p->gj_count[kbig])
        kbig = k;
    const bool forkc = p->mode != HIPLA_MODE_MIXED && p->gj_n.size () > 1;
    if (forkc)
    {
      HIPLA_CHECK (hipEventRecord (p->ev_fork, st));
      HIPLA_CHECK (hipStreamWaitEvent (p->fork_st, p->ev_fork, 0));
    }
    for (size_t k = 0; k < p->gj_n.size () && p->mode != HIPLA_MODE_MIXED; k++)
    {
      hipStream_t cs = (forkc && k != kbig) ? p->fork_st : st;
      GjArgs g;
      g.offA = p->d_offA;
      g.list = p->d_gj_list + p->gj_first[k];
      g.ncells = p->gj_count[k];
      g.nsz = p->gj_n[k];
      g.sb = 0;
      g.A = dA;
      g.M = dInv;
      g.codes = dcodes;
      g.flag = p->d_gj_flag;
      g.fb = p->d_gj_fb;
      HIPLA_CHECK (gj_enqueue_class (g, cs));
    }
    if (forkc)
    {
      HIPLA_CHECK (hipEventRecord (p->ev_join, p->fork_st));
      HIPLA_CHECK (hipStreamWaitEvent (st, p->ev_join, 0));
    }
  }
  if (p->inv_nrest > 0)
  {
    a.list = p->d_inv_rest;   // largest first
    a.nlist = p->inv_nrest;
    HIPLA_CHECK (launch_invert_f64 (a, p->inv_rest_nmax, p->inv_slots, p->d_inv_work, p->inv_wstride, dInv, st));
  }
  if (!p->gj_n.empty ())
  {
    // cells the sweep handed back (not diagonally dominant): a persistent grid over the list
    a.list = p->d_gj_fb + 1;
    a.nlist = p->ncells;
    a.dcount = p->d_gj_fb;
    HIPLA_CHECK (launch_invert_f64 (a, nmax_all, p->inv_slots, p->d_inv_work, p->inv_wstride, dInv, st));
  }
  return 0;
}

hipError_t setup_invert_attrs (size_t)
{
  return hipSuccess;
}

// ============================================================================================
// adjoint: A^T lambda = g with stored factors, dL/dA = -lambda x^T
// ============================================================================================
template <int MAXM>
__global__ __launch_bounds__ (64) void k_adjoint_f64 (const int *nv, const int64_t *offA, const int64_t *offV,
                                                      const int *list, int nlist, const double *LU, const int *ipiv,
                                                      const double *X, const double *G, double *Lam, double *GradA)
{
  const int lane = threadIdx.x;
  for (int item = blockIdx.x; item < nlist; item += gridDim.x)
  {
    const int cell = list[item];
    const int n = nv[cell];
    if (n > 64 * MAXM)
      continue;
    const double *F = LU + offA[cell];
    const int *pv = ipiv + offV[cell];
    const double *g = G + offV[cell];
    // a singular cell (an exact zero on U's diagonal, solve code 4) has no adjoint: lambda and
    // dL/dA are written as zeros instead of the infinities a division by the zero pivot gives
    bool zpiv = false;
    for (int i = lane; i < n; i += 64)
      zpiv |= F[(size_t) i * n + i] == 0.0;
    if (__ballot (zpiv))
    {
      for (int i = lane; i < n; i += 64)
        Lam[offV[cell] + i] = 0.0;
      if (GradA)
        for (int64_t e = lane; e < (int64_t) n * n; e += 64)
          GradA[offA[cell] + e] = 0.0;
      continue;
    }
    double v[MAXM];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      v[m] = (lane + 64 * m < n) ? g[lane + 64 * m] : 0.0;
    // U^T y = g: y_i = v_i / U_ii, then v_j -= U_ij y_i for j > i (row i of U)
    for (int i = 0; i < n; i++)
    {
      const double *Ui = F + (size_t) i * n;
      double vi = 0.0;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
        if (i >> 6 == m)
          vi = v[m];
      const double yi = rl_f64 (vi, i & 63) / Ui[i];
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int j = lane + 64 * m;
        if (j == i)
          v[m] = yi;
        else if (j > i && j < n)
          v[m] = fma (-Ui[j], yi, v[m]);
      }
    }
    // L^T z = y (unit): z_i = v_i, then v_j -= L_ij z_i for j < i (row i of L), i descending
    for (int i = n - 1; i >= 0; i--)
    {
      const double *Li = F + (size_t) i * n;
      double vi = 0.0;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
        if (i >> 6 == m)
          vi = v[m];
      const double zi = rl_f64 (vi, i & 63);
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        const int j = lane + 64 * m;
        if (j < i)
          v[m] = fma (-Li[j], zi, v[m]);
      }
    }
    // lambda = P^T z: undo the row swaps in reverse order (ipiv[k]: row swapped with k at step k),
    // exchanging entries across lanes with readlane
    for (int k = n - 1; k >= 0; k--)
    {
      const int p = pv[k];
      if (p == k)
        continue;
      double vk = 0.0, vp = 0.0;
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        if (k >> 6 == m)
          vk = v[m];
        if (p >> 6 == m)
          vp = v[m];
      }
      const double ak = rl_f64 (vk, k & 63), ap = rl_f64 (vp, p & 63);
#pragma unroll
      for (int m = 0; m < MAXM; m++)
      {
        if (lane + 64 * m == k)
          v[m] = ap;
        if (lane + 64 * m == p)
          v[m] = ak;
      }
    }
    double *lam = Lam + offV[cell];
#pragma unroll
    for (int m = 0; m < MAXM; m++)
      if (lane + 64 * m < n)
        lam[lane + 64 * m] = v[m];
    if (GradA)
    {
      const double *x = X + offV[cell];
      double *gA = GradA + offA[cell];
      for (int i = 0; i < n; i++)
      {
        double vi = 0.0;
#pragma unroll
        for (int m = 0; m < MAXM; m++)
          if (i >> 6 == m)
            vi = v[m];
        const double li = rl_f64 (vi, i & 63);
        for (int j = lane; j < n; j += 64)
          gA[(size_t) i * n + j] = -li * x[j];
      }
    }
  }
}

extern "C" int hipla_adjoint (hipla_plan p, const double *dX, const double *dG, double *dLambda, double *dGradA,
                              void *stream)
{
  if (!p || !dX || !dG || !dLambda)
  {
    hipla_set_detail ("hipla_adjoint: bad argument");
    return 5;
  }
  if (!(p->flags & HIPLA_FLAG_KEEP_FACTORS) || p->mode != HIPLA_MODE_FP64 || !p->solved)
  {
    hipla_set_detail ("hipla_adjoint needs an FP64 plan created with HIPLA_FLAG_KEEP_FACTORS and a prior solve");
    return 5;
  }
  const int nmax = *std::max_element (p->n.begin (), p->n.end ());
  const DeviceInfo &dev = hipla_device_info ();
  const int grid = std::min (p->ncells, 8 * dev.num_cu);
  hipStream_t st = (hipStream_t) stream;
  if (nmax <= 64)
    hipLaunchKernelGGL ((k_adjoint_f64<1>), dim3 (grid), dim3 (64), 0, st, p->d_n, p->d_offA, p->d_offV, p->d_list,
                        p->ncells, p->d_LU, p->d_piv, dX, dG, dLambda, dGradA);
  else if (nmax <= 256)
    hipLaunchKernelGGL ((k_adjoint_f64<4>), dim3 (grid), dim3 (64), 0, st, p->d_n, p->d_offA, p->d_offV, p->d_list,
                        p->ncells, p->d_LU, p->d_piv, dX, dG, dLambda, dGradA);
  else if (nmax <= 1024)
    hipLaunchKernelGGL ((k_adjoint_f64<16>), dim3 (grid), dim3 (64), 0, st, p->d_n, p->d_offA, p->d_offV, p->d_list,
                        p->ncells, p->d_LU, p->d_piv, dX, dG, dLambda, dGradA);
  else
  {
    hipla_set_detail ("hipla_adjoint: n = %d above 1024", nmax);
    return 5;
  }
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}

// ============================================================================================
// macro-atom emissivities without N^-1: out_j = (N^-T a)_j R_jj e_j
// ============================================================================================
// The reference inverts N = I - Q per cell (source/macro_accelerate.c:315), scales its columns by
// the diagonal of R (B = N^-1 R, :324-334) and contracts B with the absorbed energies
// (source/macro_gen_f.c:490-506: emiss_j = (sum_i a_i B_ij) e_j).  Since a^T N^-1 R = (N^-T a)^T R,
// one factorisation and one transposed solve give the same numbers (to rounding) for n^3/3 + 2n^2
// flops instead of the inverse's 4n^3/3.
__global__ __launch_bounds__ (256) void k_macro_emiss (const int *nv, const int64_t *offV, const int *list, int nlist,
                                                      const int *codes, const double *Lam, const double *Rd,
                                                      const double *Em, double *Out)
{
  for (int item = blockIdx.x; item < nlist; item += gridDim.x)
  {
    const int cell = list[item];
    if (codes[cell] == 4)
      continue;
    const int64_t o = offV[cell];
    const int n = nv[cell];
    for (int j = threadIdx.x; j < n; j += blockDim.x)
    {
      const double t = Lam[o + j] * Rd[o + j];
      Out[o + j] = t * Em[o + j];
    }
  }
}

extern "C" int hipla_macro_emiss (hipla_plan p, const double *dN, const double *dRdiag, const double *dAbs,
                                  const double *dEmit, double *dOut, int *dcodes, void *stream)
{
  if (!p || !dN || !dRdiag || !dAbs || !dEmit || !dOut || !dcodes)
  {
    hipla_set_detail ("hipla_macro_emiss: bad argument");
    return 5;
  }
  if (!(p->flags & HIPLA_FLAG_KEEP_FACTORS) || p->mode != HIPLA_MODE_FP64)
  {
    hipla_set_detail ("hipla_macro_emiss needs an FP64 plan created with HIPLA_FLAG_KEEP_FACTORS");
    return 5;
  }
  // factor N (GSL order; the solve's x = N^-1 a lands in dOut and is overwritten below), then
  // N^T lambda = a on the kept factors
  int rc = hipla_solve (p, dN, dAbs, dOut, dcodes, nullptr, stream);
  if (rc)
    return rc;
  rc = hipla_adjoint (p, dOut, dAbs, p->d_lam, nullptr, stream);
  if (rc)
    return rc;
  const DeviceInfo &dev = hipla_device_info ();
  const int grid = std::max (1, std::min (p->ncells, 4 * dev.num_cu));
  hipLaunchKernelGGL (k_macro_emiss, dim3 (grid), dim3 (256), 0, (hipStream_t) stream, p->d_n, p->d_offV, p->d_list,
                      p->ncells, dcodes, p->d_lam, dRdiag, dEmit, dOut);
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}
