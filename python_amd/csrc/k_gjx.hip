// k_gjx.hip -- the mixed-precision inverse of the macro-atom matrices: the blocked Gauss-Jordan
// sweep of k_gj.hip in FP32 with its block products as split BF16x3 MFMA products (FP32
// accumulation), then FP64 residual refinement of the inverse.
//
// Sweep.  The same four steps per 64 x 64 diagonal block as the FP64 sweep, on an FP32 copy Mf of
// M = A^T (the row-major input read column-major).  A block product P = L R of FP32 blocks takes
// L = Lh + Ll, R = Rh + Rl with Lh = bf16(L), Ll = bf16(L - Lh) (likewise R) and forms
// Lh Rh + Lh Rl + Ll Rh on v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the Ll Rl term,
// about 2^-16 of the product, is dropped).  The result X0 ~ M^-1 is accurate to about kappa times
// 1e-5.
// Refinement.  X_0 = X0 in FP64; then twice
//   R_k = I - M X_k                 FP64 block GEMM on v_mfma_f64_16x16x4f64 (K = n), stored FP32
//   X_k+1 = X_k + X0 R_k            BF16x3 block GEMM (K = n), added to X_k in FP64
// Each step multiplies the error by about ||I - X0 M|| (kappa times 1e-5), so two steps bring the
// inverse from about kappa 1e-5 to within rounding of FP64 for kappa up to ~1e3.  The residual is
// the one FP64 product; the correction, whose size is that of R, tolerates BF16x3.  Cells that
// are not row diagonally dominant go to the GSL-order FP64 kernel, as in the FP64 sweep.
#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

#include <algorithm>

#define GX_NB 64
#define GX_KC 32
#define GX_P (GX_KC + 4)        // LDS pitch (floats) of both operand halves: [row or col][k]
#define GX_PD (GX_KC + 2)       // LDS pitch (doubles) of the FP64 GEMM's halves

typedef float gx_f4 __attribute__ ((ext_vector_type (4)));
typedef __bf16 gx_bf8 __attribute__ ((ext_vector_type (8)));
typedef double gx_d4 __attribute__ ((ext_vector_type (4)));

__device__ __forceinline__ int gx_item (int bid, int G)
{
  const int x = bid & 7, i = bid >> 3;
  const int q0 = G >> 3, rem = G & 7;
  return x < rem ? x * (q0 + 1) + i : rem * (q0 + 1) + (x - rem) * q0 + i;
}

// ---- FP32 copy + row dominance test (hand-off list shared with the FP64 sweep's format) ----
__global__ __launch_bounds__ (256) void k_gjx_prepare (GjxArgs g)
{
  const int cell = g.list[blockIdx.x];
  const int n = g.nsz;
  const double *a = g.A + g.offA[cell];
  float *m = g.Mf + g.offA[cell];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int s_bad;
  if (threadIdx.x == 0)
    s_bad = 0;
  __syncthreads ();
  int bad = 0;
  for (int r = wv; r < n; r += 4)
  {
    double off = 0.0, dg = 0.0;
    for (int c = lane; c < n; c += 64)
    {
      const double v = a[(size_t) r * n + c];
      m[(size_t) r * n + c] = (float) v;
      if (c == r)
        dg = fabs (v);
      else
        off += fabs (v);
    }
    for (int o = 32; o > 0; o >>= 1)
    {
      off += __shfl_xor (off, o);
      dg += __shfl_xor (dg, o);
    }
    bad |= !(dg >= off);
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

// ---- FP32 sweep of the diagonal block ----
__global__ __launch_bounds__ (256) void k_gjx_diag (GjxArgs g)
{
  const int cell = g.list[blockIdx.x];
  if (g.flag[cell])
    return;
  const int n = g.nsz, k0 = g.sb * GX_NB;
  float *m = g.Mf + g.offA[cell];
  __shared__ float B[GX_NB][GX_NB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < GX_NB * GX_NB; idx += 256)
    B[idx >> 6][idx & 63] = m[(size_t) (k0 + (idx >> 6)) * n + k0 + (idx & 63)];
  __syncthreads ();
  const int j = tid & 63, r0 = (tid >> 6) * 16;
  for (int k = 0; k < GX_NB; k++)
  {
    const float piv = B[k][k];
    if (piv == 0.0f)
    {
      if (tid == 0)
      {
        g.flag[cell] = 2;
        g.codes[cell] = 1;
      }
      return;
    }
    const float inv = 1.0f / piv;
    const float u = B[j][k];
    float f[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
      f[i] = B[k][r0 + i];
    __syncthreads ();
    if (j == k)
    {
#pragma unroll
      for (int i = 0; i < 16; i++)
        B[k][r0 + i] = (r0 + i == k) ? inv : -f[i] * inv;
    }
    else
    {
      const float uj = u * inv;
#pragma unroll
      for (int i = 0; i < 16; i++)
        B[j][r0 + i] = (r0 + i == k) ? uj : fmaf (-f[i], uj, B[j][r0 + i]);
    }
    __syncthreads ();
  }
  for (int idx = tid; idx < GX_NB * GX_NB; idx += 256)
    m[(size_t) (k0 + (idx >> 6)) * n + k0 + (idx & 63)] = B[idx >> 6][idx & 63];
}

// hi / lo bf16 halves of eight consecutive floats
__device__ __forceinline__ void gx_split (const float *p, gx_bf8 &hi, gx_bf8 &lo)
{
  const float4 x0 = *reinterpret_cast<const float4 *> (p);
  const float4 x1 = *reinterpret_cast<const float4 *> (p + 4);
  const float v[8] = { x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w };
#pragma unroll
  for (int e = 0; e < 8; e++)
  {
    const __bf16 h = (__bf16) v[e];
    hi[e] = h;
    lo[e] = (__bf16) (v[e] - (float) h);
  }
}

// acc (2 x 2 blocks of 16 x 16) += Ls[rows][k] Rs[cols][k] over one staged K half, BF16x3
__device__ __forceinline__ void gx_mma_half (gx_f4 (&acc)[2][2], const float (*Ls)[GX_P], const float (*Rs)[GX_P],
                                             int rw, int cw, int lr, int lk)
{
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
  {
    gx_bf8 ah, al;
    gx_split (&Ls[rw + 16 * bi + lr][8 * lk], ah, al);
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      gx_bf8 bh, bl;
      gx_split (&Rs[cw + 16 * bq + lr][8 * lk], bh, bl);
      acc[bi][bq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16 (al, bh, acc[bi][bq], 0, 0, 0);
      acc[bi][bq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16 (ah, bl, acc[bi][bq], 0, 0, 0);
      acc[bi][bq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16 (ah, bh, acc[bi][bq], 0, 0, 0);
    }
  }
}

// ---- sweep block products (MODE 2, 4, 5 as in k_gj_tile), BF16x3 on FP32 ----
template <int MODE>
__global__ __launch_bounds__ (256, 2) void k_gjx_tile (GjxArgs g)
{
  const int nb = g.nsz / GX_NB;
  const int tpc = MODE == 4 ? (nb - 1) * (nb - 1) : nb - 1;
  const int item = gx_item (blockIdx.x, gridDim.x);
  const int ci = item / tpc, t = item - ci * tpc;
  const int cell = g.list[ci];
  if (g.flag[cell])
    return;
  const int n = g.nsz, sb = g.sb;
  int ti, tj;
  if (MODE == 4)
  {
    ti = t / (nb - 1);
    tj = t - ti * (nb - 1);
    ti += ti >= sb;
    tj += tj >= sb;
  }
  else if (MODE == 2)
  {
    ti = sb;
    tj = t + (t >= sb);
  }
  else
  {
    ti = t + (t >= sb);
    tj = sb;
  }
  const int ai = MODE == 2 ? sb : ti, bj = MODE == 5 ? sb : tj;
  const float sgn = MODE == 2 ? 1.0f : -1.0f;
  float *m = g.Mf + g.offA[cell];
  __shared__ __align__ (16) float Ls[GX_NB][GX_P];   // [row][k] of the left operand, sign applied
  __shared__ __align__ (16) float Rs[GX_NB][GX_P];   // [col][k] of the right operand
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rw = (wv >> 1) * 32, cw = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  gx_f4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      if (MODE == 4)
      {
        const float4 x = *reinterpret_cast<const float4 *> (
            m + (size_t) (tj * GX_NB + cw + 16 * bq + lr) * n + ti * GX_NB + rw + 16 * bi + 4 * lk);
        acc[bi][bq] = gx_f4{ x.x, x.y, x.z, x.w };
      }
      else
        acc[bi][bq] = gx_f4{ 0.f, 0.f, 0.f, 0.f };
    }
  for (int kh = 0; kh < GX_NB; kh += GX_KC)
  {
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int k = idx >> 6, r = idx & 63;          // coalesced along rows of a column
      Ls[r][k] = sgn * m[(size_t) (sb * GX_NB + kh + k) * n + ai * GX_NB + r];
    }
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int c = idx >> 5, k = idx & 31;
      Rs[c][k] = m[(size_t) (bj * GX_NB + c) * n + sb * GX_NB + kh + k];
    }
    __syncthreads ();
    gx_mma_half (acc, Ls, Rs, rw, cw, lr, lk);
    __syncthreads ();
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
      *reinterpret_cast<float4 *> (m + (size_t) (tj * GX_NB + cw + 16 * bq + lr) * n + ti * GX_NB + rw + 16 * bi +
                                   4 * lk) = make_float4 (acc[bi][bq][0], acc[bi][bq][1], acc[bi][bq][2],
                                                          acc[bi][bq][3]);
}

// ---- X = X0 in FP64 for the swept cells ----
__global__ __launch_bounds__ (256) void k_gjx_widen (GjxArgs g)
{
  const int cell = g.list[blockIdx.x];
  if (g.flag[cell])
    return;
  const size_t nn = (size_t) g.nsz * g.nsz;
  const float *src = g.Mf + g.offA[cell];
  double *dst = g.X + g.offA[cell];
  for (size_t i = threadIdx.x; i < nn; i += 256)
    dst[i] = (double) src[i];
}

// ---- R = I - M X (FP64 MFMA, K = n), stored FP32 ----
__global__ __launch_bounds__ (256, 2) void k_gjx_resid (GjxArgs g)
{
  const int nb = g.nsz / GX_NB;
  const int tpc = nb * nb;
  const int item = gx_item (blockIdx.x, gridDim.x);
  const int ci = item / tpc, t = item - ci * tpc;
  const int cell = g.list[ci];
  if (g.flag[cell])
    return;
  const int n = g.nsz, ti = t / nb, tj = t - ti * nb;
  const double *M = g.A + g.offA[cell];       // A row-major = M column-major
  const double *X = g.X + g.offA[cell];
  float *R = g.Rf + g.offA[cell];
  __shared__ double As[GX_KC][GX_NB + 4];     // [k][row] of -M
  __shared__ double Bs[GX_NB][GX_PD];         // [col][k] of X
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rw = (wv >> 1) * 32, cw = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  gx_d4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
#pragma unroll
      for (int e = 0; e < 4; e++)
      {
        // v_mfma_f64_16x16x4f64: element e of lane l at row l/16 + 4 e (the FP32-output BF16
        // instruction keeps it at row 4 (l/16) + e)
        const int r = ti * GX_NB + rw + 16 * bi + lk + 4 * e, c = tj * GX_NB + cw + 16 * bq + lr;
        acc[bi][bq][e] = r == c ? 1.0 : 0.0;
      }
  for (int kb = 0; kb < n; kb += GX_KC)
  {
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int k = idx >> 6, r = idx & 63;
      As[k][r] = -M[(size_t) (kb + k) * n + ti * GX_NB + r];
    }
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int c = idx >> 5, k = idx & 31;
      Bs[c][k] = X[(size_t) (tj * GX_NB + c) * n + kb + k];
    }
    __syncthreads ();
#pragma unroll
    for (int k = 0; k < GX_KC; k += 4)
    {
      double av[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        av[bi] = As[k + lk][rw + 16 * bi + lr];
#pragma unroll
      for (int bq = 0; bq < 2; bq++)
        bv[bq] = Bs[cw + 16 * bq + lr][k + lk];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bq = 0; bq < 2; bq++)
          acc[bi][bq] = __builtin_amdgcn_mfma_f64_16x16x4f64 (av[bi], bv[bq], acc[bi][bq], 0, 0, 0);
    }
    __syncthreads ();
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      float *dst = R + (size_t) (tj * GX_NB + cw + 16 * bq + lr) * n + ti * GX_NB + rw + 16 * bi + lk;
#pragma unroll
      for (int e = 0; e < 4; e++)
        dst[4 * e] = (float) acc[bi][bq][e];
    }
}

// ---- X += X0 R (BF16x3, K = n), the addition in FP64 ----
__global__ __launch_bounds__ (256, 2) void k_gjx_correct (GjxArgs g)
{
  const int nb = g.nsz / GX_NB;
  const int tpc = nb * nb;
  const int item = gx_item (blockIdx.x, gridDim.x);
  const int ci = item / tpc, t = item - ci * tpc;
  const int cell = g.list[ci];
  if (g.flag[cell])
    return;
  const int n = g.nsz, ti = t / nb, tj = t - ti * nb;
  const float *X0 = g.Mf + g.offA[cell];
  const float *R = g.Rf + g.offA[cell];
  double *X = g.X + g.offA[cell];
  __shared__ __align__ (16) float Ls[GX_NB][GX_P];
  __shared__ __align__ (16) float Rs[GX_NB][GX_P];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rw = (wv >> 1) * 32, cw = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  gx_f4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
      acc[bi][bq] = gx_f4{ 0.f, 0.f, 0.f, 0.f };
  for (int kb = 0; kb < n; kb += GX_KC)
  {
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int k = idx >> 6, r = idx & 63;
      Ls[r][k] = X0[(size_t) (kb + k) * n + ti * GX_NB + r];
    }
    for (int idx = tid; idx < GX_KC * GX_NB; idx += 256)
    {
      const int c = idx >> 5, k = idx & 31;
      Rs[c][k] = R[(size_t) (tj * GX_NB + c) * n + kb + k];
    }
    __syncthreads ();
    gx_mma_half (acc, Ls, Rs, rw, cw, lr, lk);
    __syncthreads ();
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bq = 0; bq < 2; bq++)
    {
      double *dst = X + (size_t) (tj * GX_NB + cw + 16 * bq + lr) * n + ti * GX_NB + rw + 16 * bi + 4 * lk;
      const double2 x0 = *reinterpret_cast<const double2 *> (dst);
      const double2 x1 = *reinterpret_cast<const double2 *> (dst + 2);
      *reinterpret_cast<double2 *> (dst) = make_double2 (x0.x + (double) acc[bi][bq][0], x0.y + (double) acc[bi][bq][1]);
      *reinterpret_cast<double2 *> (dst + 2) =
          make_double2 (x1.x + (double) acc[bi][bq][2], x1.y + (double) acc[bi][bq][3]);
    }
}

// every launch of one size class of the mixed inverse
hipError_t gjx_enqueue_class (const GjxArgs &g0, int refine_steps, hipStream_t st)
{
  if (g0.ncells <= 0)
    return hipSuccess;
  GjxArgs g = g0;
  const int nb = g.nsz / GX_NB;
  hipLaunchKernelGGL (k_gjx_prepare, dim3 (g.ncells), dim3 (256), 0, st, g);
  hipError_t e = hipGetLastError ();
  for (int s = 0; s < nb && e == hipSuccess; s++)
  {
    g.sb = s;
    hipLaunchKernelGGL (k_gjx_diag, dim3 (g.ncells), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gjx_tile<2>, dim3 (g.ncells * (nb - 1)), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gjx_tile<4>, dim3 (g.ncells * (nb - 1) * (nb - 1)), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gjx_tile<5>, dim3 (g.ncells * (nb - 1)), dim3 (256), 0, st, g);
    e = hipGetLastError ();
  }
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL (k_gjx_widen, dim3 (g.ncells), dim3 (256), 0, st, g);
  for (int r = 0; r < refine_steps; r++)
  {
    hipLaunchKernelGGL (k_gjx_resid, dim3 (g.ncells * nb * nb), dim3 (256), 0, st, g);
    hipLaunchKernelGGL (k_gjx_correct, dim3 (g.ncells * nb * nb), dim3 (256), 0, st, g);
  }
  return hipGetLastError ();
}
