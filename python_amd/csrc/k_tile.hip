// k_tile.hip -- register-tiled FP64 LU factorization in GSL's operation order (gfx950), for
// orders above one wave's reach (64 < n <= 175).
//
// One workgroup of NW waves per matrix, [A | b] held in registers: lane l of wave w holds rows
// l + 64 r (r < R) and columns w*CPW .. w*CPW + CPW - 1 (column n is b).  Rows never move:
// each keeps its GSL position (pos) in a register, so GSL's pivot choice -- first strict maximum
// of |a| in position order (cblas/source_iamax_r.h:30-35) -- and its row swaps are reproduced on
// positions.  Step k:
//   1. the wave owning column k finds the pivot with a DPP reduction over its lanes, forms the
//      multipliers l_i = a_ik / a_pk (reciprocal multiply when |a_pk| >= DBL_MIN,
//      linalg/lu.c:162-176), keeps them in place and publishes them, with the pivot row and
//      position, in an LDS slot of the step's parity;
//   2. after one barrier every wave takes the pivot row's values for its own columns from the
//      lane that holds that row (through a wave-private LDS slot) and updates its block,
//      a_ij <- a_ij - l_i u_j with the product rounded first (no FMA contraction).
// Every element receives its updates in ascending k, the order of GSL's recursive LU
// (linalg/lu.c:204-288), so the factors equal GSL's.  The factors (GSL row order) and the
// forward-substituted b go to the factor store; k_bsub_f64 back-substitutes and checks the
// residual.
#include <float.h>

#include "hipla_dev.h"
#include "hipla_kernels.h"

template <int R, int CPW, int NW, bool KEEP>
__global__ __launch_bounds__ (64 * NW, (NW == 8) ? 4 : 4) void k_tile_f64 (SolveArgs s)
{
  __shared__ double s_l[2][64 * R];          // multipliers of the step, by physical row
  __shared__ int s_pv[2][4];                 // pivot row, pivot position, zero-pivot flag
  __shared__ double s_u[NW][CPW + 1];        // per-wave: the pivot row's values of its columns
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
          // ---- 1. the owner of column k: pivot and multipliers ----
          if (wv == kb)
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
                s_l[buf][row] = l;
              }
            }
            if (lane == 0)
            {
              s_pv[buf][0] = p;
              s_pv[buf][1] = pq;
              s_pv[buf][2] = piv == 0.0;
              if (KEEP && s.piv)
                s.piv[s.offV[cell] + k] = pq;
            }
          }
          __syncthreads ();
          const int p = __builtin_amdgcn_readfirstlane (s_pv[buf][0]);
          const int pq = __builtin_amdgcn_readfirstlane (s_pv[buf][1]);
          singular |= s_pv[buf][2];

          // ---- 2. positions: GSL swaps the rows at positions k and pq ----
#pragma unroll
          for (int r = 0; r < R; r++)
          {
            const int row = lane + 64 * r;
            if (row == p)
              pos[r] = k;
            else if (pos[r] == k)
              pos[r] = pq;
          }

          // ---- 3. the pivot row's values of this wave's columns, from the lane holding it ----
          if (lane == (p & 63))
          {
            const int ps = p >> 6;
#pragma unroll
            for (int c = 0; c < CPW; c++)
            {
              double col[R];
#pragma unroll
              for (int r = 0; r < R; r++)
                col[r] = a[r][c];
              s_u[wv][c] = pick_slot<R> (col, ps);
            }
          }
          double lr[R];
#pragma unroll
          for (int r = 0; r < R; r++)
            lr[r] = (pos[r] > k && pos[r] < n) ? s_l[buf][lane + 64 * r] : 0.0;

          // ---- 4. update: rows below the pivot, columns right of k ----
#pragma unroll
          for (int c = 0; c < CPW; c++)
          {
            if (c0 + c > k && c0 + c <= n)
            {
              const double u = s_u[wv][c];
#pragma unroll
              for (int r = 0; r < R; r++)
                if (pos[r] > k && pos[r] < n)
                  a[r][c] = a[r][c] - lr[r] * u;
            }
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
