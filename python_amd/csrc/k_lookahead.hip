// k_lookahead.hip -- FP64 partial-pivot LU in GSL's operation order, n <= 256: one workgroup of
// four waves per matrix, a one-wave panel and a look-ahead schedule (gfx950).
//
// The factorization is the right-looking blocked LU of k_blocked_f64: 16-column panels, rows
// tracked by position (no row moves inside a panel), and the factors plus the forward-substituted
// b left in the factor store in GSL's row order for k_bsub_f64.  Every element receives its
// updates a - l*u in ascending k with the product rounded first, so the values equal GSL's
// recursive LU (linalg/lu.c:204-288) bit for bit, with GSL's pivot choice (first strict max of
// |a| by current position, cblas/source_iamax_r.h:30-35) and multiplier rule (reciprocal
// multiply when |pivot| >= DBL_MIN, linalg/lu.c:162-176).  What differs is who does what, when:
//
//  * A panel is factored by ONE wave: lane l holds the panel rows l, l+64, l+128, l+192 (16
//    columns and b) in registers.  A pivot step is a choice among the lane's four rows, a wave
//    reduction, the pivot read into scalar registers and the rest of the pivot row broadcast
//    through the LDS crossbar (ds_bpermute): no barrier.
//  * Look-ahead: while wave (p+1) mod 4 factors panel p+1, the other three waves apply panel p to
//    the columns right of it; the panel chain overlaps that trailing update, and the panel wave
//    takes its share of the update once panel p+1 is out.
//  * The kernel runs for up to one cell per CU, so a workgroup has the CU to itself: 140 KB of LDS
//    hold two panel buffers (the panel being applied and the one being factored), the U12 rows
//    and the next panel's updated rows.
//  * The trailing update runs in units of one 16-column strip x all remaining rows: a lane keeps
//    its four columns' U12 values in registers and reads one multiplier pair per row and two
//    steps from LDS; units are handed out from an LDS counter.
//
// Epoch p (panel columns k0 .. k1-1; panel p already factored, in LDS and in the factor store in
// position order) is three stages, one barrier apart:
//   U. one thread per column outside the panel moves the rows the panel's pivoting moved (the
//      panel rows, and at most 16 rows displaced below it) and, right of the panel, forms
//      U12 = L11^-1 A12 in ascending order, into the factor store and LDS (Us);
//   V. four lanes per row (U12 of their four columns in registers) apply panel p's 16 updates to
//      the columns of panel p+1 and put the rows, with b, in LDS (Sn);
//   T. wave (p+1) mod 4 factors panel p+1 from Sn and writes it out (the other panel buffer, the
//      factor store, b, its position map, GSL's ipiv); the other waves apply A22 -= L21 U12 to
//      the columns from k1 + 16 on, joined by the panel wave when it is done.
// The first panel reads A and b directly; a row of A that no pivoting moved is read from A until
// the first trailing update has written it to the factor store.
//
// Against k_blocked_f64<256> (four-wave panel steps about 2 400 cycles apart, two workgroups per
// CU filling each other's gaps), this kernel is faster for one cell (0.61 against 0.86 ms at
// n = 256, back substitution included) and slower when several cells share each CU beside other
// size classes, so it runs for up to one cell per CU.
#include <hip/hip_runtime.h>

#include <float.h>

#include "hipla_dev.h"
#include "hipla_kernels.h"
#include "hipla_plan.h"

#define LA_NB 16     // panel width
#define LA_PP 18     // LDS pitch of a panel row (doubles): 16-byte rows, conflict-free 16-byte reads
#define LA_MR 4      // rows per lane
#define LA_MAXN 256  // largest order
#define LA_UP 256    // LDS pitch of a U12 row (doubles)
#define LA_NONE 0x7fffffff

__device__ __forceinline__ int la_uni (int v)
{
  return __builtin_amdgcn_readfirstlane (v);
}

// Keeps a value in a register of its own: products formed before the subtractions that use them
// stay distinct registers (left alone, the compiler funnels every product through one temporary,
// and each subtraction waits on the multiply issued just before it).
__device__ __forceinline__ void la_hold (double &x)
{
  __asm__ volatile ("" : "+v"(x));
}

// The held panel: v[m][0 .. 15] the panel columns and v[m][16] b of the row at relative
// position mp[m] (LA_NONE: no row).
struct LaPanel
{
  double v[LA_MR][LA_NB + 1];
  int mp[LA_MR];
};

// Row slot sb (wave-uniform) of lane wl, columns KK .. 16, to every lane (scalar registers).
// lane src's double in every lane, through the LDS crossbar (no LDS memory): the results land in
// vector registers, and one instruction per dword moves a whole row's value
__device__ __forceinline__ double la_bcast (double v, int src)
{
  const unsigned long long u = (unsigned long long) __double_as_longlong (v);
  const int lo = __builtin_amdgcn_ds_bpermute (src << 2, (int) (unsigned) u);
  const int hi = __builtin_amdgcn_ds_bpermute (src << 2, (int) (unsigned) (u >> 32));
  return __longlong_as_double ((long long) (((unsigned long long) (unsigned) hi << 32) | (unsigned) lo));
}

// the pivot row's columns KK+1 .. 16 (b last) from lane wl's slot sb, in every lane, and its
// pivot U_KK,KK in scalar registers (the multiplier rule branches on it)
template <int KK>
__device__ __forceinline__ double la_pivot_row (const LaPanel &pn, int sb, int wl, double (&uu)[LA_NB + 1])
{
  switch (sb)
  {
  case 0:
#pragma unroll
    for (int c = KK + 1; c <= LA_NB; c++)
      uu[c] = la_bcast (pn.v[0][c], wl);
    return rl_f64 (pn.v[0][KK], wl);
  case 1:
#pragma unroll
    for (int c = KK + 1; c <= LA_NB; c++)
      uu[c] = la_bcast (pn.v[1][c], wl);
    return rl_f64 (pn.v[1][KK], wl);
  case 2:
#pragma unroll
    for (int c = KK + 1; c <= LA_NB; c++)
      uu[c] = la_bcast (pn.v[2][c], wl);
    return rl_f64 (pn.v[2][KK], wl);
  default:
#pragma unroll
    for (int c = KK + 1; c <= LA_NB; c++)
      uu[c] = la_bcast (pn.v[3][c], wl);
    return rl_f64 (pn.v[3][KK], wl);
  }
}
template <int KK>
__device__ __forceinline__ void la_step (LaPanel &pn, int lane, int *ipiv, int &singular)
{
  long long key = -1;
  int bp = LA_NONE, bs = 0;
#pragma unroll
  for (int m = 0; m < LA_MR; m++)
    if (pn.mp[m] >= KK && pn.mp[m] != LA_NONE)
    {
      const long long q = pivot_key (pn.v[m][KK]);
      if (q > key || (q == key && pn.mp[m] < bp))
      {
        key = q;
        bp = pn.mp[m];
        bs = m;
      }
    }
  const int wl = max (wave_pivot_lane (key >= 0, key, bp), 0);
  const int pq = __builtin_amdgcn_readlane (bp, wl);
  const int sb = __builtin_amdgcn_readlane (bs, wl);
  double uu[LA_NB + 1];
  const double piv = la_pivot_row<KK> (pn, sb, wl, uu);
  singular |= (piv == 0.0);
  double l[LA_MR];
  if (fabs (piv) >= DBL_MIN)
  {
    const double inv = 1.0 / piv;
#pragma unroll
    for (int m = 0; m < LA_MR; m++)
      l[m] = pn.v[m][KK] * inv;
  }
  else
  {
#pragma unroll
    for (int m = 0; m < LA_MR; m++)
      l[m] = pn.v[m][KK] / piv;
  }
#pragma unroll
  for (int m = 0; m < LA_MR; m++)
  {
    const bool is_piv = lane == wl && m == sb;
    // every row still below the pivot, the row at position KK (moving to pq) included
    if (!is_piv && pn.mp[m] >= KK && pn.mp[m] != LA_NONE)
    {
      pn.v[m][KK] = l[m];
      double pr[LA_NB + 1];
#pragma unroll
      for (int c = KK + 1; c <= LA_NB; c++)
      {
        pr[c] = l[m] * uu[c];
        la_hold (pr[c]);
      }
#pragma unroll
      for (int c = KK + 1; c <= LA_NB; c++)
        pn.v[m][c] = pn.v[m][c] - pr[c];
    }
    if (is_piv)
      pn.mp[m] = KK;
    else if (pn.mp[m] == KK)
      pn.mp[m] = pq;
  }
  if (lane == 0)
    ipiv[KK] = pq;
}

template <int KK>
__device__ __forceinline__ void la_factor_from (LaPanel &pn, int nbk, int lane, int *ipiv, int &singular)
{
  if constexpr (KK < LA_NB)
  {
    if (KK < nbk)
    {
      la_step<KK> (pn, lane, ipiv, singular);
      la_factor_from<KK + 1> (pn, nbk, lane, ipiv, singular);
    }
  }
}

// A row of the trailing matrix at position q: in the first epoch (k0 == 0) a row no pivoting
// moved is still only in A.
__device__ __forceinline__ const double *la_src_row (const double *A, const double *W, const int *phys, bool first,
                                                    int q, int n)
{
  const double *src = (first && phys[q] == q) ? A : W;
  return src + (size_t) q * n;
}

// acc[m][0 .. H-1] -= l[m] * u[0 .. H-1] for the row slots m that hold rows (64 m < rows, a
// wave-uniform test): each row's products in registers of their own, then its subtractions.
template <int H>
__device__ __forceinline__ void la_rank1 (double (&acc)[LA_MR][H], const double (&l)[LA_MR], const double (&u)[H],
                                          int rows)
{
#pragma unroll
  for (int m = 0; m < LA_MR; m++)
    if (64 * m < rows)
    {
      double pr[H];
#pragma unroll
      for (int e = 0; e < H; e++)
      {
        pr[e] = l[m] * u[e];
        la_hold (pr[e]);
      }
#pragma unroll
      for (int e = 0; e < H; e++)
        acc[m][e] = acc[m][e] - pr[e];
    }
}

// One trailing unit: the 16 columns c0 .. c0+15 (those below n) of rows k1 .. n-1, minus the
// products of panel p (multipliers in P rows 16 .., U12 in Us), 16 steps in ascending order.
// Lane (g, h) = (lane >> 2, lane & 3) owns columns c0 + 4h .. c0 + 4h + 3, whose 16 U12 values
// stay in registers for the whole unit, and in row tile t the rows k1 + 32 t + g + 16 r (r < 2):
// four lanes share a row, so a load or store instruction covers 16 rows in 32-byte runs, and a
// step pair reads one multiplier pair per row from LDS (the 16 rows of a read 36 banks apart,
// four lanes on each address).
#define LA_TR 2
__device__ __forceinline__ void la_trail_strip (const double *A, double *W, const double *P, const double *Us,
                                                const int *phys, bool first, int n, int k1, int c0, int lane)
{
  const int rows1 = n - k1;
  const int g = lane >> 2, h = lane & 3;
  const int cl = c0 + 4 * h;                 // the lane's first column
  const bool full = c0 + LA_NB <= n;         // wave-uniform
  double u[LA_NB][4];
#pragma unroll
  for (int kk = 0; kk < LA_NB; kk++)
  {
    const double *us = Us + kk * LA_UP + (cl - k1);
    const double2 a = *reinterpret_cast<const double2 *> (us);
    const double2 b = *reinterpret_cast<const double2 *> (us + 2);
    u[kk][0] = a.x;
    u[kk][1] = a.y;
    u[kk][2] = b.x;
    u[kk][3] = b.y;
  }
  const int nt = (rows1 + 16 * LA_TR - 1) / (16 * LA_TR);
  auto load_tile = [&] (int t, double (&d)[LA_TR][4]) {
#pragma unroll
    for (int r = 0; r < LA_TR; r++)
    {
      const int i = min (16 * LA_TR * t + g + 16 * r, rows1 - 1);
      const double *row = la_src_row (A, W, phys, first, k1 + i, n);
      if (full)
      {
#pragma unroll
        for (int e = 0; e < 4; e++)
          d[r][e] = row[cl + e];
      }
      else
      {
#pragma unroll
        for (int e = 0; e < 4; e++)
          d[r][e] = row[min (cl + e, n - 1)];
      }
    }
  };
  for (int t = 0; t < nt; t++)
  {
    double acc[LA_TR][4];
    load_tile (t, acc);
    int pr[LA_TR];
#pragma unroll
    for (int r = 0; r < LA_TR; r++)
      pr[r] = (LA_NB + min (16 * LA_TR * t + g + 16 * r, rows1 - 1)) * LA_PP;
#pragma clang loop unroll(full)
    for (int kk = 0; kk < LA_NB; kk += 2)
    {
      double2 l2[LA_TR];
#pragma unroll
      for (int r = 0; r < LA_TR; r++)
        l2[r] = *reinterpret_cast<const double2 *> (P + pr[r] + kk);
#pragma unroll
      for (int s = 0; s < 2; s++)
      {
#pragma unroll
        for (int r = 0; r < LA_TR; r++)
        {
          const double l = s ? l2[r].y : l2[r].x;
#pragma unroll
          for (int e = 0; e < 4; e++)
            acc[r][e] = acc[r][e] - l * u[kk + s][e];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < LA_TR; r++)
    {
      const int i = 16 * LA_TR * t + g + 16 * r;
      if (i < rows1)
      {
        double *row = W + (size_t) (k1 + i) * n;
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (full || cl + e < n)
            row[cl + e] = acc[r][e];
      }
    }
  }
}

template <bool KEEP>
__global__ __launch_bounds__ (256, 1) void k_lookahead_f64 (SolveArgs s)
{
  extern __shared__ double smem[];
  // two panel buffers by parity ([LA_MAXN][LA_PP]: multipliers and U11 by relative position), the
  // U12 rows ([LA_NB][LA_UP], columns relative to k1) and the next panel's rows ([LA_MAXN][LA_PP]:
  // its 16 columns and b, updated, by relative position)
  double *Pbuf = smem;
  double *Us = smem + 2 * LA_MAXN * LA_PP;
  double *Sn = Us + LA_NB * LA_UP;
  __shared__ int phys[2][LA_MAXN];           // per panel parity: relative position -> row it held
  __shared__ int mv_dst[2][LA_NB], mv_src[2][LA_NB], s_nmv[2];
  __shared__ int s_ipiv[LA_NB];
  __shared__ int s_next, s_sing;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = la_uni (tid >> 6);
  const int neff = list_len (s);
  if (!list_in_window (s, neff))
    return;
  if (tid == 0)
    s_sing = 0;
  __syncthreads ();

  for (int item = blockIdx.x; item < neff; item += gridDim.x)
  {
    const int cell = s.list[item];
    const int n = s.n[cell];
    if (n < 1 || n > LA_MAXN)
    {
      if (tid == 0)
        s.codes[cell] = 5;
      continue;
    }
    const double *A = s.A + s.offA[cell];
    const double *b = s.B + s.offV[cell];
    double *W = s.F + s.offA[cell];
    double *wb = s.Y + s.offV[cell];
    const int ns = (n + LA_NB - 1) / LA_NB;

    // epoch -1 has only stage T: wave 0 factors the first panel straight from A and b
    for (int p = -1; p < ns; p++)
    {
      const int k0 = p * LA_NB;
      const int nbk = min (LA_NB, n - k0);
      const int k1 = k0 + nbk;                  // p = -1: 0
      const int rows1 = n - k1;
      const int pb = p & 1;
      const double *P = Pbuf + pb * LA_MAXN * LA_PP;
      const int *ph = phys[pb];
      const bool first = p == 0;
      if (p >= 0)
      {
        __syncthreads ();
        const int nmv = la_uni (s_nmv[pb]);
        if (tid == 0)
          s_next = 0;

        // ---- U. row moves outside the panel, U12 right of it ----
        {
          const int c = tid;
          if (c < n && (c < k0 || c >= k1))
          {
            const double *S = first ? A : W;
            const int last = nbk - 1;
            double u[LA_NB];
#pragma unroll
            for (int kk = 0; kk < LA_NB; kk++)
              u[kk] = S[(size_t) (k0 + la_uni (ph[min (kk, last)])) * n + c];
            if (nmv > 0)
            {
              double dv[LA_NB];
#pragma unroll
              for (int i = 0; i < LA_NB; i++)
                dv[i] = S[(size_t) (k0 + la_uni (mv_src[pb][min (i, nmv - 1)])) * n + c];
#pragma unroll
              for (int i = 0; i < LA_NB; i++)
                if (i < nmv)
                  W[(size_t) (k0 + la_uni (mv_dst[pb][i])) * n + c] = dv[i];
            }
            if (c >= k1)
            {
              // a column right of the panel exists only when the panel is full
#pragma unroll
              for (int r = 1; r < LA_NB; r++)
              {
                const double *Lr = P + r * LA_PP;
#pragma unroll
                for (int kk = 0; kk < r; kk++)
                  u[r] = u[r] - Lr[kk] * u[kk];
              }
#pragma unroll
              for (int r = 0; r < LA_NB; r++)
              {
                W[(size_t) (k0 + r) * n + c] = u[r];
                Us[r * LA_UP + (c - k1)] = u[r];
              }
            }
            else
            {
#pragma unroll
              for (int kk = 0; kk < LA_NB; kk++)
                if (kk < nbk && la_uni (ph[min (kk, last)]) != kk)
                  W[(size_t) (k0 + kk) * n + c] = u[kk];
            }
          }
        }
        __syncthreads ();

        // ---- V. the next panel's columns minus panel p's products into Sn: lane (g, h) = (tid >> 2,
        //      tid & 3) owns four of the 16 columns, with their U12 values in registers, and rows
        //      g, g + 64, ...; a row's four lanes read 128 contiguous bytes ----
        if (p + 1 < ns)
        {
          const int nbk1 = min (LA_NB, rows1);
          const int g = tid >> 2, h = tid & 3;
          double u[LA_NB][4];
#pragma unroll
          for (int kk = 0; kk < LA_NB; kk++)
          {
            const double2 x = *reinterpret_cast<const double2 *> (Us + kk * LA_UP + 4 * h);
            const double2 y = *reinterpret_cast<const double2 *> (Us + kk * LA_UP + 4 * h + 2);
            u[kk][0] = x.x;
            u[kk][1] = x.y;
            u[kk][2] = y.x;
            u[kk][3] = y.y;
          }
          for (int r = g; r < rows1; r += 64)
          {
            const double *row = la_src_row (A, W, ph, first, k1 + r, n) + k1;
            double a[4];
#pragma unroll
            for (int e = 0; e < 4; e++)
              a[e] = 4 * h + e < nbk1 ? row[min (4 * h + e, nbk1 - 1)] : 0.0;
            const double *Lr = P + (LA_NB + r) * LA_PP;
            // panel p's 16 updates in ascending order (panel p is full: a column follows it)
#pragma unroll
            for (int kk = 0; kk < LA_NB; kk += 2)
            {
              const double2 l2 = *reinterpret_cast<const double2 *> (Lr + kk);
#pragma unroll
              for (int e = 0; e < 4; e++)
                a[e] = a[e] - l2.x * u[kk][e];
#pragma unroll
              for (int e = 0; e < 4; e++)
                a[e] = a[e] - l2.y * u[kk + 1][e];
            }
            double *dst = Sn + r * LA_PP + 4 * h;
            *reinterpret_cast<double2 *> (dst) =
              make_double2 (4 * h < nbk1 ? a[0] : 0.0, 4 * h + 1 < nbk1 ? a[1] : 0.0);
            *reinterpret_cast<double2 *> (dst + 2) =
              make_double2 (4 * h + 2 < nbk1 ? a[2] : 0.0, 4 * h + 3 < nbk1 ? a[3] : 0.0);
            if (h == 0)
              Sn[r * LA_PP + LA_NB] = wb[k1 + r];
          }
        }
        __syncthreads ();
      }

      // ---- T. panel p+1 on one wave, A22 on the others ----
      if (p + 1 < ns)
      {
        if (wv == ((p + 1) & 3))
        {
          LaPanel pn;
          int singular = 0;
          const int nbk1 = min (LA_NB, rows1);
#pragma unroll
          for (int m = 0; m < LA_MR; m++)
          {
            const int r = lane + 64 * m;
            pn.mp[m] = r < rows1 ? r : LA_NONE;
            const int rc = min (r, rows1 - 1);
            if (p < 0)
            {
              const double *row = A + (size_t) rc * n;
#pragma unroll
              for (int c = 0; c < LA_NB; c++)
                pn.v[m][c] = c < nbk1 ? row[min (c, nbk1 - 1)] : 0.0;
              pn.v[m][LA_NB] = b[rc];
            }
            else
            {
              const double *src = Sn + rc * LA_PP;
#pragma unroll
              for (int c = 0; c < LA_NB; c += 2)
              {
                const double2 t = *reinterpret_cast<const double2 *> (src + c);
                pn.v[m][c] = t.x;
                pn.v[m][c + 1] = t.y;
              }
              pn.v[m][LA_NB] = src[LA_NB];
            }
          }
          __builtin_amdgcn_s_setprio (2);
          la_factor_from<0> (pn, nbk1, lane, s_ipiv, singular);
          __builtin_amdgcn_s_setprio (0);

          // publish panel p+1 in position order: multipliers and U11 to the other panel buffer, its
          // columns to the factor store, b to Y, the position map, the rows pivoting moved below the
          // panel, GSL's ipiv
          const int nb1 = (p + 1) & 1;
          double *Pn = Pbuf + nb1 * LA_MAXN * LA_PP;
          int *phn = phys[nb1];
#pragma unroll
          for (int m = 0; m < LA_MR; m++)
          {
            const int r = lane + 64 * m;
            if (r < rows1)
            {
              const int q = pn.mp[m];
              double *pr = Pn + q * LA_PP;
#pragma unroll
              for (int c = 0; c < LA_NB; c += 2)
                *reinterpret_cast<double2 *> (pr + c) = make_double2 (pn.v[m][c], pn.v[m][c + 1]);
              double *wr = W + (size_t) (k1 + q) * n + k1;
#pragma unroll
              for (int c = 0; c < LA_NB; c++)
                if (c < nbk1)
                  wr[c] = pn.v[m][c];
              wb[k1 + q] = pn.v[m][LA_NB];
              phn[q] = r;
            }
          }
          const bool d = lane < nbk1 && pn.mp[0] >= nbk1;   // rows 0 .. nbk1-1 are all in slot 0
          const unsigned long long mask = __ballot (d);
          if (d)
          {
            const int i = __builtin_amdgcn_mbcnt_hi ((unsigned) (mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo ((unsigned) mask, 0u));
            mv_dst[nb1][i] = pn.mp[0];
            mv_src[nb1][i] = lane;
          }
          if (lane == 0)
          {
            s_nmv[nb1] = __popcll (mask);
            if (singular)
              s_sing = 1;
          }
          if (KEEP && s.piv && lane < nbk1)
            s.piv[s.offV[cell] + k1 + lane] = k1 + s_ipiv[lane];
        }
        // the panel wave, once panel p+1 is out, takes strips too
        if (p >= 0)
        {
          const int nunit = ns - p - 2;
          for (;;)
          {
            int u = 0;
            if (lane == 0)
              u = atomicAdd (&s_next, 1);
            u = __builtin_amdgcn_readlane (u, 0);
            if (u >= nunit)
              break;
            la_trail_strip (A, W, P, Us, ph, first, n, k1, k1 + LA_NB * (u + 1), lane);
          }
        }
      }
    }
    __syncthreads ();
    if (tid == 0)
    {
      s.codes[cell] = s_sing ? 4 : 0;
      s_sing = 0;
    }
  }
}

int lookahead_f64_max_cells ()
{
  return hipla_device_info ().num_cu;
}

size_t lookahead_f64_smem ()
{
  return (size_t) (3 * LA_MAXN * LA_PP + LA_NB * LA_UP) * sizeof (double);
}

hipError_t launch_lookahead_f64 (const SolveArgs &a, int grid, bool keep, hipStream_t st)
{
  if (a.nlist <= 0)
    return hipSuccess;
  if (!a.F || !a.Y)
    return hipErrorInvalidValue;
  const size_t sm = lookahead_f64_smem ();
  if (keep)
    hipLaunchKernelGGL ((k_lookahead_f64<true>), dim3 (grid), dim3 (256), sm, st, a);
  else
    hipLaunchKernelGGL ((k_lookahead_f64<false>), dim3 (grid), dim3 (256), sm, st, a);
  return hipGetLastError ();
}

hipError_t setup_lookahead_attrs (size_t max_smem)
{
  const void *ks[2] = { (const void *) k_lookahead_f64<false>, (const void *) k_lookahead_f64<true> };
  for (const void *k : ks)
  {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes (&fa, k) != hipSuccess)
      continue;
    const size_t dyn = max_smem > fa.sharedSizeBytes ? max_smem - fa.sharedSizeBytes : 0;
    (void) hipFuncSetAttribute (k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) dyn);
    (void) hipGetLastError ();
  }
  return hipSuccess;
}
