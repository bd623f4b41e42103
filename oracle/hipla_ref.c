/*
 * oracle/hipla_ref.c -- CPU oracle for the per-cell dense solve / inverse path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check and the
 * cpu_baseline leg of bench.py call it to produce the expected answers and the CPU timing.
 * The product library (python_amd/csrc, libhipla.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of the reference path:
 *   source/matrix_cpu.c:96-204   cpu_solve_matrix   (LU, solve, residual check, codes 0/2/3/4)
 *   source/matrix_cpu.c:223-257  cpu_invert_matrix  (LU, LU_invert, row-major copy-out)
 * on top of GSL 2.6 as vendored in software/gsl-2.6:
 *   linalg/lu.c:61-115    gsl_linalg_LU_decomp      linalg/lu.c:128-191  LU_decomp_L2
 *   linalg/lu.c:204-288   LU_decomp_L3              linalg/lu.c:290-360  LU_solve / LU_svx
 *   linalg/lu.c:414-481   LU_invert / LU_invx       linalg/lu.c:539-578  singular / apply_pivots
 *   linalg/invtri.c:37-229  tri_invert (L2 + L3)    linalg/trimult.c:44-48,203-334  tri_UL
 *   linalg/recurse.h:1-10   split rule and crossover
 *   cblas/source_{iamax_r,scal_r,ger,gemm_r,trsm_r,trsv_r,trmv_r,trmm_r,gemv_r,dot_r}.h
 *   permutation/permute_source.c:20-140
 * Every loop keeps the reference's operation order (no FMA contraction: build with
 * -ffp-contract=off), so results are those of the reference's GSL build on x86-64.
 *
 * Parity status of each entry point is listed in its comment ("pinned by" = the checks in
 * tests/test_oracle.py that would catch a dropped term, a wrong index or a transposed operand).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hipla_ref.h"

/* linalg/recurse.h:2,6-7 */
#define REF_SPLIT(n) (((n) >= 16) ? (((n) + 8) / 16) * 8 : (n) / 2)
#define REF_CROSSOVER 24

/* GSL status codes used by the path (gsl_errno.h): GSL_EDOM = 1, GSL_ENOMEM = 8, GSL_ESING = 21 */
#define REF_GSL_EDOM 1
#define REF_GSL_ENOMEM 8
#define REF_GSL_ESING 21

/* residual tolerance of the solve check: EPSILON, source/python.h:98 */
#define REF_EPSILON 1.e-6

/* A row-major sub-matrix view (what gsl_matrix_submatrix hands to the cblas routines). */
typedef struct
{
  double *d;
  size_t ld, r, c;
} rview;

static rview
sub (rview a, size_t i0, size_t j0, size_t r, size_t c)
{
  rview v = { a.d + i0 * a.ld + j0, a.ld, r, c };
  return v;
}

#define AT(v, i, j) ((v).d[(i) * (v).ld + (j)])

/* cblas/source_iamax_r.h:20-39: first index of the strict maximum of |x|, 0 when none is > 0 */
static size_t
ref_iamax_col (rview a, size_t col, size_t row0, size_t n)
{
  double max = 0.0;
  size_t result = 0;
  for (size_t i = 0; i < n; i++)
  {
    double v = fabs (AT (a, row0 + i, col));
    if (v > max)
    {
      max = v;
      result = i;
    }
  }
  return result;
}

static void
swap_rows (rview a, size_t i, size_t j)
{
  for (size_t c = 0; c < a.c; c++)
  {
    double t = AT (a, i, c);
    AT (a, i, c) = AT (a, j, c);
    AT (a, j, c) = t;
  }
}

/* linalg/lu.c:553-578 apply_pivots */
static void
apply_pivots (rview a, const unsigned *ipiv, size_t n)
{
  for (size_t i = 0; i < n; i++)
    if (ipiv[i] != i)
      swap_rows (a, i, ipiv[i]);
}

/* linalg/lu.c:128-191 LU_decomp_L2 (LAPACK dgetf2); dger from cblas/source_ger.h:25-36 */
static void
lu_decomp_L2 (rview A, unsigned *ipiv)
{
  const size_t M = A.r, N = A.c, minMN = M < N ? M : N;
  for (size_t j = 0; j < minMN; ++j)
  {
    size_t jp = j + ref_iamax_col (A, j, j, M - j);
    ipiv[j] = (unsigned) jp;
    if (jp != j)
      swap_rows (A, j, jp);
    if (j < M - 1)
    {
      double Ajj = AT (A, j, j);
      if (fabs (Ajj) >= DBL_MIN)
      {
        double s = 1.0 / Ajj;   /* dscal: X *= alpha, cblas/source_scal_r.h */
        for (size_t i = j + 1; i < M; i++)
          AT (A, i, j) *= s;
      }
      else
      {
        for (size_t i = 1; i < M - j; ++i)
          AT (A, j + i, j) /= Ajj;
      }
    }
    if (j < minMN - 1)
    {
      /* dger(-1, x = A[j+1:, j], y = A[j, j+1:], A22): A_ik += y_k * (alpha * x_i) */
      for (size_t i = j + 1; i < M; i++)
      {
        const double tmp = -1.0 * AT (A, i, j);
        for (size_t k = j + 1; k < N; k++)
          AT (A, i, k) += AT (A, j, k) * tmp;
      }
    }
  }
}

/* cblas/source_trsm_r.h, Left/Lower/NoTrans, alpha = 1: B := inv(TriL(A)) B */
static void
trsm_left_lower_notrans (int nonunit, rview A, rview B)
{
  for (size_t i = 0; i < B.r; i++)
  {
    if (nonunit)
    {
      double Aii = AT (A, i, i);
      for (size_t j = 0; j < B.c; j++)
        AT (B, i, j) /= Aii;
    }
    for (size_t k = i + 1; k < B.r; k++)
    {
      const double Aki = AT (A, k, i);
      for (size_t j = 0; j < B.c; j++)
        AT (B, k, j) -= Aki * AT (B, i, j);
    }
  }
}

/* cblas/source_trsm_r.h:157-183, Right/Upper/NoTrans, alpha = 1: B := B inv(TriU(A)) */
static void
trsm_right_upper_notrans (int nonunit, rview A, rview B)
{
  for (size_t i = 0; i < B.r; i++)
    for (size_t j = 0; j < B.c; j++)
    {
      if (nonunit)
        AT (B, i, j) /= AT (A, j, j);
      const double Bij = AT (B, i, j);
      for (size_t k = j + 1; k < B.c; k++)
        AT (B, i, k) -= AT (A, j, k) * Bij;
    }
}

/* cblas/source_gemm_r.h:70-83, NoTrans/NoTrans, beta = 1: C += alpha A B, skipping zero alpha*A_ik */
static void
gemm_nn (double alpha, rview A, rview B, rview C)
{
  if (alpha == 0.0)
    return;
  for (size_t k = 0; k < A.c; k++)
    for (size_t i = 0; i < C.r; i++)
    {
      const double temp = alpha * AT (A, i, k);
      if (temp != 0.0)
        for (size_t j = 0; j < C.c; j++)
          AT (C, i, j) += temp * AT (B, k, j);
    }
}

/* cblas/source_trmm_r.h:43-63, Left/Upper/NoTrans: B := alpha TriU(A) B */
static void
trmm_left_upper_notrans (int nonunit, double alpha, rview A, rview B)
{
  for (size_t i = 0; i < B.r; i++)
    for (size_t j = 0; j < B.c; j++)
    {
      double temp = nonunit ? AT (A, i, i) * AT (B, i, j) : AT (B, i, j);
      for (size_t k = i + 1; k < B.r; k++)
        temp += AT (A, i, k) * AT (B, k, j);
      AT (B, i, j) = alpha * temp;
    }
}

/* cblas/source_trmm_r.h:178-199, Right/Lower/NoTrans: B := alpha B TriL(A) */
static void
trmm_right_lower_notrans (int nonunit, double alpha, rview A, rview B)
{
  for (size_t i = 0; i < B.r; i++)
    for (size_t j = 0; j < B.c; j++)
    {
      double temp = nonunit ? AT (A, j, j) * AT (B, i, j) : AT (B, i, j);
      for (size_t k = j + 1; k < B.c; k++)
        temp += AT (A, k, j) * AT (B, i, k);
      AT (B, i, j) = alpha * temp;
    }
}

/* linalg/lu.c:204-288 LU_decomp_L3 (ReLAPACK dgetrf recursion) */
static void
lu_decomp_L3 (rview A, unsigned *ipiv)
{
  const size_t M = A.r, N = A.c;
  if (N <= REF_CROSSOVER)
  {
    lu_decomp_L2 (A, ipiv);
    return;
  }
  const size_t N1 = REF_SPLIT (N), N2 = N - N1, M2 = M - N1;
  rview A11 = sub (A, 0, 0, N1, N1), A12 = sub (A, 0, N1, N1, N2);
  rview A21 = sub (A, N1, 0, M2, N1), A22 = sub (A, N1, N1, M2, N2);
  rview AL = sub (A, 0, 0, M, N1), AR = sub (A, 0, N1, M, N2);
  lu_decomp_L3 (AL, ipiv);
  apply_pivots (AR, ipiv, N1);
  trsm_left_lower_notrans (0, A11, A12);
  gemm_nn (-1.0, A21, A12, A22);
  lu_decomp_L3 (A22, ipiv + N1);
  apply_pivots (A21, ipiv + N1, N2);
  for (size_t i = 0; i < N2; ++i)
    ipiv[N1 + i] += (unsigned) N1;
}

/* ipiv -> gsl_permutation and signum, linalg/lu.c:93-109 */
static int
ipiv_to_perm (const unsigned *ipiv, size_t n, size_t *perm)
{
  int signum = 1;
  for (size_t i = 0; i < n; i++)
    perm[i] = i;
  for (size_t i = 0; i < n; ++i)
  {
    size_t pivi = ipiv[i];
    if (perm[pivi] != perm[i])
    {
      size_t t = perm[pivi];
      perm[pivi] = perm[i];
      perm[i] = t;
      signum = -signum;
    }
  }
  return signum;
}

/* gsl_linalg_LU_decomp, linalg/lu.c:61-115 (square case).  Pinned by: P.L.U == A invariant,
 * bit-for-bit agreement with the independent right-looking restatement below, the reference
 * fixtures, GSL's Hilbert / Vandermonde known answers and exact rational solves. */
int
ref_lu_decomp (double *a, int n, size_t *perm, unsigned *ipiv, int *signum)
{
  if (n <= 0)
    return -1;
  rview A = { a, (size_t) n, (size_t) n, (size_t) n };
  lu_decomp_L3 (A, ipiv);
  int s = ipiv_to_perm (ipiv, (size_t) n, perm);
  if (signum)
    *signum = s;
  return 0;
}

/* ---- independent restatement: plain right-looking GEPP (the GPU kernels' order) ----
 *
 * Every element a(pos, j) receives its updates for k = 0, 1, ... in ascending order, each as
 * a - l*u with the product rounded first, exactly as in the recursion above, so the values
 * agree with the recursion.  The recursion's dgemm (cblas/source_gemm_r.h:76) additionally skips
 * exactly-zero multipliers while its dger / dtrsm do not; that decides only signs of zero results
 * and whether inf*0 produces a NaN.  This restatement (like the GPU kernels) always applies the
 * update. */
int
ref_lu_decomp_rightlooking (double *a, int n, size_t *perm, unsigned *ipiv, int *signum)
{
  if (n <= 0)
    return -1;
  rview A = { a, (size_t) n, (size_t) n, (size_t) n };
  for (int k = 0; k < n; k++)
  {
    size_t p = (size_t) k + ref_iamax_col (A, (size_t) k, (size_t) k, (size_t) (n - k));
    ipiv[k] = (unsigned) p;
    if (p != (size_t) k)
      swap_rows (A, (size_t) k, p);
    if (k < n - 1)
    {
      double piv = AT (A, k, k);
      if (fabs (piv) >= DBL_MIN)
      {
        double s = 1.0 / piv;
        for (int i = k + 1; i < n; i++)
          AT (A, i, k) *= s;
      }
      else
        for (int i = k + 1; i < n; i++)
          AT (A, i, k) /= piv;
    }
    for (int i = k + 1; i < n; i++)
    {
      const double l = AT (A, i, k);
      for (int j = k + 1; j < n; j++)
      {
        const double prod = l * AT (A, k, j);
        AT (A, i, j) = AT (A, i, j) - prod;
      }
    }
  }
  int s = ipiv_to_perm (ipiv, (size_t) n, perm);
  if (signum)
    *signum = s;
  return 0;
}

/* linalg/lu.c:539-551 */
static int
lu_singular (const double *lu, int n)
{
  for (int i = 0; i < n; i++)
    if (lu[(size_t) i * n + i] == 0)
      return 1;
  return 0;
}

/* gsl_linalg_LU_solve / LU_svx, linalg/lu.c:290-360: x = P b, dtrsv lower-unit forward,
 * dtrsv upper-nonunit backward (cblas/source_trsv_r.h: row dot products, ascending j). */
int
ref_lu_solve (const double *lu, int n, const size_t *perm, const double *b, double *x)
{
  if (lu_singular (lu, n))
    return REF_GSL_EDOM;
  for (int i = 0; i < n; i++)   /* gsl_permute_vector: OUT[i] = IN[perm[i]] */
    x[i] = b[perm[i]];
  for (int i = 1; i < n; i++)
  {
    double tmp = x[i];
    for (int j = 0; j < i; j++)
      tmp -= lu[(size_t) i * n + j] * x[j];
    x[i] = tmp;
  }
  x[n - 1] = x[n - 1] / lu[(size_t) (n - 1) * n + (n - 1)];
  for (int i = n - 2; i >= 0; i--)
  {
    double tmp = x[i];
    for (int j = i + 1; j < n; j++)
      tmp -= lu[(size_t) i * n + j] * x[j];
    x[i] = tmp / lu[(size_t) i * n + i];
  }
  return 0;
}

/* gsl_blas_dgemv(NoTrans, 1, A, x, 0, y), cblas/source_gemv_r.h:52-66 */
static void
gemv_n (const double *a, int n, const double *x, double *y)
{
  for (int i = 0; i < n; i++)
  {
    double temp = 0.0;
    for (int j = 0; j < n; j++)
      temp += x[j] * a[(size_t) i * n + j];
    y[i] = 0.0;
    y[i] += 1.0 * temp;
  }
}

/* The residual test of source/matrix_cpu.c:155-185: first failing row only sets the code. */
int
ref_residual_code (const double *a_orig, int n, const double *b, const double *x)
{
  double *t = (double *) malloc (sizeof (double) * (size_t) n);
  int code = 0, n_error = 0;
  if (!t)
    return REF_GSL_ENOMEM;
  gemv_n (a_orig, n, x, t);
  for (int i = 0; i < n; i++)
  {
    double test_val = t[i];
    if (b[i] > 0.0)
    {
      if (n_error == 0 && fabs ((test_val - b[i]) / test_val) > REF_EPSILON)
      {
        code = 2;
        n_error += 1;
      }
    }
    else if (n_error == 0 && fabs (test_val - b[i]) > REF_EPSILON)
    {
      code = 3;
      n_error += 1;
    }
  }
  free (t);
  return code;
}

/* cpu_solve_matrix, source/matrix_cpu.c:96-204.  a is overwritten with the LU factors, as the
 * reference does (matrix_cpu.c:133).  Returns 0, 2, 3 or 4 (x untouched on 4); -1 for n <= 0,
 * which the reference does not guard (gsl_permutation_alloc(0) fails there). */
int
ref_solve_matrix (double *a, const double *b, int n, double *x)
{
  if (n <= 0)
    return -1;
  size_t nn = (size_t) n * n;
  double *test = (double *) malloc (sizeof (double) * nn);
  double *sol = (double *) malloc (sizeof (double) * (size_t) n);
  size_t *perm = (size_t *) malloc (sizeof (size_t) * (size_t) n);
  unsigned *ipiv = (unsigned *) malloc (sizeof (unsigned) * (size_t) n);
  int code;
  if (!test || !sol || !perm || !ipiv)
  {
    free (test);
    free (sol);
    free (perm);
    free (ipiv);
    return REF_GSL_ENOMEM;
  }
  memcpy (test, a, sizeof (double) * nn);
  ref_lu_decomp (a, n, perm, ipiv, NULL);
  if (ref_lu_solve (a, n, perm, b, sol))
  {
    code = 4;
  }
  else
  {
    code = ref_residual_code (test, n, b, sol);
    memcpy (x, sol, sizeof (double) * (size_t) n);
  }
  free (test);
  free (sol);
  free (perm);
  free (ipiv);
  return code;
}

/* ---------------- inverse: linalg/lu.c:436-481 LU_invx ---------------- */

/* linalg/invtri.c:73-157 triangular_inverse_L2 (LAPACK dtrti2) with dtrmv from
 * cblas/source_trmv_r.h (NoTrans upper: ascending rows; NoTrans lower: descending rows). */
static void
tri_inverse_L2 (int upper, int nonunit, rview T)
{
  const size_t N = T.r;
  if (upper)
  {
    for (size_t i = 0; i < N; ++i)
    {
      double aii;
      if (nonunit)
      {
        AT (T, i, i) = 1.0 / AT (T, i, i);
        aii = -AT (T, i, i);
      }
      else
        aii = -1.0;
      if (i > 0)
      {
        /* v = T[0:i, i]; v := TriU(T[0:i,0:i]) v; v *= aii */
        for (size_t r = 0; r < i; r++)
        {
          double temp = 0.0;
          for (size_t j = r + 1; j < i; j++)
            temp += AT (T, j, i) * AT (T, r, j);
          if (nonunit)
            AT (T, r, i) = temp + AT (T, r, i) * AT (T, r, r);
          else
            AT (T, r, i) += temp;
        }
        for (size_t r = 0; r < i; r++)
          AT (T, r, i) *= aii;
      }
    }
  }
  else
  {
    for (size_t ii = 0; ii < N; ++ii)
    {
      size_t j = N - ii - 1;
      double ajj;
      if (nonunit)
      {
        AT (T, j, j) = 1.0 / AT (T, j, j);
        ajj = -AT (T, j, j);
      }
      else
        ajj = -1.0;
      if (j < N - 1)
      {
        /* m = T[j+1:, j+1:], v = T[j+1:, j]; v := TriL(m) v (rows descending); v *= ajj */
        const size_t m = N - j - 1;
        for (size_t r = m; r > 0 && r--;)
        {
          double temp = 0.0;
          for (size_t c = 0; c < r; c++)
            temp += AT (T, j + 1 + c, j) * AT (T, j + 1 + r, j + 1 + c);
          if (nonunit)
            AT (T, j + 1 + r, j) = temp + AT (T, j + 1 + r, j) * AT (T, j + 1 + r, j + 1 + r);
          else
            AT (T, j + 1 + r, j) += temp;
        }
        for (size_t r = 0; r < m; r++)
          AT (T, j + 1 + r, j) *= ajj;
      }
    }
  }
}

/* trsm Left/Lower/NoTrans restricted to the lower triangle of A (used on T22 in place) */
static void
trsm_left_lower_tri (int nonunit, rview A, rview B)
{
  trsm_left_lower_notrans (nonunit, A, B);
}

/* linalg/invtri.c:168-229 triangular_inverse_L3 */
static void
tri_inverse_L3 (int upper, int nonunit, rview T)
{
  const size_t N = T.r;
  if (N <= REF_CROSSOVER)
  {
    tri_inverse_L2 (upper, nonunit, T);
    return;
  }
  const size_t N1 = REF_SPLIT (N), N2 = N - N1;
  rview T11 = sub (T, 0, 0, N1, N1), T12 = sub (T, 0, N1, N1, N2);
  rview T21 = sub (T, N1, 0, N2, N1), T22 = sub (T, N1, N1, N2, N2);
  tri_inverse_L3 (upper, nonunit, T11);
  if (!upper)
  {
    trmm_right_lower_notrans (nonunit, -1.0, T11, T21);
    trsm_left_lower_tri (nonunit, T22, T21);
  }
  else
  {
    trmm_left_upper_notrans (nonunit, -1.0, T11, T12);
    trsm_right_upper_notrans (nonunit, T22, T12);
  }
  tri_inverse_L3 (upper, nonunit, T22);
}

/* linalg/invtri.c:231-243 triangular_singular */
static int
tri_singular (rview T)
{
  for (size_t i = 0; i < T.r; ++i)
    if (AT (T, i, i) == 0.0)
      return REF_GSL_ESING;
  return 0;
}

/* linalg/trimult.c:203-264 triangular_mult_L2 (Upper: U*L) with ddot (cblas/source_dot_r.h),
 * dgemv Trans / NoTrans (cblas/source_gemv_r.h) and dscal. */
static void
tri_mult_L2 (rview A)
{
  const size_t N = A.r;
  if (N == 1)
    return;
  for (size_t i = 0; i < N; ++i)
  {
    const double Uii = AT (A, i, i);
    if (i < N - 1)
    {
      const size_t m = N - i - 1;
      double tmp = 0.0;
      for (size_t t = 0; t < m; t++)
        tmp += AT (A, i + 1 + t, i) * AT (A, i, i + 1 + t);
      AT (A, i, i) += tmp;
      if (i > 0)
      {
        /* ll = A[i, 0:i] := Uii*ll + L_BL^T ur,  L_BL = A[i+1:, 0:i], ur = A[i, i+1:] */
        if (Uii == 0.0)
          for (size_t c = 0; c < i; c++)
            AT (A, i, c) = 0.0;
        else if (Uii != 1.0)
          for (size_t c = 0; c < i; c++)
            AT (A, i, c) *= Uii;
        for (size_t r = 0; r < m; r++)
        {
          const double temp = 1.0 * AT (A, i, i + 1 + r);
          if (temp != 0.0)
            for (size_t c = 0; c < i; c++)
              AT (A, i, c) += temp * AT (A, i + 1 + r, c);
        }
        /* ut = A[0:i, i] += U_TR lb,  U_TR = A[0:i, i+1:], lb = A[i+1:, i] */
        for (size_t r = 0; r < i; r++)
        {
          double temp = 0.0;
          for (size_t c = 0; c < m; c++)
            temp += AT (A, i + 1 + c, i) * AT (A, r, i + 1 + c);
          AT (A, r, i) += 1.0 * temp;
        }
      }
    }
    else
    {
      for (size_t c = 0; c < N - 1; c++)
        AT (A, N - 1, c) *= Uii;
    }
  }
}

/* linalg/trimult.c:275-334 triangular_mult_L3 (Upper) */
static void
tri_mult_L3 (rview A)
{
  const size_t N = A.r;
  if (N <= REF_CROSSOVER)
  {
    tri_mult_L2 (A);
    return;
  }
  const size_t N1 = REF_SPLIT (N), N2 = N - N1;
  rview A11 = sub (A, 0, 0, N1, N1), A12 = sub (A, 0, N1, N1, N2);
  rview A21 = sub (A, N1, 0, N2, N1), A22 = sub (A, N1, N1, N2, N2);
  tri_mult_L3 (A11);
  gemm_nn (1.0, A12, A21, A11);
  trmm_right_lower_notrans (0, 1.0, A22, A12);
  trmm_left_upper_notrans (1, 1.0, A22, A21);
  tri_mult_L3 (A22);
}

/* gsl_linalg_LU_invx, linalg/lu.c:436-481, in place on the LU factors.  Pinned by: the
 * reference fixtures inverse_small (exact) and inverse_macro (7 significant digits, full
 * matrix), A.inv(A) == I, and the Neumann series sum Q^k for N = I - Q. */
int
ref_lu_invx (double *lu, int n, const size_t *perm)
{
  rview T = { lu, (size_t) n, (size_t) n, (size_t) n };
  int st;
  if (lu_singular (lu, n))
    return REF_GSL_EDOM;
  if ((st = tri_singular (T)))
    return st;
  tri_inverse_L3 (1, 1, T);
  if ((st = tri_singular (T)))
    return st;
  tri_inverse_L3 (0, 0, T);
  tri_mult_L3 (T);
  double *row = (double *) malloc (sizeof (double) * (size_t) n);
  if (!row)
    return REF_GSL_ENOMEM;
  for (int i = 0; i < n; ++i)
  {
    /* gsl_permute_vector_inverse: OUT[perm[k]] = IN[k] */
    double *ri = lu + (size_t) i * n;
    for (int k = 0; k < n; k++)
      row[perm[k]] = ri[k];
    memcpy (ri, row, sizeof (double) * (size_t) n);
  }
  free (row);
  return 0;
}

/* cpu_invert_matrix, source/matrix_cpu.c:223-257.  m is overwritten with the LU factors. */
int
ref_invert_matrix (double *m, double *inv, int n)
{
  if (n <= 0)
    return -1;
  size_t nn = (size_t) n * n;
  size_t *perm = (size_t *) malloc (sizeof (size_t) * (size_t) n);
  unsigned *ipiv = (unsigned *) malloc (sizeof (unsigned) * (size_t) n);
  double *work = (double *) malloc (sizeof (double) * nn);
  int st;
  if (!perm || !ipiv || !work)
  {
    free (perm);
    free (ipiv);
    free (work);
    return REF_GSL_ENOMEM;
  }
  ref_lu_decomp (m, n, perm, ipiv, NULL);
  memcpy (work, m, sizeof (double) * nn);
  st = ref_lu_invx (work, n, perm);
  if (st == 0)
    memcpy (inv, work, sizeof (double) * nn);
  free (perm);
  free (ipiv);
  free (work);
  return st;
}

/* ---------------- transpose solve (adjoint) ----------------
 * A^T lam = g with P A = L U:  A^T = U^T L^T P, so  U^T y = g (forward), L^T z = y (backward,
 * unit), lam = P^T z (lam[perm[i]] = z[i]).  GSL 2.6 has no transpose LU solve, so this has
 * no reference counterpart; it is pinned by the residual A^T lam - g and by central finite
 * differences of the loss in tests/test_oracle.py. */
int
ref_lu_solve_transpose (const double *lu, int n, const size_t *perm, const double *g, double *lam)
{
  if (lu_singular (lu, n))
    return REF_GSL_EDOM;
  double *y = (double *) malloc (sizeof (double) * (size_t) n);
  if (!y)
    return REF_GSL_ENOMEM;
  for (int i = 0; i < n; i++)
  {
    double tmp = g[i];
    for (int j = 0; j < i; j++)
      tmp -= lu[(size_t) j * n + i] * y[j];
    y[i] = tmp / lu[(size_t) i * n + i];
  }
  for (int i = n - 1; i >= 0; i--)
  {
    double tmp = y[i];
    for (int j = i + 1; j < n; j++)
      tmp -= lu[(size_t) j * n + i] * y[j];
    y[i] = tmp;
  }
  for (int i = 0; i < n; i++)
    lam[perm[i]] = y[i];
  free (y);
  return 0;
}

/* ---------------- double-double residual refinement (truth for ill-conditioned cells) ----
 * Not part of the reference: x is refined with r = b - A x evaluated in double-double
 * (two-sum / two-product, Dekker 1971; Knuth TAOCP vol. 2 4.2.2) and corrections solved with
 * the LU factors.  Used only to judge which of two FP64 answers is closer to the exact
 * solution when the matrix is too ill-conditioned for FP64 to settle it. */
static inline void
two_sum (double a, double b, double *s, double *e)
{
  double x = a + b, bb = x - a;
  *s = x;
  *e = (a - (x - bb)) + (b - bb);
}

static inline void
two_prod (double a, double b, double *p, double *e)
{
  /* Dekker split: 2^27 + 1 */
  const double C = 134217729.0;
  double x = a * b;
  double t = C * a, ah = t - (t - a), al = a - ah;
  t = C * b;
  double bh = t - (t - b), bl = b - bh;
  *p = x;
  *e = ((ah * bh - x) + ah * bl + al * bh) + al * bl;
}

int
ref_solve_dd_refined (const double *a, const double *b, int n, double *x, int iters)
{
  size_t nn = (size_t) n * n;
  double *lu = (double *) malloc (sizeof (double) * nn);
  size_t *perm = (size_t *) malloc (sizeof (size_t) * (size_t) n);
  unsigned *ipiv = (unsigned *) malloc (sizeof (unsigned) * (size_t) n);
  double *r = (double *) malloc (sizeof (double) * (size_t) n);
  double *d = (double *) malloc (sizeof (double) * (size_t) n);
  double *xlo = (double *) calloc ((size_t) n, sizeof (double));
  int st = 0;
  if (!lu || !perm || !ipiv || !r || !d || !xlo)
  {
    st = REF_GSL_ENOMEM;
    goto done;
  }
  memcpy (lu, a, sizeof (double) * nn);
  ref_lu_decomp (lu, n, perm, ipiv, NULL);
  if ((st = ref_lu_solve (lu, n, perm, b, x)))
    goto done;
  for (int it = 0; it < iters; it++)
  {
    for (int i = 0; i < n; i++)
    {
      /* r_i = b_i - sum_j a_ij (x_j + xlo_j) in double-double */
      double sh = b[i], sl = 0.0;
      for (int j = 0; j < n; j++)
      {
        double p, pe, s, se;
        two_prod (a[(size_t) i * n + j], x[j], &p, &pe);
        two_sum (sh, -p, &s, &se);
        sl += se - pe - a[(size_t) i * n + j] * xlo[j];
        sh = s;
      }
      r[i] = sh + sl;
    }
    if (ref_lu_solve (lu, n, perm, r, d))
      break;
    for (int j = 0; j < n; j++)
    {
      double s, e;
      two_sum (x[j], d[j] + xlo[j], &s, &e);
      x[j] = s;
      xlo[j] = e;
    }
  }
done:
  free (lu);
  free (perm);
  free (ipiv);
  free (r);
  free (d);
  free (xlo);
  return st;
}

/* ---------------- batched CPU baseline ----------------
 * One worker per core over a ragged batch, each worker running the serial per-cell
 * cpu_solve_matrix (as MPI ranks do over cells in source/para_update.c:39-62). */
int
ref_solve_batch (int ncells, const int *nsz, const int64_t *offA, const int64_t *offV, const double *A, const double *B,
                 double *X, int *codes, int nthreads)
{
  int maxn = 0;
  for (int c = 0; c < ncells; c++)
    if (nsz[c] > maxn)
      maxn = nsz[c];
  (void) nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    double *work = (double *) malloc (sizeof (double) * (size_t) maxn * (size_t) maxn);
#pragma omp for schedule(dynamic, 4)
    for (int c = 0; c < ncells; c++)
    {
      const int n = nsz[c];
      memcpy (work, A + offA[c], sizeof (double) * (size_t) n * n);
      codes[c] = ref_solve_matrix (work, B + offV[c], n, X + offV[c]);
    }
    free (work);
  }
  return 0;
}

int
ref_invert_batch (int ncells, const int *nsz, const int64_t *offA, const double *A, double *Inv, int *codes, int nthreads)
{
  int maxn = 0;
  for (int c = 0; c < ncells; c++)
    if (nsz[c] > maxn)
      maxn = nsz[c];
  (void) nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    double *work = (double *) malloc (sizeof (double) * (size_t) maxn * (size_t) maxn);
#pragma omp for schedule(dynamic, 1)
    for (int c = 0; c < ncells; c++)
    {
      const int n = nsz[c];
      memcpy (work, A + offA[c], sizeof (double) * (size_t) n * n);
      codes[c] = ref_invert_matrix (work, Inv + offA[c], n);
    }
    free (work);
  }
  return 0;
}
