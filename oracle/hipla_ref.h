/* oracle/hipla_ref.h -- CPU oracle (test infrastructure only; see hipla_ref.c). */
#ifndef HIPLA_REF_H
#define HIPLA_REF_H
#include <stddef.h>
#include <stdint.h>

int ref_lu_decomp (double *a, int n, size_t *perm, unsigned *ipiv, int *signum);
int ref_lu_decomp_rightlooking (double *a, int n, size_t *perm, unsigned *ipiv, int *signum);
int ref_lu_solve (const double *lu, int n, const size_t *perm, const double *b, double *x);
int ref_residual_code (const double *a_orig, int n, const double *b, const double *x);
int ref_solve_matrix (double *a, const double *b, int n, double *x);
int ref_lu_invx (double *lu, int n, const size_t *perm);
int ref_invert_matrix (double *m, double *inv, int n);
int ref_lu_solve_transpose (const double *lu, int n, const size_t *perm, const double *g, double *lam);
int ref_solve_dd_refined (const double *a, const double *b, int n, double *x, int iters);
int ref_solve_batch (int ncells, const int *nsz, const int64_t *offA, const int64_t *offV, const double *A,
                     const double *B, double *X, int *codes, int nthreads);
int ref_invert_batch (int ncells, const int *nsz, const int64_t *offA, const double *A, double *Inv, int *codes,
                      int nthreads);
#endif
