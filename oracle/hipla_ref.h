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
/* ionization-balance iteration (ion_ref.c).  Per-ion bits: istate in bits 0-7, z in bits 8-15,
 * then the flags below. */
#define REF_ION_ISTATE(b) ((b) & 0xff)
#define REF_ION_Z(b) (((b) >> 8) & 0xff)
#define REF_ION_FIRST (1 << 16)   /* first ion of its element (ele[].firstion)            */
#define REF_ION_MAXST (1 << 17)   /* highest stage of its element (ele[].istate_max)      */
#define REF_ION_DI (1 << 18)      /* ion[].dere_di_flag > 0                                */
#define REF_ION_DR (1 << 19)      /* ion[].drflag > 0                                      */
#define REF_ION_CXI (1 << 20)     /* a charge-exchange ionization record (cxi) for the ion */
#define REF_ION_MAXJUMP 10        /* inner-shell yields per record: prob[0..9]             */
typedef struct
{
  const int *elem, *bits;                            /* per ion                            */
  const double *pi, *rr, *di, *dr, *qr, *cxr, *cxi;  /* per ion                            */
  const int *inner_beg, *inner_cnt;                  /* per ion: records [beg, beg + cnt)  */
  const double *inner_rate, *inner_yield;            /* per record (10 yields per record)  */
  int has_cx;                                        /* n_charge_exchange > 0              */
} ref_ion_cell;
void ref_populate_ion_rate_matrix (int n, const ref_ion_cell *t, double xne, double nh1, double nh2, double *a,
                                   double *b);
int ref_ion_populations (int n, const ref_ion_cell *t, double *density, double *ne_out, int *iters, int *last_code);
int ref_ion_populations_batch (int ncells, const int *nsz, const int64_t *offV, const ref_ion_cell *all,
                               double *density, double *ne, int *status, int *iters, int *codes, int nthreads);
#endif
