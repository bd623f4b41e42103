/*
 * hipla.h -- C ABI of libhipla.so, the MI355X (gfx950) per-cell dense solve / inverse library.
 *
 * Plain C types only (pointers and sizes).  Host pointers where the reference passes host
 * arrays; device pointers (hipMalloc'd or torch tensors' data_ptr) in the batched API.
 */
#ifndef HIPLA_H
#define HIPLA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ *
 * Drop-in entry points.  Same names, signatures, argument meaning and return codes as the    *
 * reference (declared at source/templates.h:358-360, defined at source/matrix_cpu.c:273-349). *
 * ------------------------------------------------------------------------------------------ */

/* Replaces solve_matrix, source/matrix_cpu.c:309-321 (-> cpu_solve_matrix :96-204 or
 * gpu_solve_matrix, source/matrix_gpu.cu:263-307).  Solves A x = b for one row-major n x n FP64
 * matrix with GSL's partial-pivot LU operation order on the GPU.  Returns
 *   0  ok
 *   2  some row with b_i > 0 fails the relative residual check |Ax-b|_i/|Ax|_i > 1e-6 (x valid)
 *   3  some row with b_i <= 0 fails the absolute residual check |Ax-b|_i > 1e-6 (x valid)
 *   4  singular (an exact zero pivot); x is not written
 *   5  the GPU library failed (bad argument, no device, out of memory); x is not written
 * a_matrix and b_matrix are only read (the reference's CPU path overwrites a_matrix with the
 * LU factors; callers free it straight after, source/matrix_ion.c:284-285). */
int solve_matrix (double *a_matrix, double *b_matrix, int size, double *x_matrix, int nplasma);

/* Replaces invert_matrix, source/matrix_cpu.c:337-349 (-> cpu_invert_matrix :223-257 or
 * gpu_invert_matrix, source/matrix_gpu.cu:331-384).  Writes the row-major inverse.  Returns
 * 0, 1 (GSL_EDOM: singular, as gsl_linalg_LU_invx, GSL 2.6 linalg/lu.c:447-449) or 5 (GPU
 * library failure).  The output is untouched on error, as in the reference. */
int invert_matrix (double *matrix, double *inverted_matrix, int num_rows);

/* Replaces get_matrix_error_string, source/matrix_cpu.c:273-281 (strings :58-72). */
const char *get_matrix_error_string (int error_code);

/* Replace cusolver_create / cusolver_destroy (source/matrix_gpu.h:16-17, matrix_gpu.cu:124-169,
 * called at source/python.c:101-108,749-751).  hipla_init is optional: every entry point
 * initialises on first use.  Return 0 on success. */
int hipla_init (void);
int hipla_finalize (void);

/* Text of the last library-level failure in this process (code 5 above), or "". */
const char *hipla_last_error_detail (void);

/* Device resources the library has made in this process (no reference counterpart): the number
 * of device allocations, streams and events created so far, and the bytes of device memory it
 * holds now.  The drop-in entry points keep their staging buffers and a small cache of
 * per-order plans, so once every order a caller uses has been seen, solve_matrix and
 * invert_matrix allocate nothing and device_allocations stops growing.  Either pointer may be
 * NULL.  Returns 0. */
int hipla_alloc_stats (int64_t *device_allocations, int64_t *device_bytes_live);

/* No counterpart in the reference: solve_matrix for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * computation and codes as solve_matrix (FP64, GSL's order, codes 0/2/3/4 in *dcode; x not
 * written on 4), but on DEVICE pointers dA (row-major n x n), dB, dX (n) and dcode (one int),
 * asynchronous on `stream`, and with no allocation of any kind: the cell descriptors and the
 * factor store live in a static device array of 8.4 MB that each GPU holds until the process
 * ends.  1 <= n <= 1024, else 5.  Every call on a GPU uses that GPU's one array: calls that
 * overlap on one GPU (other streams, other threads) race; serialise them. */
int hipla_solve_matrix_noalloc (const double *dA, const double *dB, int n, double *dX, int *dcode, void *stream);

/* ------------------------------------------------------------------------------------------ *
 * Batched ragged API (new; the reference solves one matrix per call).  It batches the per-cell *
 * loops that call the path: wind_update's cell loop (source/wind_updates2d.c:89-153 ->        *
 * matrix_ion.c:276) and calc_all_matom_matrices (source/macro_accelerate.c:1052-1061 -> :315). *
 *                                                                                             *
 * Cell c is the row-major n[c] x n[c] FP64 matrix at A + offA[c] (in doubles) with its        *
 * right-hand side at B + offV[c] and solution at X + offV[c].  A plan is built once per batch  *
 * shape from HOST arrays n / offA / offV, then executed any number of times on DEVICE data.   *
 * ------------------------------------------------------------------------------------------ */

enum
{
  HIPLA_MODE_FP64 = 0,  /* GSL operation order in FP64: equal to the reference's results        */
  HIPLA_MODE_MIXED = 1  /* equilibrated FP32 LU + FP64 residual refinement + FP64 LU fallback     */
};

enum
{
  HIPLA_FLAG_GRAPH = 1,         /* capture the launch sequence in a hipGraph and replay it       */
  HIPLA_FLAG_KEEP_FACTORS = 2,  /* keep LU factors and pivots in the plan (for hipla_adjoint)    */
  HIPLA_FLAG_NO_FALLBACK = 4,   /* mixed mode: report stalled cells instead of re-solving them   */
  HIPLA_FLAG_PER_CELL = 8       /* plain per-matrix launches: one kernel launch per cell         */
};

/* Per-cell report of the mixed solver (device array of ncells records). */
typedef struct
{
  int32_t refine_steps; /* FP64 residual refinement steps taken (0..max_refine)                     */
  int32_t fallback;     /* 1 if the cell was re-solved by the FP64 LU after refinement stalled      */
  double rel_correction;/* ||dx||_inf / ||x||_inf of the last refinement step                      */
  double backward_err;  /* ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) at exit, in FP64     */
} hipla_cell_stats;

typedef struct hipla_plan_s *hipla_plan;

/* Build a plan for ncells cells; n, offA, offV are host arrays.  Returns 0 or 5. */
int hipla_plan_create (hipla_plan *plan, int ncells, const int *n, const int64_t *offA, const int64_t *offV,
                       int mode, int flags);
int hipla_plan_destroy (hipla_plan plan);

/* Mixed-mode knobs: refinement step cap (default 5) and the ||dx||/||x|| target (default 1e-12). */
int hipla_plan_set_refinement (hipla_plan plan, int max_refine, double tol);

/* Solve every cell on `stream` (a hipStream_t, or NULL for the default stream).  dA, dB, dX,
 * dcodes (ncells ints, solve_matrix's codes per cell) and dstats (may be NULL) are device
 * pointers.  Asynchronous: returns 0 once queued, or 5. */
int hipla_solve (hipla_plan plan, const double *dA, const double *dB, double *dX, int *dcodes,
                 hipla_cell_stats *dstats, void *stream);

/* Invert every cell (offV is ignored); dInv has A's layout and must not overlap dA.  dcodes gets
 * invert_matrix's codes per cell (0, or 1 = GSL_EDOM for an exact zero pivot); dstats is not
 * written.  Cells whose order is a multiple of 64 from 128 on and whose matrix is row diagonally
 * dominant (as N = I - Q of the macro-atom scheme is) are swept by a blocked Gauss-Jordan
 * inverse on the matrix cores: in HIPLA_MODE_FP64 with FP64 MFMA block products, in
 * HIPLA_MODE_MIXED with split BF16x3 MFMA products in FP32 followed by two FP64 residual
 * refinement steps.  Every other cell runs the GSL-order FP64 LU inverse with partial pivoting.
 * Environment: HIPLA_INVERT_SWEEP=0 sends every cell to the pivoting inverse; HIPLA_GJ_STEP=64
 * or 128 caps the FP64 sweep's step width (default 256 columns). 
 * All are within 1e-8 relative Frobenius of the reference's gsl_linalg_LU_invx on the
 * workloads tested; none is in its operation order. */
int hipla_invert (hipla_plan plan, const double *dA, double *dInv, int *dcodes, hipla_cell_stats *dstats,
                  void *stream);

/* Adjoint of the last hipla_solve on this plan (needs HIPLA_FLAG_KEEP_FACTORS, HIPLA_MODE_FP64):
 * solves A^T lambda = g per cell with the stored LU factors and writes dL/dA = -lambda x^T
 * (layout of A) into dGradA when it is not NULL.  dG, dLambda have B's layout.  A singular cell
 * (solve code 4) gets lambda = 0 and dL/dA = 0. */
int hipla_adjoint (hipla_plan plan, const double *dX, const double *dG, double *dLambda, double *dGradA,
                   void *stream);

/* The T_e gradient workload's rate model on the GPU (no counterpart in the reference, which finds
 * T_e by a root find around the solve, source/ionization.c:486).  ncells matrices of order
 * n <= 64, flat (cell c's A at c n^2, its b at c n).  Stage q joins rows r = dRow[q] and r + 1
 * (0 <= r < n - 1): A[r+1][r] = C T^0.5 exp(-E / kT), A[r][r+1] = a0 (T / 1e4)^-0.7, with C, E,
 * a0 of shape [ncells][nstage] and T per cell; each diagonal is minus its column sum; row
 * dConsS[e] is replaced by ones over the dConsM[e] ions from it and b there is dConsAb[e] (the
 * abundance rows of populate_ion_rate_matrix, source/matrix_ion.c:633-655); every other b is 0. */
int hipla_te_assemble (int ncells, int n, int nstage, const int *dRow, const double *dC, const double *dE,
                       const double *dA0, int ncons, const int *dConsS, const int *dConsM, const double *dConsAb,
                       const double *dT, double *dA, double *dB, void *stream);

/* dL/dT_e = -lambda^T (dA/dT) x per cell into dGradT[ncells], for the same model, from the
 * solution x and the adjoint's lambda (hipla_adjoint with dGradA NULL): the dense dL/dA is never
 * formed.  Abundance rows carry no T dependence. */
int hipla_te_grad (int ncells, int n, int nstage, const int *dRow, const double *dC, const double *dE,
                   const double *dA0, int ncons, const int *dConsS, const double *dT, const double *dX,
                   const double *dLambda, double *dGradT, void *stream);

/* Macro-atom emissivities without forming N^-1 (needs HIPLA_FLAG_KEEP_FACTORS, HIPLA_MODE_FP64).
 * Replaces invert_matrix + B = N^-1 R (source/macro_accelerate.c:315-334) + the contraction of
 * source/macro_gen_f.c:490-506 by one factorisation and one transposed solve per cell: with N the
 * macro-atom matrix (A's layout, order nlevels_macro + 1, the k-packet last) and, in B's layout,
 * rdiag = diag(R), abs = (matom_abs[0..], kpkt_abs), emit = (level_emit[0..], kpkt_emit):
 *   out_j = (N^-T abs)_j * rdiag_j * emit_j   (matom_emiss for j < n-1, kpkt_emiss for j = n-1).
 * dcodes gets solve_matrix's codes for N x = abs; a cell with code 4 (singular N) is not
 * written.  Equal to the reference's numbers up to rounding (not in its summation order). */
int hipla_macro_emiss (hipla_plan plan, const double *dN, const double *dRdiag, const double *dAbs,
                       const double *dEmit, double *dOut, int *dcodes, void *stream);

/* Binary ragged-batch files (host only; no counterpart in the reference, whose tests exchange
 * single matrices as 7-digit "%e" text, source/tests/test_matrix.c:52-62).  A batch is kept
 * bit-exact in the layout the plans take: a 64-byte header ("HIPLARB1", version 1, flags bit 0 =
 * B present, ncells, totA, totV), int32 n[ncells] padded to 8 bytes, int64 offA[ncells],
 * int64 offV[ncells], FP64 A[totA], then FP64 B[totV] if present; little-endian.
 * hipla_batch_write stores a batch (B may be NULL); hipla_batch_info reads the sizes (has_b may
 * be NULL); hipla_batch_read fills caller arrays of those sizes (B may be NULL to skip it).
 * Every cell's extent and the file size are checked.  0, or 5 with the reason in
 * hipla_last_error_detail. */
int hipla_batch_write (const char *path, int ncells, const int *n, const int64_t *offA, const int64_t *offV,
                       const double *A, const double *B);
int hipla_batch_info (const char *path, int *ncells, int64_t *totA, int64_t *totV, int *has_b);
int hipla_batch_read (const char *path, int *n, int64_t *offA, int64_t *offV, double *A, double *B);

/* ------------------------------------------------------------------------------------------ *
 * The ionization balance of many cells in lock step.  Restates, for a batch of cells at once,  *
 * the n_e fixed point of matrix_ion_populations (source/matrix_ion.c:219-382: DENSITY_MIN      *
 * floors, xnew, the FRACTIONAL_ERROR = 0.03 test, (xnew + xne) / 2 damping, MAXITERATIONS =    *
 * 200) with the rate matrix assembled on the GPU as populate_ion_rate_matrix does (:444-659)   *
 * and solved in FP64 in GSL's order (solve_matrix, :276).  In the reference the per-ion         *
 * coefficients (pi_rates, rr_rates, di/dr/qrecomb_coeffs, charge exchange, inner-shell rates   *
 * and yields) are computed from atomic data before the loop (:98-183) and are constants of the  *
 * cell inside it; here the caller supplies them.                                               *
 * ------------------------------------------------------------------------------------------ */
#define HIPLA_ION_MAXN 512            /* ions per cell                                          */
/* per-ion bits: istate (1 = neutral) in bits 0-7, z in bits 8-15, and: */
#define HIPLA_ION_FIRST (1 << 16)     /* first ion of its element (ele[].firstion)              */
#define HIPLA_ION_MAXST (1 << 17)     /* highest stage of its element (ele[].istate_max)        */
#define HIPLA_ION_DI (1 << 18)        /* has a collisional-ionization rate (dere_di_flag > 0)    */
#define HIPLA_ION_DR (1 << 19)        /* has a dielectronic-recombination rate (drflag > 0)      */
#define HIPLA_ION_CXI (1 << 20)       /* has a charge-exchange ionization rate (by H II)         */
#define HIPLA_ION_MAXJUMP 10          /* inner-shell electron yields per record (prob[0..9])      */

/* Device pointers; per-ion arrays in B's layout (cell c's ions at offV[c]).  elem is an element
 * id in 0 .. 2^30 - 1 that is unique within a cell (the reference compares ion[].z).  Inner-shell record k of
 * ion i (inner_beg[i] <= k < inner_beg[i] + inner_cnt[i], global record indices) ionizes i with
 * rate inner_rate[k] into the stages i + d, d = 1 .. min(z - istate + 1, 11) - 1, with yields
 * inner_yield[10 k + d - 1].  has_cx: n_charge_exchange > 0 (charge-exchange recombination by
 * H I on every ion above hydrogen, coefficient cxr). */
typedef struct
{
  const int32_t *elem, *bits;
  const double *pi, *rr, *di, *dr, *qr, *cxr, *cxi;
  const int32_t *inner_beg, *inner_cnt;
  const double *inner_rate, *inner_yield;
  int has_cx;
} hipla_ion_rates;

typedef struct hipla_ion_s *hipla_ion;

/* A context for ncells cells of n[c] ions (1 <= n[c] <= HIPLA_ION_MAXN); host arrays.  0 or 5. */
int hipla_ion_create (hipla_ion *ctx, int ncells, const int *n, const int64_t *offV);
int hipla_ion_destroy (hipla_ion ctx);

/* Run every cell's iteration to its end.  dDensity (B's layout) holds the absolute ion densities
 * to start from (xplasma->density) and receives the new ones for cells that converge; dNe
 * (ncells) gets get_ne of the new densities, dStatus matrix_ion_populations' return value (0, or
 * -1 for a singular solve or MAXITERATIONS, density then untouched), dIters the iteration count
 * at the end, dCodes the last solve_matrix code.  Synchronous: it reads the number of cells still
 * iterating back after every few iterations and returns once none is left.  0 or 5. */
int hipla_ion_populations (hipla_ion ctx, const hipla_ion_rates *rates, double *dDensity, double *dNe, int *dStatus,
                           int *dIters, int *dCodes, void *stream);

/* Time the kernel variants for this plan's size histogram and keep the fastest per size class:
 * in mixed plans FP32 LU + FP64 refinement against FP64 straight away, and in both modes the FP64
 * kernel shape (hipla_plan_set_variant) of every class solved in FP64.  The choices are cached per
 * process, keyed on (kernel family, largest order, cell count), and, if HIPLA_TUNE_CACHE names a
 * file, across processes.  Runs synchronously on `stream`.  dA..dX must be valid inputs as for
 * hipla_solve. */
int hipla_plan_autotune (hipla_plan plan, const double *dA, const double *dB, double *dX, int *dcodes,
                         void *stream);

/* The FP64 kernel shape of size-class group `group` (hipla_plan_describe lists the groups and
 * their "variant"), as hipla_plan_autotune would choose it: 0 the default; 1 the other shape of
 * the group's family (two waves per workgroup instead of four -- one instead of two for n <= 64 --
 * for the one-wave-per-matrix kernels, the blocked kernel instead of the register-tiled one for
 * 64 < n <= 175, 512 threads instead of 256 for the blocked kernel at n <= 256).  Every shape gives
 * the same values.  0, or 5 for a shape the group does not have. */
int hipla_plan_set_variant (hipla_plan plan, int group, int variant);

/* Profiling helper (no reference counterpart): launch each size-class group of the plan alone
 * `reps` times on the plan's own stream, bracketed by HIP events, and write the average ms per
 * launch into ms_out[0 .. ngroups-1].  With nout >= ngroups + 2 also ms_out[ngroups] = the FP64
 * re-solve pass of a mixed plan (0 for FP64 plans) and ms_out[ngroups+1] = a whole hipla_solve;
 * with nout == ngroups only the size classes are launched.  nout >= ngroups. */
int hipla_plan_time_groups (hipla_plan plan, const double *dA, const double *dB, double *dX, int *dcodes,
                            hipla_cell_stats *dstats, int reps, double *ms_out, int nout);

/* Describe the plan's schedule (size classes, kernel variant per class, tiles) into buf. */
int hipla_plan_describe (hipla_plan plan, char *buf, int buflen);

#ifdef __cplusplus
}
#endif

#endif /* HIPLA_H */
