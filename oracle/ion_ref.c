/*
 * oracle/ion_ref.c -- CPU oracle of the ionization-balance iteration (TEST INFRASTRUCTURE ONLY).
 *
 * The checker for the batched n_e iteration: tests/ and the cpu_baseline leg of bench.py call it;
 * the product library never links, loads or calls it.  It restates, loop for loop and in the
 * reference's operation order (build with -ffp-contract=off):
 *   source/matrix_ion.c:444-659  populate_ion_rate_matrix  (rate matrix and b at xne, nh1, nh2)
 *   source/matrix_ion.c:219-382  the n_e fixed point of matrix_ion_populations (DENSITY_MIN
 *                                floors, xnew, the FRACTIONAL_ERROR test, (xnew + xne) / 2
 *                                damping, MAXITERATIONS) around cpu_solve_matrix
 *   source/matrix_ion.c:84-96,129  element densities and fractional populations
 *   source/saha.c:821-832         get_ne
 * with the per-ion coefficients (pi_rates, rr_rates, di/dr/qrecomb_coeffs, charge exchange,
 * inner-shell rates and electron yields) as inputs instead of the atomic-data physics that
 * computes them (source/matrix_ion.c:98-183, compute_*_coeffs): inside the iteration they are
 * constants of the cell.
 *
 * Element membership uses a per-ion element id where the reference compares ion[].z: the two
 * are the same when each element appears once, and the id keeps synthetic layouts that repeat
 * an element's stage count apart.  z itself still gives the inner-shell electron count and the
 * hydrogen tests, as in the reference.  At most one charge-exchange ionization record per ion.
 *
 * Parity is pinned in tests/test_oracle_ion.py: detailed balance along element chains
 * (ionization over recombination, charge exchange included), R x = 0 for a rate matrix built
 * independently from a transition list (inner-shell Auger jumps included), a hand-written 3-ion
 * matrix for the dielectronic-recombination flag, the hydrogen fixed point in closed form, the
 * iteration count of the damping sequence in closed form, and the singular path.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hipla_ref.h"

/* source/atomic.h:9, source/python.h:1211-1212 */
#define REF_DENSITY_MIN 1.e-20
#define REF_MAXITERATIONS 200
#define REF_FRACTIONAL_ERROR 0.03

static int
inner_nelec (int bits)
{
  /* source/matrix_ion.c:605-607 */
  int n_elec = REF_ION_Z (bits) - REF_ION_ISTATE (bits) + 1;
  if (n_elec > 11)
    n_elec = 11;
  return n_elec;
}

/* populate_ion_rate_matrix, source/matrix_ion.c:444-659, for one cell of n ions.  Each process
 * is its own pass over the ions, as in the reference, so every entry receives its terms in the
 * reference's order (the reference's double loops over (mm, nn) only act on mm = nn +- 1). */
void
ref_populate_ion_rate_matrix (int n, const ref_ion_cell *t, double xne, double nh1, double nh2, double *a, double *b)
{
  const int *el = t->elem, *bits = t->bits;
  memset (a, 0, sizeof (double) * (size_t) n * (size_t) n);
#define A_(r, c) a[(size_t) (r) * (size_t) n + (size_t) (c)]
  /* photoionization out of / into a stage (:474-494) */
  for (int mm = 0; mm < n; mm++)
    if (!(bits[mm] & REF_ION_MAXST))
      A_ (mm, mm) -= t->pi[mm];
  for (int nn = 0; nn + 1 < n; nn++)
    if (!(bits[nn] & REF_ION_MAXST) && el[nn + 1] == el[nn])
      A_ (nn + 1, nn) += t->pi[nn];
  /* direct (collisional) ionization, xne-dependent (:498-518) */
  for (int mm = 0; mm < n; mm++)
    if (!(bits[mm] & REF_ION_MAXST) && (bits[mm] & REF_ION_DI))
      A_ (mm, mm) -= (xne * t->di[mm]);
  for (int nn = 0; nn + 1 < n; nn++)
    if (!(bits[nn] & REF_ION_MAXST) && el[nn + 1] == el[nn] && (bits[nn] & REF_ION_DI))
      A_ (nn + 1, nn) += (xne * t->di[nn]);
  /* radiative + three-body recombination (:523-541) */
  for (int mm = 0; mm < n; mm++)
    if (!(bits[mm] & REF_ION_FIRST))
      A_ (mm, mm) -= xne * (t->rr[mm] + xne * t->qr[mm]);
  for (int nn = 1; nn < n; nn++)
    if (!(bits[nn] & REF_ION_FIRST) && el[nn - 1] == el[nn])
      A_ (nn - 1, nn) += xne * (t->rr[nn] + xne * t->qr[nn]);
  /* dielectronic recombination (:545-567): the populating term tests the flag of the lower ion
   * (ion[mm], mm = nn - 1), the depopulating one that of the ion itself */
  for (int mm = 0; mm < n; mm++)
    if (!(bits[mm] & REF_ION_FIRST) && (bits[mm] & REF_ION_DR))
      A_ (mm, mm) -= (xne * t->dr[mm]);
  for (int nn = 1; nn < n; nn++)
    if (!(bits[nn] & REF_ION_FIRST) && el[nn - 1] == el[nn] && (bits[nn - 1] & REF_ION_DR))
      A_ (nn - 1, nn) += (xne * t->dr[nn]);
  /* charge-exchange recombination with neutral hydrogen, every ion above hydrogen (:571-581);
   * the reference's ion 0 is hydrogen, so mm - 1 >= 0 there */
  if (t->has_cx)
    for (int mm = 1; mm < n; mm++)
      if (REF_ION_Z (bits[mm]) != 1 && REF_ION_ISTATE (bits[mm]) != 0)
      {
        A_ (mm, mm) -= t->cxr[mm] * nh1;
        A_ (mm - 1, mm) += t->cxr[mm] * nh1;
      }
  /* charge-exchange ionization by ionized hydrogen (:585-591), one record per ion */
  for (int mm = 0; mm + 1 < n; mm++)
    if (bits[mm] & REF_ION_CXI)
    {
      A_ (mm, mm) -= t->cxi[mm] * nh2;
      A_ (mm + 1, mm) += t->cxi[mm] * nh2;
    }
  /* inner-shell ionization with Auger electron yields (:594-617), records in order */
  for (int mm = 0; mm < n; mm++)
    for (int r = t->inner_beg[mm]; r < t->inner_beg[mm] + t->inner_cnt[mm]; r++)
    {
      const double rate = t->inner_rate[r];
      A_ (mm, mm) -= rate;
      const int n_elec = inner_nelec (bits[mm]);
      /* the reference relies on its atomic data to keep mm + d inside the matrix */
      for (int d = 1; d < n_elec && mm + d < n; d++)
        A_ (mm + d, mm) += rate * t->inner_yield[(size_t) r * REF_ION_MAXJUMP + (size_t) (d - 1)];
    }
  /* the first row of each element becomes the abundance constraint (:633-655) */
  for (int nn = 0; nn < n; nn++)
  {
    if (bits[nn] & REF_ION_FIRST)
    {
      b[nn] = 1.0;
      for (int mm = 0; mm < n; mm++)
        A_ (nn, mm) = (el[mm] == el[nn]) ? 1.0 : 0.0;
    }
    else
      b[nn] = 0.0;
  }
#undef A_
}

/* get_ne, source/saha.c:821-832 */
static double
ref_get_ne (int n, const int *bits, const double *density)
{
  double ne = 0;
  for (int k = 0; k < n; k++)
    ne += density[k] * (REF_ION_ISTATE (bits[k]) - 1);
  return ne;
}

/* The n_e iteration of matrix_ion_populations, source/matrix_ion.c:219-382, for one cell whose
 * coefficients t gives.  density (absolute ion densities, n) is updated on success, as
 * xplasma->density is (:366-377); *ne_out = get_ne of the new densities (:378), *iters = the
 * iteration count at the break, *last_code = the last solve_matrix code.  Returns 0, or -1
 * (a singular solve, :287-291, or MAXITERATIONS, :363-381), leaving density untouched. */
int
ref_ion_populations (int n, const ref_ion_cell *t, double *density, double *ne_out, int *iters, int *last_code)
{
  const int *el = t->elem, *bits = t->bits;
  double *elem_dens = (double *) calloc ((size_t) n, sizeof (double));
  double *newden = (double *) malloc (sizeof (double) * (size_t) n);
  double *a = (double *) malloc (sizeof (double) * (size_t) n * (size_t) n);
  double *bb = (double *) malloc (sizeof (double) * (size_t) n);
  double *pop = (double *) malloc (sizeof (double) * (size_t) n);
  int status = 0, niterate = 0;
  *last_code = 0;
  *iters = 0;
  if (!elem_dens || !newden || !a || !bb || !pop)
  {
    status = -2;
    goto done;
  }
  for (int mm = 0; mm < n; mm++)
    if (el[mm] < 0 || el[mm] >= n)
    {
      status = -2;
      goto done;
    }
  /* element densities and fractional populations (:84-96, :129) */
  for (int mm = 0; mm < n; mm++)
    elem_dens[el[mm]] = elem_dens[el[mm]] + density[mm];
  for (int mm = 0; mm < n; mm++)
    newden[mm] = density[mm] / elem_dens[el[mm]];
  double nh1 = 0, nh2 = 0;
  double xne = ref_get_ne (n, bits, density);   /* :208 */
  double xnew = 0.0;
  while (niterate < REF_MAXITERATIONS)
  {
    /* neutral / ionized hydrogen densities for charge exchange (:231-241) */
    for (int nn = 0; nn < n; nn++)
    {
      if (REF_ION_Z (bits[nn]) == 1 && REF_ION_ISTATE (bits[nn]) == 1)
        nh1 = newden[nn] * elem_dens[el[nn]];
      if (REF_ION_Z (bits[nn]) == 1 && REF_ION_ISTATE (bits[nn]) == 2)
        nh2 = newden[nn] * elem_dens[el[nn]];
    }
    ref_populate_ion_rate_matrix (n, t, xne, nh1, nh2, a, bb);
    const int err = ref_solve_matrix (a, bb, n, pop);   /* :276 */
    *last_code = err;
    if (err == 4)
    {
      status = -1;
      goto done;
    }
    /* new fractional populations with the DENSITY_MIN floor (:307-336) */
    for (int nn = 0; nn < n; nn++)
    {
      newden[nn] = pop[nn];
      if (newden[nn] < REF_DENSITY_MIN)
        newden[nn] = REF_DENSITY_MIN;
    }
    /* the electron density they imply (:341-349) */
    xnew = 0.0;
    for (int nn = 0; nn < n; nn++)
      xnew += newden[nn] * (REF_ION_ISTATE (bits[nn]) - 1) * elem_dens[el[nn]];
    if (xnew < REF_DENSITY_MIN)
      xnew = REF_DENSITY_MIN;
    if (fabs ((xne - xnew) / (xnew)) < REF_FRACTIONAL_ERROR || xnew < 1.e-6)   /* :352 */
      break;
    xne = (xnew + xne) / 2.;   /* :357 */
    niterate++;
    if (niterate == REF_MAXITERATIONS)
    {
      status = -1;
      goto done;
    }
  }
  /* absolute densities and n_e from them (:366-378) */
  for (int nn = 0; nn < n; nn++)
    density[nn] = newden[nn] * elem_dens[el[nn]];
  *ne_out = ref_get_ne (n, bits, density);
done:
  *iters = niterate;
  free (elem_dens);
  free (newden);
  free (a);
  free (bb);
  free (pop);
  return status;
}

/* A ragged batch of cells (ions of cell c at offV[c], inner-shell records indexed globally), one
 * worker per core, each running the serial per-cell iteration as the reference's MPI ranks do
 * over cells (source/para_update.c:39-62).  status/ne/iters/codes per cell. */
int
ref_ion_populations_batch (int ncells, const int *nsz, const int64_t *offV, const ref_ion_cell *all, double *density,
                           double *ne, int *status, int *iters, int *codes, int nthreads)
{
  (void) nthreads;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 2)
  for (int c = 0; c < ncells; c++)
  {
    const int64_t o = offV[c];
    ref_ion_cell t = *all;
    t.elem += o;
    t.bits += o;
    t.pi += o;
    t.rr += o;
    t.di += o;
    t.dr += o;
    t.qr += o;
    t.cxr += o;
    t.cxi += o;
    t.inner_beg += o;
    t.inner_cnt += o;
    status[c] = ref_ion_populations (nsz[c], &t, density + o, ne + c, iters + c, codes + c);
  }
  return 0;
}
