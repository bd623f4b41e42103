// k_te.hip -- the temperature-dependent ionization matrices of the T_e gradient workload,
// assembled on the GPU, and dL/dT_e from the adjoint's lambda without forming dL/dA.
//
// The rate model (the proxy for the rates calc_te's root find moves, source/ionization.c:486):
// stage q of an element joins row r = row[q] (lower ion) and r + 1:
//   ionization r -> r+1 at A[r+1][r] = C T^0.5 exp(-E / (k T))
//   recombination r+1 -> r at A[r][r+1] = a0 (T / 1e4)^-0.7
// each diagonal is minus its column's off-diagonal sum, and the first row s of every element is
// replaced by the abundance row (ones over the element's m ions, b[s] = ab), as
// populate_ion_rate_matrix does (source/matrix_ion.c:633-655).  A cell is one wave, lane t owning
// column t (n <= 64): every row is written as one coalesced 512-byte store.
//
// The gradient: with A x = b, A^T lambda = dL/dx and b independent of T,
//   dL/dT = -lambda^T (dA/dT) x,
// and dA/dT has four entries per stage (the two rates' derivatives and their diagonal
// counterparts), zero in abundance rows; one term per stage and lane, then a wave sum.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "hipla_dev.h"
#include "hipla_plan.h"

#define TE_KB_EV 8.617333262e-5     // Boltzmann constant, eV / K

struct TeArgs
{
  int ncells, n, nstage, ncons;
  const int *row;         // [nstage] lower row of each stage
  const double *C, *E, *a0;   // [ncells][nstage]
  const int *cons_s, *cons_m; // [ncons] first row and ion count of each element
  const double *cons_ab;  // [ncons]
  const double *T;        // [ncells]
};

// stage of lower row r on this lane (-1: none) and whether row r is an abundance row
__device__ __forceinline__ void te_row_info (const TeArgs &a, int r, int &stage, bool &cons)
{
  stage = -1;
  for (int q = 0; q < a.nstage; q++)
    if (a.row[q] == r)
      stage = q;
  cons = false;
  for (int e = 0; e < a.ncons; e++)
    if (a.cons_s[e] == r)
      cons = true;
}

__global__ __launch_bounds__ (64) void k_te_assemble (TeArgs a, double *A, double *B)
{
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int n = a.n;
  const double T = a.T[c];
  int q;
  bool cons_t;
  te_row_info (a, t, q, cons_t);
  // the rates of the stage whose lower row is t
  double ion = 0.0, rec = 0.0;
  if (t < n && q >= 0)
  {
    const size_t k = (size_t) c * a.nstage + q;
    ion = a.C[k] * sqrt (T) * exp (-a.E[k] / (TE_KB_EV * T));
    rec = a.a0[k] * pow (T / 1e4, -0.7);
  }
  // column t: ion at row t+1 (this lane's stage), rec at row t-1 (the previous lane's stage)
  const double rec_up = __shfl_up (rec, 1);
  const double rec_col = t > 0 ? rec_up : 0.0;
  const double diag = -(rec_col + ion);
  double *Ac = A + (size_t) c * n * n;
  for (int i = 0; i < n; i++)
  {
    // row i: an abundance row holds ones over its element's ions, every other row the rates
    int es = -1, em = 0;
    for (int e = 0; e < a.ncons; e++)
      if (a.cons_s[e] == i)
      {
        es = a.cons_s[e];
        em = a.cons_m[e];
      }
    double v;
    if (es >= 0)
      v = (t >= es && t < es + em) ? 1.0 : 0.0;
    else
      v = (i == t + 1) ? ion : (i == t - 1) ? rec_col : (i == t) ? diag : 0.0;
    if (t < n)
      Ac[(size_t) i * n + t] = v;
  }
  if (t < n)
  {
    double bv = 0.0;
    for (int e = 0; e < a.ncons; e++)
      if (a.cons_s[e] == t)
        bv = a.cons_ab[e];
    B[(size_t) c * n + t] = bv;
  }
}

__global__ __launch_bounds__ (64) void k_te_grad (TeArgs a, const double *X, const double *Lam, double *G)
{
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int n = a.n;
  const double T = a.T[c];
  int q;
  bool cons_t;
  te_row_info (a, t, q, cons_t);
  const double *x = X + (size_t) c * n;
  const double *lam = Lam + (size_t) c * n;
  double term = 0.0;
  if (t < n && q >= 0 && t + 1 < n)
  {
    const size_t k = (size_t) c * a.nstage + q;
    const double xe = a.E[k] / (TE_KB_EV * T);
    const double ion = a.C[k] * sqrt (T) * exp (-xe);
    const double rec = a.a0[k] * pow (T / 1e4, -0.7);
    const double dion = ion * (0.5 / T + xe / T);
    const double drec = rec * (-0.7 / T);
    int q1;
    bool cons_t1;
    te_row_info (a, t + 1, q1, cons_t1);
    // abundance rows carry no rate terms
    const double lr = cons_t ? 0.0 : lam[t];
    const double lr1 = cons_t1 ? 0.0 : lam[t + 1];
    // dA(t+1, t) = dion, dA(t, t) -= dion, dA(t, t+1) = drec, dA(t+1, t+1) -= drec
    term = dion * (lr1 - lr) * x[t] + drec * (lr - lr1) * x[t + 1];
  }
  // wave sum
  for (int o = 32; o > 0; o >>= 1)
    term += __shfl_xor (term, o);
  if (t == 0)
    G[c] = -term;
}

static bool te_args_ok (const TeArgs &a, bool need_rows, const char *who)
{
  if (a.ncells < 0 || a.n < 1 || a.n > 64 || a.nstage < 0 || a.ncons < 0 || !a.T ||
      (a.nstage > 0 && (!a.row || !a.C || !a.E || !a.a0)) || (a.ncons > 0 && !a.cons_s) ||
      (need_rows && a.ncons > 0 && (!a.cons_m || !a.cons_ab)))
  {
    hipla_set_detail ("%s: bad argument (n must be 1..64)", who);
    return false;
  }
  return true;
}

extern "C" int hipla_te_assemble (int ncells, int n, int nstage, const int *dRow, const double *dC, const double *dE,
                                  const double *dA0, int ncons, const int *dConsS, const int *dConsM,
                                  const double *dConsAb, const double *dT, double *dA, double *dB, void *stream)
{
  const TeArgs a{ ncells, n, nstage, ncons, dRow, dC, dE, dA0, dConsS, dConsM, dConsAb, dT };
  if (!te_args_ok (a, true, "hipla_te_assemble"))
    return 5;
  if (!dA || !dB)
  {
    hipla_set_detail ("hipla_te_assemble: bad argument (A, b)");
    return 5;
  }
  if (ncells == 0)
    return 0;
  hipLaunchKernelGGL (k_te_assemble, dim3 (ncells), dim3 (64), 0, (hipStream_t) stream, a, dA, dB);
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}

extern "C" int hipla_te_grad (int ncells, int n, int nstage, const int *dRow, const double *dC, const double *dE,
                              const double *dA0, int ncons, const int *dConsS, const double *dT, const double *dX,
                              const double *dLambda, double *dGradT, void *stream)
{
  const TeArgs a{ ncells, n, nstage, ncons, dRow, dC, dE, dA0, dConsS, nullptr, nullptr, dT };
  if (!te_args_ok (a, false, "hipla_te_grad"))
    return 5;
  if (!dX || !dLambda || !dGradT)
  {
    hipla_set_detail ("hipla_te_grad: bad argument (x, lambda, gradient)");
    return 5;
  }
  if (ncells == 0)
    return 0;
  hipLaunchKernelGGL (k_te_grad, dim3 (ncells), dim3 (64), 0, (hipStream_t) stream, a, dX, dLambda, dGradT);
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}
