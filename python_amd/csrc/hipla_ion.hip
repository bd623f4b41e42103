// hipla_ion.hip -- the ionization balance of every cell in lock step: the n_e fixed point of
// matrix_ion_populations (source/matrix_ion.c:219-382) with the rate matrix assembled on the GPU
// (populate_ion_rate_matrix, :444-659) and only the unconverged cells solved in each iteration.
//
// The reference runs the fixed point one cell at a time and, under CUDA, pays a device
// allocation and a host<->device copy of the whole matrix per iteration (source/matrix_gpu.cu:
// 263-307).  Here the per-ion coefficients stay in HBM; one iteration is four launches over all
// cells plus the FP64 GSL-order solves of the cells still iterating (one branch per kernel route,
// side by side):
//   k_ion_reset     clears the per-route cell lists and the active count (a kernel, so the
//                   iteration can live in a hipGraph)
//   k_ion_assemble  one workgroup per active cell: nh1 / nh2 from the current fractions, the
//                   twelve possibly non-zero entries of every column formed in the reference's
//                   order into LDS, then the n x n matrix and b written row by row (coalesced);
//                   the cell is appended to the list of the FP64 kernel its order routes to
//   solves          the FP64 kernels over those lists (device-side counts, persistent grids)
//   k_ion_update    one wave per active cell: DENSITY_MIN floors, xnew, the FRACTIONAL_ERROR
//                   test, the (xnew + xne) / 2 damping, MAXITERATIONS, and on convergence the
//                   absolute densities and get_ne (source/saha.c:821-832)
// Several iterations are captured in one graph; the host reads the active count after each
// replay and stops when no cell is left.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "hipla_kernels.h"
#include "hipla_plan.h"

#define ION_ISTATE(b) ((b) & 0xff)
#define ION_Z(b) (((b) >> 8) & 0xff)
#define ION_FIRST (1 << 16)
#define ION_MAXST (1 << 17)
#define ION_DI (1 << 18)
#define ION_DR (1 << 19)
#define ION_CXI (1 << 20)
#define ION_MAXJUMP 10
#define ION_NCOL (ION_MAXJUMP + 2)   // per column: (j-1, j), (j, j), (j+1 .. j+10, j)
#define ION_RFIRST (1 << 30)         // assembly's row flags: an element's abundance row
#define ION_DENSITY_MIN 1.e-20       // source/atomic.h:9
#define ION_MAXITERATIONS 200        // source/python.h:1211
#define ION_FRACTIONAL_ERROR 0.03    // source/python.h:1212
#define ION_ASM_THREADS 256
#define ION_CHUNK 4                  // iterations per graph replay
#define ION_ROUTES 5                 // FP64 kernels: wave 16 / 32 / 64, workgroup tile, blocked

struct IonArgs
{
  const int *n;
  const int64_t *offA, *offV;
  int ncells;
  hipla_ion_rates r;
  double *A, *B;
  const double *X;
  const int *codes;
  const double *density_in;
  double *density;   // out (written on convergence); may alias density_in
  double *ne;
  int *status, *iters, *codes_out;
  double *xne, *elem_dens, *newden;
  int *iter, *active;
  const int *route;
  int *lists;        // ION_ROUTES lists of [count, cells...], stride listcap + 1
  int listcap;
  int *nactive;
};

// ---- initial state: element densities, fractions, get_ne (source/matrix_ion.c:84-96,129,208) ----
// The ordered sums run on one lane over values staged in LDS (formed in parallel, each rounded
// exactly as the reference rounds it), so only the additions are serial.
__global__ __launch_bounds__ (64) void k_ion_init (IonArgs s)
{
  __shared__ int s_el[HIPLA_ION_MAXN];
  __shared__ double s_d[HIPLA_ION_MAXN];
  const int lane = threadIdx.x;
  for (int cell = blockIdx.x; cell < s.ncells; cell += gridDim.x)
  {
    const int n = s.n[cell];
    const int64_t o = s.offV[cell];
    const double *d = s.density_in + o;
    const int *bt = s.r.bits + o;
    for (int i = lane; i < n; i += 64)
    {
      s_el[i] = s.r.elem[o + i];
      s_d[i] = d[i];
    }
    __syncthreads ();
    for (int i = lane; i < n; i += 64)
    {
      // elem_dens[z] accumulated over the element's ions in ascending order
      const int e = s_el[i];
      double ed = 0.0;
      for (int j = 0; j < n; j++)
        if (s_el[j] == e)
          ed = ed + s_d[j];
      s.elem_dens[o + i] = ed;
      s.newden[o + i] = s_d[i] / ed;
    }
    __syncthreads ();
    for (int i = lane; i < n; i += 64)
      s_d[i] = s_d[i] * (ION_ISTATE (bt[i]) - 1);
    __syncthreads ();
    if (lane == 0)
    {
      double ne = 0;
      for (int k = 0; k < n; k++)
        ne += s_d[k];
      s.xne[cell] = ne;
      s.iter[cell] = 0;
      s.active[cell] = 1;
      s.status[cell] = 0;
      s.iters[cell] = 0;
      s.codes_out[cell] = 0;
    }
    __syncthreads ();
  }
}

__global__ void k_ion_reset (IonArgs s)
{
  if (threadIdx.x < ION_ROUTES)
    s.lists[threadIdx.x * (s.listcap + 1)] = 0;
  if (threadIdx.x == 0)
    *s.nactive = 0;
}

// ---- populate_ion_rate_matrix (source/matrix_ion.c:444-659) for every active cell ----
__global__ __launch_bounds__ (ION_ASM_THREADS) void k_ion_assemble (IonArgs s)
{
  extern __shared__ double smem[];
  __shared__ int s_h1, s_h2;
  const int tid = threadIdx.x;
  for (int cell = blockIdx.x; cell < s.ncells; cell += gridDim.x)
  {
    if (!s.active[cell])
      continue;
    const int n = s.n[cell];
    const int64_t o = s.offV[cell];
    // per row: element | ION_RFIRST on abundance rows, padded to whole groups of 16 rows (first:
    // 16-byte aligned for the grouped reads); then the columns' rates, elements and flags
    int *rinfo = reinterpret_cast<int *> (smem);
    double *colv = smem + ((n + 15) & ~15) / 2;                 // [n][ION_NCOL]
    int *el = reinterpret_cast<int *> (colv + (size_t) n * ION_NCOL);
    int *bt = el + n;
    if (tid == 0)
    {
      s_h1 = -1;
      s_h2 = -1;
    }
    __syncthreads ();
    for (int i = tid; i < n; i += ION_ASM_THREADS)
    {
      const int b = s.r.bits[o + i];
      const int e = s.r.elem[o + i];
      el[i] = e;
      bt[i] = b;
      // the flag bit set on abundance rows only (cleared otherwise, so no other row can match
      // an element id below 2^30 with the flag)
      rinfo[i] = (b & ION_FIRST) ? (e | ION_RFIRST) : (e & ~ION_RFIRST);
      // the reference's loop over ions keeps the last H I / H II it meets (:231-241)
      if (ION_Z (b) == 1 && ION_ISTATE (b) == 1)
        atomicMax (&s_h1, i);
      if (ION_Z (b) == 1 && ION_ISTATE (b) == 2)
        atomicMax (&s_h2, i);
    }
    __syncthreads ();
    const double xne = s.xne[cell];
    const double nh1 = s_h1 >= 0 ? s.newden[o + s_h1] * s.elem_dens[o + s_h1] : 0.0;
    const double nh2 = s_h2 >= 0 ? s.newden[o + s_h2] * s.elem_dens[o + s_h2] : 0.0;
    const hipla_ion_rates &r = s.r;
    for (int j = tid; j < n; j += ION_ASM_THREADS)
    {
      const int b = bt[j], e = el[j];
      const int64_t q = o + j;
      const bool first = (b & ION_FIRST) != 0, top = (b & ION_MAXST) != 0;
      const bool cxr_on = r.has_cx && j >= 1 && ION_Z (b) != 1 && ION_ISTATE (b) != 0;
      const bool cxi_on = (b & ION_CXI) && j + 1 < n;
      double *cv = colv + (size_t) j * ION_NCOL;
      // (j, j): every process out of ion j, in the order of the reference's passes
      double dg = 0.0;
      if (!top)
        dg -= r.pi[q];
      if (!top && (b & ION_DI))
        dg -= (xne * r.di[q]);
      if (!first)
        dg -= xne * (r.rr[q] + xne * r.qr[q]);
      if (!first && (b & ION_DR))
        dg -= (xne * r.dr[q]);
      if (cxr_on)
        dg -= r.cxr[q] * nh1;
      if (cxi_on)
        dg -= r.cxi[q] * nh2;
      const int rb = r.inner_beg[q], rc = r.inner_cnt[q];
      for (int k = rb; k < rb + rc; k++)
        dg -= r.inner_rate[k];
      cv[1] = dg;
      // (j-1, j): recombination of ion j into the stage below, and charge exchange with H I
      double up = 0.0;
      if (j >= 1)
      {
        const bool same = el[j - 1] == e;
        if (!first && same)
          up += xne * (r.rr[q] + xne * r.qr[q]);
        if (!first && same && (bt[j - 1] & ION_DR))   // the lower ion's flag (:563)
          up += (xne * r.dr[q]);
        if (cxr_on)
          up += r.cxr[q] * nh1;
      }
      cv[0] = up;
      // (j+1, j): ionization into the next stage
      double dn = 0.0;
      if (j + 1 < n)
      {
        const bool same = el[j + 1] == e;
        if (!top && same)
          dn += r.pi[q];
        if (!top && same && (b & ION_DI))
          dn += (xne * r.di[q]);
        if (cxi_on)
          dn += r.cxi[q] * nh2;
      }
      cv[2] = dn;
      for (int d = 2; d <= ION_MAXJUMP; d++)
        cv[1 + d] = 0.0;
      // inner-shell ionization with Auger yields, records in order (:594-617)
      int n_elec = ION_Z (b) - ION_ISTATE (b) + 1;
      if (n_elec > 11)
        n_elec = 11;
      for (int k = rb; k < rb + rc; k++)
      {
        const double rate = r.inner_rate[k];
        for (int d = 1; d < n_elec; d++)
          if (j + d < n)
            cv[1 + d] += rate * r.inner_yield[(int64_t) k * ION_MAXJUMP + (d - 1)];
      }
    }
    __syncthreads ();
    // write A and b; the first row of each element is its abundance constraint.  A thread writes
    // its column j: every row gets 0, or 1 on an abundance row of j's element (the row flags read
    // 16 rows at a time, one LDS round trip per 16 stores); then the rows j-1 .. j+ION_MAXJUMP
    // that are not abundance rows get the column's rates (a later store of the same thread to the
    // same address, which wins).  Each row's stores are coalesced across the workgroup.
    double *A = s.A + s.offA[cell];
    for (int j = tid; j < n; j += ION_ASM_THREADS)
    {
      const int ej = el[j];
      double cvj[ION_NCOL];
#pragma unroll
      for (int k = 0; k < ION_NCOL; k++)
        cvj[k] = colv[(size_t) j * ION_NCOL + k];
      double *Aj = A + j;
      for (int i0 = 0; i0 < n; i0 += 16)
      {
        int ri[16];
#pragma unroll
        for (int q = 0; q < 16; q += 4)
        {
          const int4 t = *reinterpret_cast<const int4 *> (rinfo + i0 + q);
          ri[q] = t.x;
          ri[q + 1] = t.y;
          ri[q + 2] = t.z;
          ri[q + 3] = t.w;
        }
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (i0 + q < n)
            Aj[(int64_t) (i0 + q) * n] = (ri[q] == (ej | ION_RFIRST)) ? 1.0 : 0.0;
      }
#pragma unroll
      for (int k = 0; k < ION_NCOL; k++)
      {
        const int i = j + k - 1;
        if (i >= 0 && i < n && !(rinfo[i] & ION_RFIRST))
          Aj[(int64_t) i * n] = cvj[k];
      }
    }
    for (int i = tid; i < n; i += ION_ASM_THREADS)
      s.B[o + i] = (bt[i] & ION_FIRST) ? 1.0 : 0.0;
    if (tid == 0)
    {
      int *lst = s.lists + (int64_t) s.route[cell] * (s.listcap + 1);
      const int pos = atomicAdd (lst, 1);
      lst[1 + pos] = cell;
    }
    __syncthreads ();
  }
}

// ---- the fixed-point step of matrix_ion_populations (source/matrix_ion.c:276-381) ----
__global__ __launch_bounds__ (64) void k_ion_update (IonArgs s)
{
  __shared__ double s_p[HIPLA_ION_MAXN];
  __shared__ int s_out;
  const int lane = threadIdx.x;
  for (int cell = blockIdx.x; cell < s.ncells; cell += gridDim.x)
  {
    if (!s.active[cell])
      continue;
    const int n = s.n[cell];
    const int64_t o = s.offV[cell];
    const int code = s.codes[cell];
    const int *bt = s.r.bits + o;
    const double *x = s.X + o;
    const double *ed = s.elem_dens + o;
    if (code == 4)
    {
      // solve_matrix failed: the cell returns -1 with its densities untouched (:287-291)
      if (lane == 0)
      {
        s.codes_out[cell] = code;
        s.status[cell] = -1;
        s.iters[cell] = s.iter[cell];
        s.active[cell] = 0;
      }
      continue;
    }
    // new fractions with the DENSITY_MIN floor (:307-336) and the terms of xnew (:341-346)
    for (int i = lane; i < n; i += 64)
    {
      double v = x[i];
      if (v < ION_DENSITY_MIN)
        v = ION_DENSITY_MIN;
      s.newden[o + i] = v;
      s_p[i] = v * (ION_ISTATE (bt[i]) - 1) * ed[i];
    }
    __syncthreads ();
    if (lane == 0)
    {
      // xnew summed in the reference's order
      double xnew = 0.0;
      for (int k = 0; k < n; k++)
        xnew += s_p[k];
      if (xnew < ION_DENSITY_MIN)
        xnew = ION_DENSITY_MIN;
      const double xne = s.xne[cell];
      s.codes_out[cell] = code;
      int outcome;   // 0 continue, 1 converged, 2 MAXITERATIONS
      if (fabs ((xne - xnew) / (xnew)) < ION_FRACTIONAL_ERROR || xnew < 1.e-6)
        outcome = 1;
      else
      {
        s.xne[cell] = (xnew + xne) / 2.;
        const int it = s.iter[cell] + 1;
        s.iter[cell] = it;
        outcome = it == ION_MAXITERATIONS ? 2 : 0;
      }
      s_out = outcome;
    }
    __syncthreads ();
    const int outcome = s_out;
    if (outcome == 1)
    {
      // absolute densities and get_ne of them (:366-378)
      for (int i = lane; i < n; i += 64)
      {
        double v = x[i];
        if (v < ION_DENSITY_MIN)
          v = ION_DENSITY_MIN;
        const double dn = v * ed[i];
        s.density[o + i] = dn;
        s_p[i] = dn * (ION_ISTATE (bt[i]) - 1);
      }
      __syncthreads ();
      if (lane == 0)
      {
        double ne = 0;
        for (int k = 0; k < n; k++)
          ne += s_p[k];
        s.ne[cell] = ne;
        s.status[cell] = 0;
        s.iters[cell] = s.iter[cell];
        s.active[cell] = 0;
      }
    }
    else if (lane == 0)
    {
      if (outcome == 2)
      {
        s.status[cell] = -1;
        s.iters[cell] = ION_MAXITERATIONS;
        s.active[cell] = 0;
      }
      else
        atomicAdd (s.nactive, 1);
    }
    __syncthreads ();
  }
}

// --------------------------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------------------------
struct hipla_ion_s
{
  int ncells = 0;
  hipla_plan plan = nullptr;   // FP64 plan over the internal matrix layout (cell geometry, stores)
  std::vector<int> n;
  int64_t totA = 0, totV = 0;
  double *dA = nullptr, *dB = nullptr, *dX = nullptr;
  int *dcodes = nullptr;
  double *d_xne = nullptr, *d_ed = nullptr, *d_newden = nullptr;
  int *d_iter = nullptr, *d_active = nullptr, *d_route = nullptr, *d_lists = nullptr, *d_nactive = nullptr;
  int *h_nactive = nullptr;    // pinned
  int route_used[ION_ROUTES] = { 0, 0, 0, 0, 0 };
  int route_nmax[ION_ROUTES] = { 0, 0, 0, 0, 0 };
  hipStream_t cap = nullptr;
  // the FP64 routes of one iteration run side by side, one stream each, forked after the assembly
  // and joined before the update (in the graph, parallel branches): once few cells are left, an
  // iteration costs its slowest route instead of the sum of all five
  hipStream_t rs[ION_ROUTES] = { nullptr, nullptr, nullptr, nullptr, nullptr };
  hipEvent_t ev_fork = nullptr, ev_join[ION_ROUTES] = { nullptr, nullptr, nullptr, nullptr, nullptr };
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // the same iterations once at most one cell per CU is left (the active set only shrinks): the
  // order 175 < n <= 256 route runs the look-ahead kernel alone
  hipGraph_t graph_lone = nullptr;
  hipGraphExec_t gexec_lone = nullptr;
  IonArgs gargs{};             // the arguments the graphs were captured with
  bool have_graph = false, have_lone = false;
};

// The stream (branch) a route's launches go to: the three one-wave routes share one, one after
// another (short even with every cell active; on streams of their own, one of them would share a
// hardware queue with the longest route and run before or after it).
static int ion_branch_of (int r)
{
  return r < 3 ? 0 : r;
}

static int ion_route_of (int n)
{
  return n <= 16 ? 0 : n <= 32 ? 1 : n <= 64 ? 2 : n <= wgr_f64_limit () ? 3 : 4;
}

static void ion_drop_graphs (hipla_ion c)
{
  if (c->gexec)
    (void) hipGraphExecDestroy (c->gexec);
  if (c->graph)
    (void) hipGraphDestroy (c->graph);
  if (c->gexec_lone)
    (void) hipGraphExecDestroy (c->gexec_lone);
  if (c->graph_lone)
    (void) hipGraphDestroy (c->graph_lone);
  c->gexec = c->gexec_lone = nullptr;
  c->graph = c->graph_lone = nullptr;
  c->have_graph = c->have_lone = false;
}

static void ion_free (hipla_ion c)
{
  ion_drop_graphs (c);
  void *bufs[] = { c->dA, c->dB, c->dX, c->dcodes, c->d_xne, c->d_ed, c->d_newden, c->d_iter, c->d_active, c->d_route,
                   c->d_lists, c->d_nactive };
  for (void *b : bufs)
    if (b)
      (void) hipla_dev_free (b);
  if (c->h_nactive)
    (void) hipHostFree (c->h_nactive);
  if (c->cap)
    (void) hipStreamDestroy (c->cap);
  for (int r = 0; r < ION_ROUTES; r++)
  {
    if (c->rs[r])
      (void) hipStreamDestroy (c->rs[r]);
    if (c->ev_join[r])
      (void) hipEventDestroy (c->ev_join[r]);
  }
  if (c->ev_fork)
    (void) hipEventDestroy (c->ev_fork);
  if (c->plan)
    (void) hipla_plan_destroy (c->plan);
}

extern "C" int hipla_ion_create (hipla_ion *out, int ncells, const int *n, const int64_t *offV)
{
  if (!out || ncells <= 0 || !n || !offV)
  {
    hipla_set_detail ("hipla_ion_create: bad argument");
    return 5;
  }
  for (int c = 0; c < ncells; c++)
    if (n[c] <= 0 || n[c] > HIPLA_ION_MAXN || offV[c] < 0 || offV[c] > INT64_MAX - n[c])
    {
      hipla_set_detail ("hipla_ion_create: cell %d has n=%d offV=%lld (1 <= n <= %d)", c, n[c], (long long) offV[c],
                        HIPLA_ION_MAXN);
      return 5;
    }
  int rc = hipla_ensure_init ();
  if (rc)
    return rc;
  *out = nullptr;
  hipla_ion c = new hipla_ion_s ();
  c->ncells = ncells;
  c->n.assign (n, n + ncells);
  std::vector<int64_t> offA (ncells);
  std::vector<int> route (ncells);
  for (int i = 0; i < ncells; i++)
  {
    offA[i] = c->totA;
    c->totA += (int64_t) n[i] * n[i];
    c->totV = std::max<int64_t> (c->totV, offV[i] + n[i]);
    route[i] = ion_route_of (n[i]);
    c->route_used[route[i]] = 1;
    c->route_nmax[route[i]] = std::max (c->route_nmax[route[i]], n[i]);
  }
#define ICHECK(expr)                                                            \
  do                                                                            \
  {                                                                             \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
    {                                                                           \
      hipla_set_detail ("%s failed: %s", #expr, hipGetErrorString (_e));        \
      ion_free (c);                                                             \
      delete c;                                                                 \
      return 5;                                                                 \
    }                                                                           \
  } while (0)
  rc = hipla_plan_create (&c->plan, ncells, n, offA.data (), offV, HIPLA_MODE_FP64, 0);
  if (rc)
  {
    delete c;
    return rc;
  }
  ICHECK (hipla_dev_alloc (&c->dA, sizeof (double) * c->totA));
  ICHECK (hipla_dev_alloc (&c->dB, sizeof (double) * c->totV));
  ICHECK (hipla_dev_alloc (&c->dX, sizeof (double) * c->totV));
  ICHECK (hipla_dev_alloc (&c->dcodes, sizeof (int) * ncells));
  ICHECK (hipla_dev_alloc (&c->d_xne, sizeof (double) * ncells));
  ICHECK (hipla_dev_alloc (&c->d_ed, sizeof (double) * c->totV));
  ICHECK (hipla_dev_alloc (&c->d_newden, sizeof (double) * c->totV));
  ICHECK (hipla_dev_alloc (&c->d_iter, sizeof (int) * ncells));
  ICHECK (hipla_dev_alloc (&c->d_active, sizeof (int) * ncells));
  ICHECK (hipla_dev_alloc (&c->d_route, sizeof (int) * ncells));
  ICHECK (hipla_dev_alloc (&c->d_lists, sizeof (int) * ION_ROUTES * ((size_t) ncells + 1)));
  ICHECK (hipla_dev_alloc (&c->d_nactive, sizeof (int)));
  ICHECK (hipHostMalloc ((void **) &c->h_nactive, sizeof (int), hipHostMallocDefault));
  ICHECK (hipMemcpy (c->d_route, route.data (), sizeof (int) * ncells, hipMemcpyHostToDevice));
  ICHECK (hipStreamCreateWithFlags (&c->cap, hipStreamNonBlocking));
  hipla_count_resource ();
  ICHECK (hipEventCreateWithFlags (&c->ev_fork, hipEventDisableTiming));
  hipla_count_resource ();
  for (int r = 0; r < ION_ROUTES; r++)
  {
    const int br = ion_branch_of (r);
    if (c->route_used[r] && !c->rs[br])
    {
      ICHECK (hipStreamCreateWithFlags (&c->rs[br], hipStreamNonBlocking));
      ICHECK (hipEventCreateWithFlags (&c->ev_join[br], hipEventDisableTiming));
      hipla_count_resource ();
      hipla_count_resource ();
    }
  }
#undef ICHECK
  *out = c;
  return 0;
}

extern "C" int hipla_ion_destroy (hipla_ion c)
{
  if (!c)
    return 0;
  ion_free (c);
  delete c;
  return 0;
}

static IonArgs ion_args (hipla_ion c, const hipla_ion_rates *r, double *dDensity, double *dNe, int *dStatus,
                         int *dIters, int *dCodes)
{
  IonArgs a{};
  a.n = c->plan->d_n;
  a.offA = c->plan->d_offA;
  a.offV = c->plan->d_offV;
  a.ncells = c->ncells;
  a.r = *r;
  a.A = c->dA;
  a.B = c->dB;
  a.X = c->dX;
  a.codes = c->dcodes;
  a.density_in = dDensity;
  a.density = dDensity;
  a.ne = dNe;
  a.status = dStatus;
  a.iters = dIters;
  a.codes_out = dCodes;
  a.xne = c->d_xne;
  a.elem_dens = c->d_ed;
  a.newden = c->d_newden;
  a.iter = c->d_iter;
  a.active = c->d_active;
  a.route = c->d_route;
  a.lists = c->d_lists;
  a.listcap = c->ncells;
  a.nactive = c->d_nactive;
  return a;
}

static size_t ion_asm_smem (int nmax)
{
  // rinfo padded to whole groups of 16 rows, colv, el, bt
  return (size_t) ((nmax + 15) & ~15) * sizeof (int) + (size_t) nmax * ION_NCOL * sizeof (double) +
         2 * (size_t) nmax * sizeof (int);
}

// one iteration: reset, assemble, the FP64 solves of the listed cells, update.  lone: at most
// lookahead_f64_max_cells () cells are active.
static int ion_enqueue_iteration (hipla_ion c, const IonArgs &a, hipStream_t st, bool lone)
{
  const DeviceInfo &dev = hipla_device_info ();
  const int nmax = *std::max_element (c->n.begin (), c->n.end ());
  hipLaunchKernelGGL (k_ion_reset, dim3 (1), dim3 (64), 0, st, a);
  HIPLA_CHECK (hipGetLastError ());
  const int agrid = std::max (1, std::min (c->ncells, 4 * dev.num_cu));
  hipLaunchKernelGGL (k_ion_assemble, dim3 (agrid), dim3 (ION_ASM_THREADS), ion_asm_smem (nmax), st, a);
  HIPLA_CHECK (hipGetLastError ());
  HIPLA_CHECK (hipEventRecord (c->ev_fork, st));
  // three branches (ion_branch_of), largest orders first
  bool forked[ION_ROUTES] = { false, false, false, false, false };
  for (int r = ION_ROUTES - 1; r >= 0; r--)
  {
    if (!c->route_used[r])
      continue;
    const int br = ion_branch_of (r);
    hipStream_t rst = c->rs[br];
    if (!forked[br])
      HIPLA_CHECK (hipStreamWaitEvent (rst, c->ev_fork, 0));
    forked[br] = true;
    SolveArgs sa;
    sa.n = a.n;
    sa.offA = a.offA;
    sa.offV = a.offV;
    sa.A = c->dA;
    sa.B = c->dB;
    sa.X = c->dX;
    sa.codes = c->dcodes;
    sa.list = c->d_lists + (int64_t) r * (c->ncells + 1) + 1;
    sa.nlist = c->ncells;
    sa.dcount = c->d_lists + (int64_t) r * (c->ncells + 1);
    sa.LU = nullptr;
    sa.piv = nullptr;
    sa.F = c->plan->d_F;
    sa.Y = c->plan->d_Y;
    if (r < 3)
      HIPLA_CHECK (launch_wave_f64 (sa, r == 0 ? 16 : r == 1 ? 32 : 64, false, 2 * dev.num_cu, rst));
    else if (r == 3)
      HIPLA_CHECK (launch_wgr_f64 (sa, c->route_nmax[3], false, dev.num_cu, rst));
    else if (lone && c->route_nmax[4] <= LA_ROUTE_MAXN)
      HIPLA_CHECK (launch_lookahead_route_f64 (sa, c->route_nmax[4], false, rst));
    else
      HIPLA_CHECK (launch_blocked_f64 (sa, c->route_nmax[4], blocked_f64_wgs_per_cu (c->route_nmax[4]) * dev.num_cu,
                                       false, rst));
  }
  for (int br = 0; br < ION_ROUTES; br++)
    if (forked[br])
    {
      HIPLA_CHECK (hipEventRecord (c->ev_join[br], c->rs[br]));
      HIPLA_CHECK (hipStreamWaitEvent (st, c->ev_join[br], 0));
    }
  const int ugrid = std::max (1, std::min (c->ncells, 8 * dev.num_cu));
  hipLaunchKernelGGL (k_ion_update, dim3 (ugrid), dim3 (64), 0, st, a);
  HIPLA_CHECK (hipGetLastError ());
  return 0;
}

// the pointers and flag a captured graph depends on
static bool same_args (const IonArgs &x, const IonArgs &y)
{
  const hipla_ion_rates &p = x.r, &q = y.r;
  return p.elem == q.elem && p.bits == q.bits && p.pi == q.pi && p.rr == q.rr && p.di == q.di && p.dr == q.dr &&
         p.qr == q.qr && p.cxr == q.cxr && p.cxi == q.cxi && p.inner_beg == q.inner_beg &&
         p.inner_cnt == q.inner_cnt && p.inner_rate == q.inner_rate && p.inner_yield == q.inner_yield &&
         p.has_cx == q.has_cx && x.density == y.density && x.ne == y.ne && x.status == y.status &&
         x.iters == y.iters && x.codes_out == y.codes_out;
}

// capture ION_CHUNK iterations into a graph and instantiate it
static int ion_capture (hipla_ion c, const IonArgs &a, bool lone, hipGraph_t *graph, hipGraphExec_t *gexec)
{
  HIPLA_CHECK (hipStreamBeginCapture (c->cap, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < ION_CHUNK && rc == 0; k++)
    rc = ion_enqueue_iteration (c, a, c->cap, lone);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture (c->cap, &g);
  if (rc)
  {
    if (g)
      (void) hipGraphDestroy (g);
    return rc;
  }
  HIPLA_CHECK (e);
  *graph = g;
  HIPLA_CHECK (hipGraphInstantiate (gexec, g, nullptr, nullptr, 0));
  return 0;
}

extern "C" int hipla_ion_populations (hipla_ion c, const hipla_ion_rates *rates, double *dDensity, double *dNe,
                                      int *dStatus, int *dIters, int *dCodes, void *stream)
{
  if (!c || !rates || !dDensity || !dNe || !dStatus || !dIters || !dCodes || !rates->elem || !rates->bits ||
      !rates->pi || !rates->rr || !rates->di || !rates->dr || !rates->qr || !rates->cxr || !rates->cxi ||
      !rates->inner_beg || !rates->inner_cnt)
  {
    hipla_set_detail ("hipla_ion_populations: bad argument");
    return 5;
  }
  hipStream_t st = (hipStream_t) stream;
  IonArgs a = ion_args (c, rates, dDensity, dNe, dStatus, dIters, dCodes);
  const DeviceInfo &dev = hipla_device_info ();
  hipLaunchKernelGGL (k_ion_init, dim3 (std::max (1, std::min (c->ncells, 8 * dev.num_cu))), dim3 (64), 0, st, a);
  HIPLA_CHECK (hipGetLastError ());
  // ION_CHUNK iterations per graph replay, captured once per argument set (the lone variant when
  // first needed)
  if (!same_args (a, c->gargs))
    ion_drop_graphs (c);
  c->gargs = a;
  bool lone = false;
  // at most MAXITERATIONS solves per cell: every cell has left after that many iterations
  const int max_replays = (ION_MAXITERATIONS + ION_CHUNK - 1) / ION_CHUNK + 1;
  for (int k = 0; k < max_replays; k++)
  {
    bool &have = lone ? c->have_lone : c->have_graph;
    if (!have)
    {
      int rc = ion_capture (c, a, lone, lone ? &c->graph_lone : &c->graph, lone ? &c->gexec_lone : &c->gexec);
      if (rc)
        return rc;
      have = true;
    }
    HIPLA_CHECK (hipGraphLaunch (lone ? c->gexec_lone : c->gexec, st));
    HIPLA_CHECK (hipMemcpyAsync (c->h_nactive, c->d_nactive, sizeof (int), hipMemcpyDeviceToHost, st));
    HIPLA_CHECK (hipStreamSynchronize (st));
    if (*c->h_nactive == 0)
      return 0;
    // cells only leave the iteration, so from here on every list holds at most this many
    if (*c->h_nactive <= lookahead_f64_max_cells ())
      lone = true;
  }
  hipla_set_detail ("hipla_ion_populations: %d cells still active after %d iterations", *c->h_nactive,
                    max_replays * ION_CHUNK);
  return 5;
}
