// hipla_plan.hip -- batched ragged plans: size-class routing, fork-join launch, hipGraph replay.
//
// The reference solves one cell per call inside its per-cell loops (source/wind_updates2d.c:89-153,
// source/macro_accelerate.c:1052-1061).  A plan takes the whole ragged batch at once: every cell is
// routed by its order n to a kernel shape (wave-per-matrix, LDS-resident, or blocked), each shape's
// cells are sorted largest first so the hardware dispatcher approximates longest-processing-time
// scheduling, the shapes run as concurrent branches of one fork-join launch, and the launch
// sequence is captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "hipla_kernels.h"
#include "hipla_plan.h"
#include "hipla_tune.h"

static const char *kind_name (int k)
{
  switch (k)
  {
  case KK_WAVE_F64:
    return "wave_f64";
  case KK_WGR_F64:
    return "tile_f64";
  case KK_BLOCKED_F64:
    return "blocked_f64";
  case KK_WAVE_MIX:
    return "wave_mixed";
  case KK_LDS_MIX:
    return "blk_mixed";
  case KK_BLOCKED_MIX:
    return "fp64_over_256";
  default:
    return "?";
  }
}

static void route_fp64 (int n, int &kind, int &param)
{
  if (n <= 16)
  {
    kind = KK_WAVE_F64;
    param = 16;
  }
  else if (n <= 32)
  {
    kind = KK_WAVE_F64;
    param = 32;
  }
  else if (n <= 64)
  {
    kind = KK_WAVE_F64;
    param = 64;
  }
  else if (n <= wgr_f64_limit ())
  {
    kind = KK_WGR_F64;
    param = 0;
  }
  else
  {
    kind = KK_BLOCKED_F64;
    param = BLK_NB;
  }
}

void plan_route_fp64 (int n, int *kind, int *param)
{
  route_fp64 (n, *kind, *param);
}

static int free_plan_device (hipla_plan p)
{
  if (p->gexec)
    (void) hipGraphExecDestroy (p->gexec);
  if (p->graph)
    (void) hipGraphDestroy (p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
  void *bufs[] = { p->d_n,   p->d_offA, p->d_offV, p->d_list, p->d_list_desc, p->d_inv_work, p->d_fallback,
                   p->d_LU,  p->d_piv,  p->own_F ? p->d_F : nullptr, p->d_Y, p->d_lam,
                   p->d_gj_list, p->d_gj_flag, p->d_gj_fb, p->d_inv_rest, p->d_gx_Mf, p->d_gx_Rf };
  for (void *b : bufs)
    if (b)
      (void) hipla_dev_free (b);
  for (auto s : p->aux)
    (void) hipStreamDestroy (s);
  p->aux.clear ();
  if (p->fork_st)
    (void) hipStreamDestroy (p->fork_st);
  if (p->ev_fork)
    (void) hipEventDestroy (p->ev_fork);
  if (p->ev_join)
    (void) hipEventDestroy (p->ev_join);
  p->fork_st = nullptr;
  p->ev_fork = p->ev_join = nullptr;
  return 0;
}

extern "C" int hipla_plan_destroy (hipla_plan plan)
{
  if (!plan)
    return 0;
  free_plan_device (plan);
  delete plan;
  return 0;
}

extern "C" int hipla_plan_create (hipla_plan *out, int ncells, const int *n, const int64_t *offA, const int64_t *offV,
                                  int mode, int flags)
{
  if (!out || ncells <= 0 || !n || !offA || (!offV && ncells > 0))
  {
    hipla_set_detail ("hipla_plan_create: bad argument");
    return 5;
  }
  if (mode != HIPLA_MODE_FP64 && mode != HIPLA_MODE_MIXED)
  {
    hipla_set_detail ("hipla_plan_create: unknown mode %d", mode);
    return 5;
  }
  // every cell's extent is checked before anything else (so offA + n^2 cannot wrap: the offsets
  // are compared with the room the int64 extents leave)
  for (int c = 0; c < ncells; c++)
  {
    const int64_t k = n[c];
    if (k <= 0 || k > 65536 || offA[c] < 0 || offV[c] < 0 || offA[c] > INT64_MAX - k * k ||
        offV[c] > INT64_MAX - k)
    {
      hipla_set_detail ("hipla_plan_create: cell %d has n=%d offA=%lld offV=%lld", c, n[c], (long long) offA[c],
                        (long long) offV[c]);
      return 5;
    }
  }
  int rc = hipla_ensure_init ();
  if (rc)
    return rc;
  *out = nullptr;
  hipla_plan p = new hipla_plan_s ();
  p->ncells = ncells;
  p->mode = mode;
  p->flags = flags;
  p->n.assign (n, n + ncells);
  p->offA.assign (offA, offA + ncells);
  p->offV.assign (offV, offV + ncells);
  for (int c = 0; c < ncells; c++)
  {
    p->totA = std::max<int64_t> (p->totA, offA[c] + (int64_t) n[c] * n[c]);
    p->totV = std::max<int64_t> (p->totV, offV[c] + n[c]);
  }

  // ---- route every cell to a (kind, param) group ----
  std::vector<int> kind (ncells), param (ncells);
  for (int c = 0; c < ncells; c++)
  {
    if (mode == HIPLA_MODE_MIXED)
      route_mixed (n[c], kind[c], param[c]);
    else
      route_fp64 (n[c], kind[c], param[c]);
  }
  std::vector<int> idx (ncells);
  std::iota (idx.begin (), idx.end (), 0);
  // Groups of one kind in ascending param, except the mixed plan's column-block groups: those are
  // in descending param, so that when adjacent groups are merged into one FP64 launch its list
  // still runs largest cells first.
  std::stable_sort (idx.begin (), idx.end (), [&] (int a, int b) {
    if (kind[a] != kind[b])
      return kind[a] < kind[b];
    if (param[a] != param[b])
      return kind[a] == KK_LDS_MIX ? param[a] > param[b] : param[a] < param[b];
    return n[a] > n[b];
  });
  p->list = idx;
  const DeviceInfo &dev = hipla_device_info ();
  for (int i = 0; i < ncells;)
  {
    int j = i;
    while (j < ncells && kind[idx[j]] == kind[idx[i]] && param[idx[j]] == param[idx[i]])
      j++;
    Group g;
    g.kind = kind[idx[i]];
    g.param = param[idx[i]];
    g.first = i;
    g.count = j - i;
    g.nmax = n[idx[i]];
    g.stream = (int) p->groups.size ();
    // persistent grid of groups that run (or, in mixed plans, may be routed to) the blocked
    // FP64 kernel; it factors each cell in place in the plan's factor store
    if (g.kind == KK_BLOCKED_F64 || g.kind == KK_BLOCKED_MIX ||
        (mode == HIPLA_MODE_MIXED && g.nmax > wgr_f64_limit ()))
      g.grid = std::min (g.count, blocked_f64_wgs_per_cu (g.nmax) * dev.num_cu);
    p->groups.push_back (g);
    i = j;
  }
  // the one-wave kernel for 16 < n <= 32 takes two-wave workgroups by default: 9.6 KB of LDS, so
  // they fit beside the blocked kernel's workgroups on a CU.  With no class on the blocked kernel
  // nothing runs beside them, and four-wave workgroups are slightly faster (the autotuner still
  // measures both).
  bool has_blocked = false;
  for (const Group &g : p->groups)
    has_blocked = has_blocked || g.nmax > wgr_f64_limit ();
  if (!has_blocked)
    for (Group &g : p->groups)
      if (fp64_family_of (g.nmax) == 1)
        g.variant = 1;

  // mixed mode: FP64 routes the fallback re-solve can need
  if (mode == HIPLA_MODE_MIXED)
  {
    for (int c = 0; c < ncells; c++)
    {
      if (kind[c] == KK_BLOCKED_MIX)
        continue;
      int fk, fp;
      route_fp64 (n[c], fk, fp);
      const int r = fk == KK_WAVE_F64 ? (fp == 16 ? 0 : fp == 32 ? 1 : 2) : fk == KK_WGR_F64 ? 3 : 4;
      p->fb_route_used[r] = 1;
    }
  }

  // ---- device copies and workspaces ----
#define PCHECK(expr)                                                                       \
  do                                                                                       \
  {                                                                                        \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
    {                                                                                      \
      hipla_set_detail ("%s failed: %s", #expr, hipGetErrorString (_e));                   \
      free_plan_device (p);                                                                \
      delete p;                                                                            \
      return 5;                                                                            \
    }                                                                                      \
  } while (0)
  PCHECK (hipla_dev_alloc (&p->d_n, sizeof (int) * ncells));
  PCHECK (hipla_dev_alloc (&p->d_offA, sizeof (int64_t) * ncells));
  PCHECK (hipla_dev_alloc (&p->d_offV, sizeof (int64_t) * ncells));
  PCHECK (hipla_dev_alloc (&p->d_list, sizeof (int) * ncells));
  PCHECK (hipMemcpy (p->d_n, n, sizeof (int) * ncells, hipMemcpyHostToDevice));
  PCHECK (hipMemcpy (p->d_offA, offA, sizeof (int64_t) * ncells, hipMemcpyHostToDevice));
  PCHECK (hipMemcpy (p->d_offV, offV, sizeof (int64_t) * ncells, hipMemcpyHostToDevice));
  PCHECK (hipMemcpy (p->d_list, p->list.data (), sizeof (int) * ncells, hipMemcpyHostToDevice));
  {
    std::vector<int> desc (ncells);
    std::iota (desc.begin (), desc.end (), 0);
    std::stable_sort (desc.begin (), desc.end (), [&] (int a, int b) { return n[a] > n[b]; });
    PCHECK (hipla_dev_alloc (&p->d_list_desc, sizeof (int) * ncells));
    PCHECK (hipMemcpy (p->d_list_desc, desc.data (), sizeof (int) * ncells, hipMemcpyHostToDevice));
  }
  if (mode == HIPLA_MODE_MIXED)
    PCHECK (hipla_dev_alloc (&p->d_fallback, sizeof (int) * 5 * (ncells + 1)));
  if (flags & HIPLA_FLAG_KEEP_FACTORS)
  {
    PCHECK (hipla_dev_alloc (&p->d_LU, sizeof (double) * p->totA));
    PCHECK (hipla_dev_alloc (&p->d_piv, sizeof (int) * p->totV));
    PCHECK (hipla_dev_alloc (&p->d_lam, sizeof (double) * p->totV));
  }
  // factor and y stores of the workgroup factor kernels (cells above 64), read by the
  // back-substitution kernel; kept factors double as the factor store
  {
    int nbig = 0;
    for (int c = 0; c < ncells; c++)
      nbig |= n[c] > 64;
    if (nbig)
    {
      if (p->d_LU)
        p->d_F = p->d_LU;
      else
      {
        PCHECK (hipla_dev_alloc (&p->d_F, sizeof (double) * p->totA));
        p->own_F = true;
      }
      PCHECK (hipla_dev_alloc (&p->d_Y, sizeof (double) * p->totV));
    }
  }
  p->aux.resize (1, nullptr);
  PCHECK (hipStreamCreateWithFlags (&p->aux[0], hipStreamNonBlocking));
  PCHECK (hipStreamCreateWithFlags (&p->fork_st, hipStreamNonBlocking));
  PCHECK (hipEventCreateWithFlags (&p->ev_fork, hipEventDisableTiming));
  PCHECK (hipEventCreateWithFlags (&p->ev_join, hipEventDisableTiming));
  for (int i = 0; i < 4; i++)
    hipla_count_resource ();
#undef PCHECK
  *out = p;
  return 0;
}

extern "C" int hipla_plan_set_refinement (hipla_plan plan, int max_refine, double tol)
{
  if (!plan || max_refine < 0 || max_refine > 50 || !(tol > 0.0))
  {
    hipla_set_detail ("hipla_plan_set_refinement: bad argument");
    return 5;
  }
  // the cached graph holds kernel arguments captured with the old settings: drop it so the next
  // solve captures the new ones
  if (plan->max_refine != max_refine || plan->tol != tol)
  {
    if (plan->gexec)
      (void) hipGraphExecDestroy (plan->gexec);
    if (plan->graph)
      (void) hipGraphDestroy (plan->graph);
    plan->gexec = nullptr;
    plan->graph = nullptr;
    plan->gkind = -1;
  }
  plan->max_refine = max_refine;
  plan->tol = tol;
  return 0;
}

extern "C" int hipla_plan_describe (hipla_plan p, char *buf, int buflen)
{
  if (!p || !buf || buflen <= 0)
    return 5;
  std::string s;
  char line[256];
  snprintf (line, sizeof line, "{\"ncells\": %d, \"mode\": %d, \"flags\": %d, \"groups\": [", p->ncells, p->mode,
            p->flags);
  s += line;
  for (size_t i = 0; i < p->groups.size (); i++)
  {
    const Group &g = p->groups[i];
    // algorithmic work of the group's solves: bytes 8 (n^2 + 2 n), flops 2/3 n^3 + 2 n^2
    double bytes = 0.0, flops = 0.0;
    for (int j = g.first; j < g.first + g.count; j++)
    {
      const double nn = p->n[p->list[j]];
      bytes += 8.0 * (nn * nn + 2.0 * nn);
      flops += 2.0 / 3.0 * nn * nn * nn + 2.0 * nn * nn;
    }
    // the FP64 kernel instantiation that runs the group (FP64 plans, or a mixed size class
    // routed straight to FP64)
    char sym[96];
    fp64_variant_symbol (g.nmax, g.variant, (p->flags & HIPLA_FLAG_KEEP_FACTORS) != 0, sym, sizeof sym, g.count);
    char line2[512];
    snprintf (line2, sizeof line2,
              "%s{\"kernel\": \"%s\", \"param\": %d, \"nmax\": %d, \"count\": %d, \"grid\": %d, "
              "\"direct_fp64\": %d, \"variant\": %d, \"t_mixed_ms\": %.4f, \"t_direct_ms\": %.4f, \"bytes\": %.0f, "
              "\"flops\": %.0f, \"fp64_kernel\": \"%s\"}",
              i ? ", " : "", kind_name (g.kind), g.param, g.nmax, g.count, g.grid, g.direct, g.variant, g.t_mixed,
              g.t_direct,
              bytes, flops, sym);
    s += line2;
  }
  s += "]}";
  snprintf (buf, (size_t) buflen, "%s", s.c_str ());
  return 0;
}

// --------------------------------------------------------------------------------------------
// enqueue
// --------------------------------------------------------------------------------------------
static SolveArgs make_args (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                            const Group &g)
{
  SolveArgs a;
  a.n = p->d_n;
  a.offA = p->d_offA;
  a.offV = p->d_offV;
  a.A = dA;
  a.B = dB;
  a.X = dX;
  a.codes = dcodes;
  a.list = p->d_list + g.first;
  a.nlist = g.count;
  a.LU = p->d_LU;
  a.piv = p->d_piv;
  a.dcount = nullptr;
  a.F = p->d_F;
  a.Y = p->d_Y;
  return a;
}

static hipError_t launch_group_f64 (hipla_plan p, const SolveArgs &a, const Group &g, hipStream_t st)
{
  const bool keep = (p->flags & HIPLA_FLAG_KEEP_FACTORS) != 0;
  switch (g.kind)
  {
  case KK_WAVE_F64:
  case KK_WGR_F64:
    return launch_fp64_variant (a, g.nmax, g.variant, keep, 0, st);
  case KK_BLOCKED_F64:
    return launch_fp64_variant (a, g.nmax, g.variant, keep, std::min (g.grid, std::max (a.nlist, 1)), st);
  default:
    return hipErrorInvalidValue;
  }
}

int enqueue_fp64_group (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipStream_t st)
{
  const Group &g = p->groups[gi];
  SolveArgs a = make_args (p, dA, dB, dX, dcodes, g);
  HIPLA_CHECK (launch_group_f64 (p, a, g, st));
  return 0;
}

int plan_enqueue_solve (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipla_cell_stats *dstats, hipStream_t st)
{
  if (p->mode == HIPLA_MODE_MIXED)
    return plan_enqueue_mixed (p, dA, dB, dX, dcodes, dstats, st);
  const bool per_cell = (p->flags & HIPLA_FLAG_PER_CELL) != 0;
  // the size classes one after another on the caller's stream: each class's launch already
  // spreads over every CU, and all classes side by side only contend for LDS and registers
  // (measured slower than back to back).  The exception: a last class on the blocked kernel runs,
  // with its back substitution, on the fork stream from the start, beside the smaller classes
  // (they take the CU slots its persistent workgroups leave as they finish: 2 % faster on the
  // uneven 10k batch); the caller's stream waits for it at the end.
  const int ng = (int) p->groups.size ();
  const bool fork = !per_cell && ng > 1 && p->groups.back ().kind == KK_BLOCKED_F64;
  if (fork)
  {
    const Group &g = p->groups.back ();
    SolveArgs a = make_args (p, dA, dB, dX, dcodes, g);
    const bool keep = (p->flags & HIPLA_FLAG_KEEP_FACTORS) != 0;
    HIPLA_CHECK (hipEventRecord (p->ev_fork, st));
    HIPLA_CHECK (hipStreamWaitEvent (p->fork_st, p->ev_fork, 0));
    HIPLA_CHECK (launch_fp64_variant (a, g.nmax, g.variant, keep, std::min (g.grid, std::max (a.nlist, 1)), p->fork_st,
                                      true));
    HIPLA_CHECK (hipEventRecord (p->ev_join, p->fork_st));
  }
  for (int gi = 0; gi < ng - (fork ? 1 : 0); gi++)
  {
    const Group &g = p->groups[gi];
    SolveArgs a = make_args (p, dA, dB, dX, dcodes, g);
    if (per_cell)
    {
      // the reference's GPU path: one factor+solve launch per matrix (source/matrix_gpu.cu:263-307)
      for (int c = 0; c < g.count; c++)
      {
        SolveArgs a1 = a;
        a1.list = a.list + c;
        a1.nlist = 1;
        HIPLA_CHECK (launch_group_f64 (p, a1, g, st));
      }
    }
    else
      HIPLA_CHECK (launch_group_f64 (p, a, g, st));
  }
  if (fork)
    HIPLA_CHECK (hipStreamWaitEvent (st, p->ev_join, 0));
  return 0;
}

// --------------------------------------------------------------------------------------------
// public execution entry points (graph capture / replay)
// --------------------------------------------------------------------------------------------
static int run_maybe_graph (hipla_plan p, int kind, const void *k0, const void *k1, const void *k2, const void *k3,
                            const void *k4, hipStream_t st,
                            int (*enqueue) (hipla_plan, const void *, const void *, const void *, const void *,
                                            const void *, hipStream_t))
{
  if (!(p->flags & HIPLA_FLAG_GRAPH) || (p->flags & HIPLA_FLAG_PER_CELL))
    return enqueue (p, k0, k1, k2, k3, k4, st);
  // a solve that is one kernel launch (an FP64 plan with a single one-wave class) goes straight to
  // the stream: a one-node graph adds to both sides (900 cells of n = 30: 43.9 us per solve from
  // the graph, 39.2 launched directly; the host's part 31 us against 9)
  if (kind == 0 && p->mode == HIPLA_MODE_FP64 && p->groups.size () == 1 && p->groups[0].kind == KK_WAVE_F64)
    return enqueue (p, k0, k1, k2, k3, k4, st);
  const void *key[6] = { k0, k1, k2, k3, k4, (const void *) (intptr_t) kind };
  bool hit = p->gexec != nullptr && p->gkind == kind;
  for (int i = 0; hit && i < 6; i++)
    hit = key[i] == p->gkey[i];
  if (!hit)
  {
    if (p->gexec)
      (void) hipGraphExecDestroy (p->gexec);
    if (p->graph)
      (void) hipGraphDestroy (p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    hipStream_t cap = p->aux.back ();
    HIPLA_CHECK (hipStreamBeginCapture (cap, hipStreamCaptureModeThreadLocal));
    int rc = enqueue (p, k0, k1, k2, k3, k4, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture (cap, &g);
    if (rc)
    {
      if (g)
        (void) hipGraphDestroy (g);
      return rc;
    }
    HIPLA_CHECK (e);
    p->graph = g;
    HIPLA_CHECK (hipGraphInstantiate (&p->gexec, g, nullptr, nullptr, 0));
    for (int i = 0; i < 6; i++)
      p->gkey[i] = key[i];
    p->gkind = kind;
  }
  HIPLA_CHECK (hipGraphLaunch (p->gexec, st));
  return 0;
}

static int enq_solve (hipla_plan p, const void *a, const void *b, const void *x, const void *c, const void *s,
                      hipStream_t st)
{
  return plan_enqueue_solve (p, (const double *) a, (const double *) b, (double *) x, (int *) c,
                             (hipla_cell_stats *) s, st);
}

static int enq_invert (hipla_plan p, const void *a, const void *inv, const void *c, const void *s, const void *,
                       hipStream_t st)
{
  return plan_enqueue_invert (p, (const double *) a, (double *) inv, (int *) c, (hipla_cell_stats *) s, st);
}

extern "C" int hipla_solve (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                            hipla_cell_stats *dstats, void *stream)
{
  if (!p || !dA || !dB || !dX || !dcodes)
  {
    hipla_set_detail ("hipla_solve: bad argument");
    return 5;
  }
  int rc = run_maybe_graph (p, 0, dA, dB, dX, dcodes, dstats, (hipStream_t) stream, enq_solve);
  if (rc == 0)
  {
    p->last_A = dA;
    p->solved = 1;
  }
  return rc;
}

extern "C" int hipla_invert (hipla_plan p, const double *dA, double *dInv, int *dcodes, hipla_cell_stats *dstats,
                             void *stream)
{
  if (!p || !dA || !dInv || !dcodes)
  {
    hipla_set_detail ("hipla_invert: bad argument");
    return 5;
  }
  int rc = plan_invert_prepare (p);
  if (rc)
    return rc;
  return run_maybe_graph (p, 1, dA, dInv, dcodes, dstats, nullptr, (hipStream_t) stream, enq_invert);
}
