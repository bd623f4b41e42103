// hipla_tune.hip -- per-group timing and the size-class autotuner of libhipla plans.
//
// hipla_plan_time_groups launches every size-class group of a plan on its own, bracketed by HIP
// events on the stream the kernels run on, and reports the average milliseconds per launch.
// hipla_plan_autotune uses it to choose, per size class of a mixed plan, between
//   (a) the FP32 LU + FP64 refinement kernel followed by the FP64 re-solve of the cells it hands
//       back, and
//   (b) the FP64 GSL-order kernel straight away,
// keeping the faster.  Rate matrices whose entries span many decades make (a) hand back most
// cells of some size classes, and then (b) wins.  Choices are cached per process keyed on
// (kernel shape, largest order, cell count) and, when the environment variable HIPLA_TUNE_CACHE
// names a file, read from and appended to that file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "hipla_kernels.h"
#include "hipla_plan.h"

int enqueue_fp64_group (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipStream_t st);

namespace
{
using Key = std::tuple<int, int, int, int>;   // kind, param, nmax, count
std::mutex tune_mu;
std::map<Key, int> tune_cache;                 // -> direct (0/1)
bool file_loaded = false;

void load_file_locked ()
{
  if (file_loaded)
    return;
  file_loaded = true;
  const char *path = getenv ("HIPLA_TUNE_CACHE");
  if (!path)
    return;
  FILE *f = fopen (path, "r");
  if (!f)
    return;
  int k, pa, nm, c, d;
  while (fscanf (f, "%d %d %d %d %d", &k, &pa, &nm, &c, &d) == 5)
    tune_cache[Key (k, pa, nm, c)] = d;
  fclose (f);
}

void store_locked (const Key &key, int direct)
{
  tune_cache[key] = direct;
  const char *path = getenv ("HIPLA_TUNE_CACHE");
  if (!path)
    return;
  FILE *f = fopen (path, "a");
  if (!f)
    return;
  fprintf (f, "%d %d %d %d %d\n", std::get<0> (key), std::get<1> (key), std::get<2> (key), std::get<3> (key), direct);
  fclose (f);
}

struct EvPair
{
  hipEvent_t a = nullptr, b = nullptr;
  EvPair ()
  {
    (void) hipEventCreate (&a);
    (void) hipEventCreate (&b);
    hipla_count_resource ();
    hipla_count_resource ();
  }
  ~EvPair ()
  {
    if (a)
      (void) hipEventDestroy (a);
    if (b)
      (void) hipEventDestroy (b);
  }
};

void drop_graph (hipla_plan p)
{
  if (p->gexec)
    (void) hipGraphExecDestroy (p->gexec);
  if (p->graph)
    (void) hipGraphDestroy (p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
  p->gkind = -1;
}

// Average ms of `reps` launches of one group (mixed: counter reset + group kernel).
int time_group (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes,
                hipla_cell_stats *dstats, int reps, hipStream_t st, double &ms)
{
  EvPair ev;
  HIPLA_CHECK (hipEventRecord (ev.a, st));
  for (int r = 0; r < reps; r++)
  {
    int rc;
    if (p->mode == HIPLA_MODE_MIXED)
    {
      if ((rc = enqueue_mixed_reset (p, st)) || (rc = enqueue_mixed_group (p, gi, dA, dB, dX, dcodes, dstats, st)))
        return rc;
    }
    else if ((rc = enqueue_fp64_group (p, gi, dA, dB, dX, dcodes, st)))
      return rc;
  }
  HIPLA_CHECK (hipEventRecord (ev.b, st));
  HIPLA_CHECK (hipEventSynchronize (ev.b));
  float t = 0.0f;
  HIPLA_CHECK (hipEventElapsedTime (&t, ev.a, ev.b));
  ms = t / reps;
  return 0;
}

// Average ms of the FP64 re-solve pass for whatever the last mixed launch handed back.
int time_fallback (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes, int reps, hipStream_t st,
                   double &ms)
{
  EvPair ev;
  HIPLA_CHECK (hipEventRecord (ev.a, st));
  for (int r = 0; r < reps; r++)
  {
    int rc = enqueue_mixed_fallback (p, dA, dB, dX, dcodes, st);
    if (rc)
      return rc;
  }
  HIPLA_CHECK (hipEventRecord (ev.b, st));
  HIPLA_CHECK (hipEventSynchronize (ev.b));
  float t = 0.0f;
  HIPLA_CHECK (hipEventElapsedTime (&t, ev.a, ev.b));
  ms = t / reps;
  return 0;
}
} // namespace

namespace
{
int tune_variant (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes, int reps,
                  hipStream_t st);
}

extern "C" int hipla_plan_time_groups (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                                       hipla_cell_stats *dstats, int reps, double *ms_out, int nout)
{
  if (!p || !dA || !dB || !dX || !dcodes || !ms_out || reps <= 0)
  {
    hipla_set_detail ("hipla_plan_time_groups: bad argument");
    return 5;
  }
  const int ng = (int) p->groups.size ();
  if (nout < ng)
  {
    hipla_set_detail ("hipla_plan_time_groups: need at least %d outputs", ng);
    return 5;
  }
  hipStream_t st = p->aux.back ();
  int rc;
  for (int gi = 0; gi < ng; gi++)
    if ((rc = time_group (p, gi, dA, dB, dX, dcodes, dstats, reps, st, ms_out[gi])))
      return rc;
  if (nout < ng + 2)
    return 0;        // size classes only
  ms_out[ng] = 0.0;
  if (p->mode == HIPLA_MODE_MIXED)
  {
    // populate the hand-back lists with one whole solve, then time the FP64 pass alone
    if ((rc = plan_enqueue_mixed (p, dA, dB, dX, dcodes, dstats, st)))
      return rc;
    if ((rc = time_fallback (p, dA, dB, dX, dcodes, reps, st, ms_out[ng])))
      return rc;
  }
  EvPair ev;
  HIPLA_CHECK (hipEventRecord (ev.a, st));
  for (int r = 0; r < reps; r++)
    if ((rc = hipla_solve (p, dA, dB, dX, dcodes, dstats, st)))
      return rc;
  HIPLA_CHECK (hipEventRecord (ev.b, st));
  HIPLA_CHECK (hipEventSynchronize (ev.b));
  float t = 0.0f;
  HIPLA_CHECK (hipEventElapsedTime (&t, ev.a, ev.b));
  ms_out[ng + 1] = t / reps;
  return 0;
}

namespace
{
// The FP64 kernel shape of one group: every variant of its family (launch_fp64_variant: waves
// per workgroup of the one-wave-per-matrix kernels, register tiles or blocked for the middle
// orders, 256 or 512 threads of the blocked kernel) timed alone, the fastest kept.  All variants
// give the same values (GSL's operation order).
int tune_variant (hipla_plan p, int gi, const double *dA, const double *dB, double *dX, int *dcodes, int reps,
                  hipStream_t st)
{
  Group &g = p->groups[gi];
  const int fam = fp64_family_of (g.nmax);
  const int nv = fp64_variants (fam, g.nmax);
  if (nv <= 1)
    return 0;
  const Key key (300 + fam, (p->flags & HIPLA_FLAG_KEEP_FACTORS) ? 1 : 0, g.nmax, g.count);
  auto it = tune_cache.find (key);
  if (it != tune_cache.end ())
  {
    g.variant = it->second;
    return 0;
  }
  int best = 0, rc;
  double tbest = 1e300;
  for (int v = 0; v < nv; v++)
  {
    g.variant = v;
    double t = 0.0;
    if ((rc = time_group (p, gi, dA, dB, dX, dcodes, nullptr, reps, st, t)))
      return rc;
    if (t < tbest)
    {
      tbest = t;
      best = v;
    }
  }
  g.variant = best;
  store_locked (key, best);
  return 0;
}
} // namespace

extern "C" int hipla_plan_autotune (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                                    void *stream)
{
  if (!p || !dA || !dB || !dX || !dcodes)
  {
    hipla_set_detail ("hipla_plan_autotune: bad argument");
    return 5;
  }
  if (p->flags & HIPLA_FLAG_PER_CELL)
    return 0;   // the one-launch-per-matrix baseline keeps its shapes
  hipStream_t st = stream ? (hipStream_t) stream : p->aux.back ();
  HIPLA_CHECK (hipStreamSynchronize (st));
  std::lock_guard<std::mutex> lk (tune_mu);
  load_file_locked ();
  const int reps = 3;
  int rc;
  if (p->mode != HIPLA_MODE_MIXED)
  {
    for (int gi = 0; gi < (int) p->groups.size (); gi++)
      if ((rc = tune_variant (p, gi, dA, dB, dX, dcodes, reps, st)))
        return rc;
    drop_graph (p);
    return 0;
  }
  for (int gi = 0; gi < (int) p->groups.size (); gi++)
  {
    Group &g = p->groups[gi];
    if (g.kind == KK_BLOCKED_MIX)
      continue;
    const Key key (g.kind, g.param, g.nmax, g.count);
    auto it = tune_cache.find (key);
    if (it != tune_cache.end ())
    {
      g.direct = it->second;
      continue;
    }
    double t_mix = 0.0, t_fb = 0.0, t_dir = 0.0;
    g.direct = 0;
    if ((rc = time_group (p, gi, dA, dB, dX, dcodes, nullptr, reps, st, t_mix)))
      return rc;
    if ((rc = time_fallback (p, dA, dB, dX, dcodes, reps, st, t_fb)))
      return rc;
    g.direct = 1;
    if ((rc = time_group (p, gi, dA, dB, dX, dcodes, nullptr, reps, st, t_dir)))
      return rc;
    g.direct = (t_dir < t_mix + t_fb) ? 1 : 0;
    g.t_mixed = t_mix + t_fb;
    g.t_direct = t_dir;
    store_locked (key, g.direct);
  }
  // the FP64 kernel shape of every size class that goes straight to FP64
  for (int gi = 0; gi < (int) p->groups.size (); gi++)
  {
    const Group &g = p->groups[gi];
    if ((g.kind == KK_BLOCKED_MIX || g.direct) && (rc = tune_variant (p, gi, dA, dB, dX, dcodes, reps, st)))
      return rc;
  }
  drop_graph (p);
  return 0;
}

extern "C" int hipla_plan_set_variant (hipla_plan p, int group, int variant)
{
  if (!p || group < 0 || group >= (int) p->groups.size () || variant < 0)
  {
    hipla_set_detail ("hipla_plan_set_variant: bad argument");
    return 5;
  }
  Group &g = p->groups[group];
  if (variant >= fp64_variants (fp64_family_of (g.nmax), g.nmax))
  {
    hipla_set_detail ("hipla_plan_set_variant: group %d (n <= %d) has %d kernel shapes", group, g.nmax,
                      fp64_variants (fp64_family_of (g.nmax), g.nmax));
    return 5;
  }
  g.variant = variant;
  drop_graph (p);
  return 0;
}
