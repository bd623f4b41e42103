// k_mixed.hip -- placeholder until the mixed-precision kernels land.
#include "hipla_kernels.h"
#include "hipla_plan.h"
#include "hipla_tune.h"

void route_mixed (int n, int &kind, int &param)
{
  kind = KK_LDS_MIX;
  param = n;
}
int64_t blocked_mixed_ws_floats (int nmax)
{
  return (int64_t) nmax * nmax;
}
int plan_enqueue_mixed (hipla_plan, const double *, const double *, double *, int *, hipla_cell_stats *, hipStream_t)
{
  hipla_set_detail ("mixed mode not available");
  return 5;
}
hipError_t setup_mixed_attrs (size_t)
{
  return hipSuccess;
}
