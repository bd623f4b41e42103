// k_invert.hip -- placeholder until the inverse kernels land.
#include "hipla_kernels.h"
#include "hipla_plan.h"

int plan_enqueue_invert (hipla_plan, const double *, double *, int *, hipla_cell_stats *, hipStream_t)
{
  hipla_set_detail ("invert not available");
  return 5;
}
hipError_t setup_invert_attrs (size_t)
{
  return hipSuccess;
}
extern "C" int hipla_adjoint (hipla_plan, const double *, const double *, double *, double *, void *)
{
  hipla_set_detail ("adjoint not available");
  return 5;
}
