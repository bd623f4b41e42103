// hipla_tune.h -- size-class routing and autotuned tiles for the mixed-precision solver.
#pragma once

#include <stdint.h>

#include "hipla_plan.h"

void route_mixed (int n, int &kind, int &param);
int64_t blocked_mixed_ws_floats (int nmax);
int plan_enqueue_mixed (hipla_plan p, const double *dA, const double *dB, double *dX, int *dcodes,
                        hipla_cell_stats *dstats, hipStream_t st);
