// Resident block-step plans measured on MI355X by
// `python -m hermite_amd.autotune --plan c1 c2 c3 c4 c5` (hermite_amd/tuning/mi355x_plan.json):
// order, total particles, clusters, then hhr_set_plan's and hhr_set_split_waves'
// parameters.  hhr_create takes the row of its order and kind (one cluster or an
// ensemble) nearest in total particles, within a factor of 4.  Only plans that beat
// the built-in one by 1.02x or more are rows; for these the built-in plan
// was kept (within run-to-run noise of the best found):
//   order 4, N 1024, 1 cluster(s) (c1, fp64): best {128, 512, 8192, 256, 8192, 16384} measured 1.003x the built-in plan
//   order 4, N 549993, 256 cluster(s) (c4, fp64): best {64, 128, 8192, 512, 8192, 8192} measured 1.003x the built-in plan
//   order 4, N 549993, 256 cluster(s) (c4, split): best {64, 128, 8192, 512, 16384, 32768} measured 1.003x the built-in plan
//   order 6, N 65536, 1 cluster(s) (c2, fp64): best {128, 512, 8192, 256, 8192, 16384} measured 1.008x the built-in plan
//   order 6, N 1048576, 1 cluster(s) (c5, fp64): best {64, 128, 8192, 256, 8192, 16384} measured 1.002x the built-in plan
//   order 8, N 262144, 1 cluster(s) (c3, fp64): best {128, 512, 8192, 256, 8192, 16384} measured 1.002x the built-in plan
#pragma once

namespace hh {

struct PlanRow {
    int order, n_total, nclu;
    int v[6];  // jpar_max smax task_target jpar_target wj4_max wj2_max
};

static const PlanRow kPlanTable[] = {
    {0, 0, 0, {0, 0, 0, 0, 0, 0}},  // end of table
};

}  // namespace hh
