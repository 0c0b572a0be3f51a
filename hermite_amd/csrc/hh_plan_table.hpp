// Resident block-step plans measured on MI355X by
// `python -m hermite_amd.autotune --plan c1 c2 c3 c4 c5` (hermite_amd/tuning/mi355x_plan.json):
// order, total particles, clusters, then hhr_set_plan's and hhr_set_split_waves'
// parameters.  hhr_create takes the row of its order and kind (one cluster or an
// ensemble) nearest in total particles, within a factor of 4.
#pragma once

namespace hh {

struct PlanRow {
    int order, n_total, nclu;
    int v[6];  // jpar_max smax task_target jpar_target wj4_max wj2_max
};

static const PlanRow kPlanTable[] = {
    {4, 1024, 1, {128, 256, 8192, 256, 8192, 8192}},  // c1, fp64: 1.008x the built-in plan
    {4, 549993, 256, {64, 128, 8192, 512, 8192, 32768}},  // c4, fp64: 1.004x the built-in plan
    {6, 65536, 1, {128, 512, 8192, 256, 8192, 16384}},  // c2, fp64: 1.013x the built-in plan
    {6, 1048576, 1, {64, 128, 8192, 256, 8192, 16384}},  // c5, fp64: 1.002x the built-in plan
    {8, 262144, 1, {128, 256, 8192, 256, 16384, 16384}},  // c3, fp64: 1.002x the built-in plan
};

}  // namespace hh
