// CPU check of the tuning cache (ops-fft_amd/lib/tune_cache.cpp): a plan-level choice written for
// one uneven plan is read back for the same groups and not for other groups; step choices are
// written beside it.  Built and run by tests/test_tune_cache.py against the library's registries.
#include <cstdio>
#include <cstring>

#include "tune_cache.h"

int main() {
    ofx::PlanDesc d;
    d.rank = 1;
    d.type = ACLFFT_C2C;
    d.uneven = true;
    d.ulen = {60, 1024, 4096};
    d.ubatch = {3, 7, 2};
    d.batch = 3;
    d.auto_precision = true;  // the tuner chooses the precision: the key names "auto"
    ofx::PlanDesc chosen = d;
    chosen.uneven_form = 1;
    chosen.sched = 0;
    chosen.blocks_per_cu = 3;
    chosen.class_set = 1;
    chosen.precision = 2;
    ofx::store_plan_choice(chosen);
    ofx::PlanDesc got;
    if (!ofx::load_plan_choice(d, &got)) {
        std::printf("no entry read back\n");
        return 1;
    }
    if (got.uneven_form != 1 || got.sched != 0 || got.blocks_per_cu != 3 || got.class_set != 1 || got.precision != 2) {
        std::printf("wrong entry: %d %d %d %d %d\n", got.uneven_form, got.sched, got.blocks_per_cu, got.class_set,
                    got.precision);
        return 2;
    }
    ofx::PlanDesc other = d;
    other.ubatch[1] = 8;
    if (ofx::load_plan_choice(other, &got)) {
        std::printf("entry found for other groups\n");
        return 3;
    }
    ofx::PlanDesc fixed = d;
    fixed.auto_precision = false;
    if (ofx::load_plan_choice(fixed, &got)) {
        std::printf("entry found for a fixed precision\n");
        return 4;
    }
    std::printf("ok\n");
    return 0;
}
