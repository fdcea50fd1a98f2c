// Host-only check of the plan builders for the sanitizer test (tests/test_host_sanitize.py): for
// each degree given, build the plan, the reverse plan and the fused reverse step's tables, and
// validate the tables against the plan's transitions.
#include <cstdio>
#include <cstdarg>
#include <vector>
#include "plan_host.hpp"
using namespace eggman;
int eggman::fail(int code, const char *fmt, ...) { va_list a; va_start(a, fmt); vfprintf(stderr, fmt, a); va_end(a); fprintf(stderr, "\n"); return code; }
int main(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
        int K = atoi(argv[i]);
        PlanLayout lay; if (plan_layout(K, 4, &lay)) return 1;
        std::vector<int32_t> starts(lay.nstates + 1), trans(lay.ntrans), levels(lay.nlevels);
        build_plan_host(lay, starts.data(), trans.data(), levels.data(), 8);
        RevPlanHost rp; if (build_rev_plan(lay, trans.data(), 256, 128, 64, &rp)) return 1;
        FusePlanHost fp; if (build_fuse_plan(lay, rp, 64, &fp)) return 1;
        int64_t e = 0, t = 0;
        int rc = validate_fuse_plan(lay, trans.data(), rp, fp, &e, &t);
        printf("K=%d rc=%d entries=%ld terms=%ld ntrans=%ld\n", K, rc, (long)e, (long)t, (long)lay.ntrans);
        if (rc) return 2;
    }
    return 0;
}
