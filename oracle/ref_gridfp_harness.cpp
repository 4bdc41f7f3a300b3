// ref_gridfp_harness.cpp -- exports the REFERENCE's own Grid-FP transition
// functions (compiled from /root/reference/src/common/*.hpp where they lie,
// by oracle/ref.mk) through a C ABI, so tests can pin this repository's
// oracle against them.  TEST INFRASTRUCTURE ONLY; output goes to oracle/_ref/.
#include <cstdint>
#include "gridfp_reverse.hpp"
#include "gridfp_transition.hpp"

using namespace oneesan::gridfp;

extern "C" {
int ref_include(uint64_t m, int width, int p, uint64_t* out, int* blocked) {
    IncludeResult z = include_horizontal(m, width, p);
    *out = z.mate;
    *blocked = z.blocked ? 1 : 0;
    return z.valid ? 1 : 0;
}
uint64_t ref_blocked_exclude(uint64_t b, int p) { return blocked_exclude(b, p); }
int ref_reverse_block_preds(uint64_t b, int width, int p, uint64_t* out, int cap) {
    int n = 0;
    reverse_block_predecessors<true>(b, width, p, [&](MateID m) { if (n < cap) out[n] = m; ++n; });
    return n;
}
int ref_reverse_boundary_preds(uint64_t t, int width, uint64_t* out, int cap) {
    int n = 0;
    reverse_boundary_main_predecessors(t, width, [&](MateID m) { if (n < cap) out[n] = m; ++n; });
    return n;
}
}
