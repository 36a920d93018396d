// Exact unsigned division by a divisor fixed before a launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plb {

// Division by a divisor fixed for a launch, as a multiply-high and two shifts
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant): exact for every 32-bit
// dividend and divisors up to 2^31.  On block-uniform values it stays on the
// scalar unit instead of the vector unit's emulated division.
struct FastDiv {
    uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}

__device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

}  // namespace plb
