// fastdiv.h -- division by a divisor fixed at launch as a multiply-high and
// a shift (host and device; the CPU tests build it with g++ alone).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

namespace lbm {

// Division by a wave-uniform divisor fixed at launch, as a multiply-high and a
// shift on the scalar unit (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, section 4: l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(m, n) + n) >> l with the
// sum in 33 bits).  A plain division by a run-time value compiles to a
// float-reciprocal sequence on the vector unit, several instructions per
// wavefront in the 16-bit kernels, which are bound by the vector unit.
struct FastDiv {
    uint32_t d, m, l;
    // d >= 1 (the callers' divisors are row and segment counts)
    static FastDiv of(uint32_t d)
    {
        FastDiv f;
        f.d = d;
        f.l = 0;
        while ((1ull << f.l) < d) ++f.l;
        f.m = (uint32_t)((((1ull << f.l) - d) << 32) / d + 1);
        return f;
    }
    LBM_HD uint32_t div(uint32_t n) const
    {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t t = __umulhi(m, n);
#else
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
#endif
        return (uint32_t)(((uint64_t)t + n) >> l);
    }
};

} // namespace lbm
