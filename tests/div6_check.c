/* Exhaustive check, over all 2^32 float bit patterns, of the exact division
 * by 6 that the Jacobi kernels use (boiling-sim_amd/csrc/jacobi.hip, div6):
 * fma(x, zh, x*zl) with zh = RN(1/6), zl = RN(1/6 - zh) for finite
 * |x| >= 2^-96, the IEEE quotient otherwise.  Built and run by
 * tests/test_div6.py; every operation is an IEEE single-precision one (the
 * fused multiply-add through the hardware FMA, -ffp-contract=off for the
 * rest), so it checks the identity itself, not a compiler's rendering. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline float as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t as_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static inline int same(float a, float b) {
    if (isnan(a) && isnan(b)) return 1;
    return as_bits(a) == as_bits(b);
}

/* counts[0]: div6 differs from x / 6.0f (must be 0)
 * counts[1]: x reached the fma path
 * counts[2]: the one-term form x * zh differs from x / 6.0f
 * counts[3]: the fma form without the |x| >= 2^-96 guard differs */
__attribute__((target("fma")))
void div6_check(long long* counts) {
    const float zh = 0x1.555556p-3f, zl = -0x1.555556p-28f;
    long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma omp parallel for schedule(static) reduction(+ : c0, c1, c2, c3)
    for (int64_t i = 0; i < (int64_t)1 << 32; ++i) {
        const float x = as_float((uint32_t)i);
        const float q = x / 6.0f;
        const float ax = fabsf(x);
        const float two = __builtin_fmaf(x, zh, x * zl);
        float d;
        if (ax >= 0x1p-96f && ax <= 0x1.fffffep127f) {
            d = two;
            ++c1;
        } else {
            d = q;
        }
        c0 += !same(d, q);
        c2 += !same(x * zh, q);
        c3 += !same(two, q);
    }
    counts[0] = c0;
    counts[1] = c1;
    counts[2] = c2;
    counts[3] = c3;
}
