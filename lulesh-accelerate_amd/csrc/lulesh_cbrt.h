/*
 * lulesh_cbrt.h -- the cube root used by the hourglass coefficient.
 *
 * The reference takes CBRT(determ) through the C library
 * (reference/C/LULESH-OMP.cc:101-103, :1365; lulesh-2.0.3/lulesh.h:45-47,
 * lulesh-2.0.3/lulesh.cc:899).  The GPU has no glibc, and glibc's cbrt is not
 * correctly rounded, so no device routine can promise its exact bits.  This
 * header defines one cube root as a fixed sequence of IEEE-754 binary64
 * operations (+ - * / and fma, each correctly rounded on x86-64 and on
 * gfx950), so the same source compiled for the host and for the device gives
 * identical bits.  It is accurate to about half an ulp; the tests bound its
 * distance from the C library's cbrt.
 *
 * Method: split x = t * 2^(3q) with t in [1, 8); a cubic least-squares fit of
 * t^(1/3) (relative error < 1.4e-2), three Newton steps y <- (2y + t/y^2)/3
 * (error ~1e-2 -> 2e-4 -> 4e-8 -> 2e-15), then one Newton step whose residual
 * y^3 - t is formed with fma so that it is exact to within an ulp of the
 * residual.  The result is y * 2^q, applied to the exponent field.
 *
 * The includer defines LULESH_FN (for example `static inline` on the host or
 * `static __device__ inline` in HIP code) before including this header.
 */
#ifndef LULESH_CBRT_H
#define LULESH_CBRT_H

#include <stdint.h>

#ifndef LULESH_FN
#error "define LULESH_FN before including lulesh_cbrt.h"
#endif

LULESH_FN double lulesh_bits_to_double(uint64_t b)
{
    union { uint64_t u; double d; } c;
    c.u = b;
    return c.d;
}

LULESH_FN uint64_t lulesh_double_to_bits(double d)
{
    union { uint64_t u; double d; } c;
    c.d = d;
    return c.u;
}

/* Cube root of a positive normal binary64 value. */
LULESH_FN double lulesh_cbrt_pos_normal(double x)
{
    const uint64_t b = lulesh_double_to_bits(x);
    const int e = (int)((b >> 52) & 0x7ffu) - 1023;
    /* q = floor(e / 3), r = e - 3q in {0, 1, 2} */
    const int q = (e >= 0) ? (e / 3) : -((2 - e) / 3);
    const int r = e - 3 * q;
    const double t = lulesh_bits_to_double((b & 0x000fffffffffffffull) |
                                           ((uint64_t)(1023 + r) << 52));
    double y = 0.7091350135760441 +
               t * (0.3400187424911607 +
                    t * (-0.037495962996826254 + t * 0.0019167009490987084));
    y = (2.0 * y + t / (y * y)) / 3.0;
    y = (2.0 * y + t / (y * y)) / 3.0;
    y = (2.0 * y + t / (y * y)) / 3.0;
    {
        const double y2 = y * y;
        const double y2lo = fma(y, y, -y2);          /* y*y = y2 + y2lo exactly */
        const double res = fma(y, y2, -t) + y * y2lo; /* y^3 - t */
        y = y - res / (3.0 * y2);
    }
    return lulesh_bits_to_double(lulesh_double_to_bits(y) +
                                 ((uint64_t)(int64_t)q << 52));
}

LULESH_FN double lulesh_cbrt(double x)
{
    const uint64_t b = lulesh_double_to_bits(x);
    const uint64_t ex = (b >> 52) & 0x7ffu;
    if (ex != 0 && ex != 0x7ffu) {
        if (x > 0.0) {
            return lulesh_cbrt_pos_normal(x);
        }
        return -lulesh_cbrt_pos_normal(-x);
    }
    if (ex == 0x7ffu || x == 0.0) {
        return x + x; /* inf, nan and signed zero pass through */
    }
    /* subnormal: scale by 2^54 (exact), take the root, scale by 2^-18 */
    {
        const double s = x * 18014398509481984.0;
        const double r = (s > 0.0) ? lulesh_cbrt_pos_normal(s)
                                   : -lulesh_cbrt_pos_normal(-s);
        return r * 3.814697265625e-06;
    }
}

#endif /* LULESH_CBRT_H */
