/*
 * lulesh_cbrt.h -- the cube root used by the hourglass coefficient.
 *
 * The reference takes CBRT(determ) through the C library
 * (reference/C/LULESH-OMP.cc:101-103, :1365; lulesh-2.0.3/lulesh.h:45-47,
 * lulesh-2.0.3/lulesh.cc:899).  The GPU has no glibc, and glibc's cbrt is not
 * correctly rounded, so no device routine can promise its exact bits.  This
 * header defines one cube root as a fixed sequence of IEEE-754 binary64
 * operations (+ - * and fma, each correctly rounded on x86-64 and on
 * gfx950), so the same source compiled for the host and for the device gives
 * identical bits.  It is correctly rounded except for roots within about
 * 1e-10 ulp of a rounding boundary; the tests check the rounding exactly and
 * bound the distance from the C library's cbrt.
 *
 * Method: split x = t * 2^(3q) with t = m * 2^r, m in [1, 2), r in {0, 1, 2}.
 * The reciprocal root z ~ t^(-1/3) starts from a cubic minimax fit of
 * m^(-1/3) on [1, 2] times 2^(-r/3) (relative error < 2.6e-4) and takes two
 * Newton steps z <- z + z (1 - t z^3) / 3 (error -> 1.3e-7 -> 3.4e-14).  Then
 * y = t z^2 is the root to about 7e-14, and one Newton step on y, whose
 * residual y^3 - t is formed with fma so that it is exact to within an ulp of
 * the residual and whose slope 1 / (3 y^2) is z^2 / 3, leaves an error near
 * 1e-26 before the final rounding.  No step divides.  The result is y * 2^q,
 * applied to the exponent field.
 *
 * The includer defines LULESH_FN (for example `static inline` on the host or
 * `static __device__ inline` in HIP code) before including this header.  It
 * may also define LULESH_CONST(c), which must return c unchanged: the GPU
 * build uses it to keep the constants in scalar registers at their use, so
 * that they do not occupy vector registers across the loops that inline the
 * root.
 */
#ifndef LULESH_CBRT_H
#define LULESH_CBRT_H

#include <stdint.h>

#ifndef LULESH_FN
#error "define LULESH_FN before including lulesh_cbrt.h"
#endif
#ifndef LULESH_CONST
#define LULESH_CONST(c) (c)
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

/* Cube root of a positive normal binary64 value (for other inputs it returns
 * some finite value). */
LULESH_FN double lulesh_cbrt_pos_normal(double x)
{
    const uint64_t b = lulesh_double_to_bits(x);
    const int e = (int)((b >> 52) & 0x7ffu) - 1023;
    /* q = floor(e / 3) for e in [-1022, 1023]; r = e - 3q in {0, 1, 2} */
    const int q = (int)((unsigned)(e + 1026) / 3u) - 342;
    const int r = e - 3 * q;
    const uint64_t frac = b & 0x000fffffffffffffull;
    const double m = lulesh_bits_to_double(frac | (1023ull << 52));
    const double t = lulesh_bits_to_double(frac | ((uint64_t)(1023 + r) << 52));
    const double third = LULESH_CONST(0.3333333333333333);
    double z = fma(m, fma(m, fma(m, LULESH_CONST(-0.04993701447458462),
                                 LULESH_CONST(0.3177574893441185)),
                          LULESH_CONST(-0.809960645949834)),
                   LULESH_CONST(1.5418870676993346));
    /* times 2^(-r/3) */
    {
        const double zc = z * ((r == 1) ? LULESH_CONST(0.7937005259840998)
                                        : LULESH_CONST(0.6299605249474366));
        z = (r == 0) ? z : zc;
    }
    for (int i = 0; i < 2; ++i) {
        const double d = fma(-t, (z * z) * z, 1.0); /* 1 - t z^3 */
        z = fma(z * d, third, z);
    }
    const double zz = z * z;
    double y = t * zz;
    {
        const double y2 = y * y;
        const double y2lo = fma(y, y, -y2);          /* y*y = y2 + y2lo exactly */
        const double res = fma(y, y2, -t) + y * y2lo; /* y^3 - t */
        y = fma(-res, zz * third, y);
    }
    return lulesh_bits_to_double(lulesh_double_to_bits(y) +
                                 ((uint64_t)(int64_t)q << 52));
}

/* Every input takes the same steps (selects, no branches): the root of |x|,
 * subnormals scaled by 2^54 into the normal range first and the root by
 * 2^-18 after (both exact), then x's sign; zeros, infinities and NaNs return
 * x + x, whatever the core computed for them. */
LULESH_FN double lulesh_cbrt(double x)
{
    const uint64_t b = lulesh_double_to_bits(x);
    const uint64_t ex = (b >> 52) & 0x7ffu;
    const uint64_t sign = b & 0x8000000000000000ull;
    const double ax = lulesh_bits_to_double(b ^ sign);
    const double s = (ex == 0) ? ax * 18014398509481984.0 : ax;
    double r = lulesh_cbrt_pos_normal(s);
    r = (ex == 0) ? r * 3.814697265625e-06 : r;
    r = lulesh_bits_to_double(lulesh_double_to_bits(r) | sign);
    return (ex == 0x7ffu || x == 0.0) ? x + x : r;
}

#endif /* LULESH_CBRT_H */
