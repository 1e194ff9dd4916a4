// sheen_math.h — float transcendentals of the shading, light-sampling,
// media and photon paths: sin, cos, tan, exp, log2 and pow.
//
// Each is evaluated in double from a short series with every operation
// exactly rounded (compiled with contraction off), then rounded once to
// float.  The series errors are below 1e-15 relative, so the float result is
// the correctly rounded value except when the true value lies within that
// distance of a float rounding boundary; and because the same operations run
// on the GPU and on the host, the HIP path and the CPU oracle get the same
// bits whatever maths library either links.  (The reference calls the
// single-precision functions of CUDA's library, which no other library
// reproduces to the last bit.)
//
// Methods:
//   sin, cos:  x = k pi/2 + r (Cody-Waite, pi/2 split in a 33-bit head and a
//              double tail, exact for |k| < 2^20), Taylor series of sin and
//              cos on |r| <= pi/4 to r^17 / r^18, quadrant by k mod 4;
//   tan:       sin / cos of the same reduction;
//   log2:      x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//              ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^23/23);
//   exp2:      z = k + f, |f| <= 1/2, 2^f = e^(f ln 2) by Taylor to 15 terms,
//              scaled by 2^k through the exponent bits;
//   exp:       exp2(x log2(e));  pow (x >= 0): exp2(y log2(x)).
// Domains: the callers pass angles and 2 pi u (|x| < 2^20 pi/2), x >= 0 to
// pow and log2, and exp arguments of any size.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SHM_FN __host__ __device__ inline
#else
#define SHM_FN inline
#endif

namespace shm {

SHM_FN double bits_to_d(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
SHM_FN uint64_t d_to_bits(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
}

// 2^k for an integer k in [-1022, 1023] (a normal double).
SHM_FN double pow2i(int k) { return bits_to_d((uint64_t)(k + 1023) << 52); }

// 2^z.
SHM_FN double exp2_d(double z) {
    if (z != z) return z;
    if (z > 1100.0) return bits_to_d(0x7FF0000000000000ull);
    if (z < -1100.0) return 0.0;
    const double k = __builtin_floor(z + 0.5);
    const double f = z - k;                     // |f| <= 1/2, exact
    const double t = f * 0.69314718055994530942;  // f ln 2
    double p = 1.0 / 1307674368000.0;             // 1/15!
    p = p * t + 1.0 / 87178291200.0;
    p = p * t + 1.0 / 6227020800.0;
    p = p * t + 1.0 / 479001600.0;
    p = p * t + 1.0 / 39916800.0;
    p = p * t + 1.0 / 3628800.0;
    p = p * t + 1.0 / 362880.0;
    p = p * t + 1.0 / 40320.0;
    p = p * t + 1.0 / 5040.0;
    p = p * t + 1.0 / 720.0;
    p = p * t + 1.0 / 120.0;
    p = p * t + 1.0 / 24.0;
    p = p * t + 1.0 / 6.0;
    p = p * t + 0.5;
    p = p * t + 1.0;
    p = p * t + 1.0;
    int ki = (int)k;
    // scale in two steps so that results near the ends of the range stay exact
    const int k1 = ki / 2, k2 = ki - k1;
    return p * pow2i(k1) * pow2i(k2);
}

// log2(x) for finite x > 0.
SHM_FN double log2_d(double x) {
    uint64_t b = d_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0) {  // subnormal: scale up by 2^64 first
        x = x * 18446744073709551616.0;
        b = d_to_bits(x);
        e = (int)((b >> 52) & 0x7FF) - 64;
    }
    e -= 1023;
    double m = bits_to_d((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);  // [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);  // |s| < 0.1716
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double ln_m = 2.0 * s * p;
    return (double)e + ln_m * 1.44269504088896340736;  // 1/ln 2
}

// x = k pi/2 + r with |r| <= pi/4 (about); q = k mod 4.
SHM_FN double reduce_pio2(double x, int& q) {
    const double k = __builtin_floor(x * 0.63661977236758134308 + 0.5);  // 2/pi
    const double hi = 1.5707963267341256;       // pi/2 to 33 bits: k * hi is exact
    const double lo = 6.077100506506192e-11;    // pi/2 - hi
    const double r = (x - k * hi) - k * lo;
    q = (int)((int64_t)k & 3);
    return r;
}
SHM_FN double sin_poly(double r) {
    const double r2 = r * r;
    double p = 1.0 / 355687428096000.0;  // 1/17!
    p = -p * r2 + 1.0 / 1307674368000.0;
    p = -p * r2 + 1.0 / 6227020800.0;
    p = -p * r2 + 1.0 / 39916800.0;
    p = -p * r2 + 1.0 / 362880.0;
    p = -p * r2 + 1.0 / 5040.0;
    p = -p * r2 + 1.0 / 120.0;
    p = -p * r2 + 1.0 / 6.0;
    p = -p * r2 + 1.0;
    return p * r;  // r - r^3/3! + ... + r^17/17!
}
SHM_FN double cos_poly(double r) {
    const double r2 = r * r;
    double p = 1.0 / 6402373705728000.0;  // 1/18!
    p = -p * r2 + 1.0 / 20922789888000.0;
    p = -p * r2 + 1.0 / 87178291200.0;
    p = -p * r2 + 1.0 / 479001600.0;
    p = -p * r2 + 1.0 / 3628800.0;
    p = -p * r2 + 1.0 / 40320.0;
    p = -p * r2 + 1.0 / 720.0;
    p = -p * r2 + 1.0 / 24.0;
    p = -p * r2 + 0.5;
    p = -p * r2 + 1.0;
    return p;  // 1 - r^2/2! + ... + r^18/18!
}
SHM_FN double sin_d(double x) {
    int q;
    const double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}
SHM_FN double cos_d(double x) {
    int q;
    const double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}
SHM_FN double tan_d(double x) {
    int q;
    const double r = reduce_pio2(x, q);
    const double s = sin_poly(r), c = cos_poly(r);
    return (q & 1) ? -c / s : s / c;
}

// Float entry points: the double value rounded once.
SHM_FN float sinf_r(float x) { return (float)sin_d((double)x); }
SHM_FN float cosf_r(float x) { return (float)cos_d((double)x); }
SHM_FN float tanf_r(float x) { return (float)tan_d((double)x); }
SHM_FN float expf_r(float x) {
    return (float)exp2_d((double)x * 1.44269504088896340736);  // x log2(e)
}
SHM_FN float log2f_r(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    return (float)log2_d((double)x);
}
// x^y for x >= 0 (the callers clamp), y finite.
SHM_FN float powf_r(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    if (x == 1.0f) return 1.0f;
    return (float)exp2_d((double)y * log2_d((double)x));
}

}  // namespace shm
