// E3M12: a 16-bit float code for the scaled population deviations
// s = k_d (f_d - w_d rho0) of TDLB_PREC_E3M12, host and device.
//
// IEEE binary16 spends 5 exponent bits on a range (2^-24 .. 65504) of which the
// deviations use 2^-9 .. 2^-1; E3M12 spends 3, and the 2 bits it saves lengthen
// the significand to 12 bits (13 significant), a quarter of fp16's relative
// step.  The code is a sign bit and a 15-bit magnitude q, the float bits of
// y = |s| + C (C = 2^-9) above those of C, rounded to nearest-even at 12
// fraction bits:
//
//     q = round((bits(|s| + C) - bits(C)) / 2^11),   s = +-(float(bits(C) + q 2^11) - C)
//
// so |s| >= C keeps 12 fraction bits (relative step 2^-12), |s| < C is linear
// with step C 2^-12 = 2^-21, zero is exact (the rest state is code 0) and the
// range is |s| < 255 C = 0.498 (f_d from -2 w_d rho0 to 4 w_d rho0 at rho0 = 1).
// The float sum |s| + C is rounded before the code is, so a value within 2^-12
// of a step from the midpoint of two codes may take the farther one (error at
// most (1/2 + 2^-12) step, unbiased); every other value takes the nearest.
// Magnitudes beyond the range, infinities and NaN saturate to +-0.498.  Every
// step is a few integer operations on the float bits (no table, no divide).
//
// The per-direction scale k = 2^L2K of the fused collision folds into the
// exponent: decode<L2K> returns s / k and encode<L2K>(v) is the code of v k,
// both exactly (C / k = 2^-(9 + L2K) is a normal float and scaling by a power
// of two commutes with the rounding of |v| + C / k).
#pragma once

#include <cmath>
#include <cstdint>
#include "lattice.h"

namespace tdlb {

TDLB_HD inline uint32_t f32_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
TDLB_HD inline float f32_from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

constexpr int kE3M12Frac = 12;       // fraction bits
constexpr int kE3M12OffsetExp = -9;  // C = 2^-9
constexpr uint16_t kE3M12MaxMag = 0x7fff;

// bits of C / 2^L2K
template <int L2K>
TDLB_HD constexpr uint32_t e3m12_base() {
    return (uint32_t)(127 + kE3M12OffsetExp - L2K) << 23;
}

// Codes are int16_t.  The decoder moves the sign bit straight to bit 31 and
// the encoder builds the code in the upper half of a 32-bit word (stored from
// there), a handful of integer operations each way.
template <int L2K>
TDLB_HD inline float e3m12_decode(int16_t h) {
    constexpr uint32_t base = e3m12_base<L2K>();
    const uint32_t w = (uint16_t)h;
    const float y = f32_from_bits(((w & kE3M12MaxMag) << (23 - kE3M12Frac)) + base);
    const float m = y - f32_from_bits(base);  // exact: y and C / k share the grid of C / k 2^-23
    return f32_from_bits(((w << 16) & 0x80000000u) | f32_bits(m));
}

template <int L2K>
TDLB_HD inline int16_t e3m12_encode(float v) {
    constexpr uint32_t base = e3m12_base<L2K>();
    constexpr int drop = 23 - kE3M12Frac;
    constexpr uint32_t rmax = ((uint32_t)kE3M12MaxMag << drop) | ((1u << drop) - 1u);
    const uint32_t yb = f32_bits(fabsf(v) + f32_from_bits(base));
    // round to nearest-even at bit `drop` of bits(y) - bits(C / k) (bits(C / k) has no bits below 23)
    uint32_t r = yb - base + ((1u << (drop - 1)) - 1u) + ((yb >> drop) & 1u);
    r = r > rmax ? rmax : r;  // saturate: magnitude code at most kE3M12MaxMag
    // magnitude code to bits 16..30, the sign of v to bit 31; the code is the upper half
    const uint32_t w = (r << (16 - drop)) | (f32_bits(v) & 0x80000000u);
    return (int16_t)(w >> 16);
}

}  // namespace tdlb
