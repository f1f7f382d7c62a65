// pgw_fold.h -- the exact reductions of the int8-limb path, one definition
// for the kernels and for their host checks.
//
// Residue x in [0,p) is written as a balanced base-256 number
//   x == y = sum_j d_j 256^j (mod p),  d_j in [-128, 127],  j < NL,
// with y = x if x <= MAXPOS else x - p (MAXPOS = 127 (256^NL - 1) / 255);
// every residue has such a representative because p <= 256^NL.  The MFMA
// sums digit products per shift s = j + l into int32 accumulators with
// |acc_s| <= 2^31 (a bound the host checks before a launch), and the
// epilogue folds sum_s acc_s 256^s back to a residue.
#pragma once

#include "pgw_internal.h"

namespace pgw {

constexpr uint32_t kM31 = 0x7fffffffu;  // 2^31 - 1

struct DigitSpec {
    int64_t maxpos;
    uint64_t bias;
};

__host__ __device__ inline DigitSpec digit_spec(uint32_t nl) {
    return nl == 2 ? DigitSpec{32639, 0x8080ULL} : DigitSpec{2139062143LL, 0x80808080ULL};
}

// bytes d_j + 128 of the representative; XOR 0x80 per byte gives the int8 digit
__host__ __device__ inline uint64_t digits_biased(uint32_t x, uint32_t p, const DigitSpec& ds) {
    int64_t y = (int64_t)x <= ds.maxpos ? (int64_t)x : (int64_t)x - (int64_t)p;
    return (uint64_t)(y + (int64_t)ds.bias);
}

// (f & M31) + (f >> 31) == f (mod 2^31 - 1), since 2^31 == 1
__host__ __device__ inline uint64_t fold31(uint64_t f) { return (f & 0x7fffffffULL) + (f >> 31); }

// v mod (2^31 - 1) for v < 2^63: the first fold leaves < 2^31 + 2^32, the
// second < 2^31 + 2, so one conditional subtraction finishes
__host__ __device__ inline uint32_t red63_m31(uint64_t v) {
    const uint64_t r = fold31(fold31(v));
    return (uint32_t)(r >= kM31 ? r - kM31 : r);
}

// sum_s x[s] 256^s mod p for the NS = 2 NL - 1 shift sums, |x[s]| <= 2^31:
//  * p = 2^31 - 1 (NL = 4): 256^s == 2^(8s mod 31), so sum_s x[s] 2^(8s mod 31)
//    is the value mod p and is exact in 64 bits: |sum| <= 2^31 (1 + 2 + 2^8 +
//    2^9 + 2^16 + 2^17 + 2^24) < 2^56 - 2^25 = p 2^25; offset by p 2^25 it lies
//    in (0, 2^57) and two folds leave at most p + 1;
//  * NL = 2 (p <= 65536): sum_s x[s] 256^s is exact, |sum| < 2^48 <= p 2^47;
//    offset by p 2^47 (< 2^64) it is one 64-bit Barrett reduction;
//  * otherwise: each shift reduced from x[s] + p 2^32 > 0, times
//    w[s] = 256^s mod p (< 2^62 each), summed four and three at a time.
template <int NL>
__host__ __device__ inline uint32_t fold_digit_shifts(const int32_t* x, const ModP& m, const uint32_t* w) {
    constexpr int NS = 2 * NL - 1;
    const uint32_t p = m.p;
    if constexpr (NL == 2) {
        uint64_t t = (uint64_t)p << 47;
#pragma unroll
        for (int s = 0; s < NS; ++s) t += (uint64_t)((int64_t)x[s] * ((int64_t)1 << (8 * s)));
        return mod_u64(t, m);
    } else {
        if (p == kM31) {
            uint64_t t = (uint64_t)p << 25;
#pragma unroll
            for (int s = 0; s < NS; ++s) t += (uint64_t)((int64_t)x[s] * ((int64_t)1 << ((8 * s) % 31)));
            const uint32_t v = (uint32_t)fold31(fold31(t));
            return v >= p ? v - p : v;
        }
        const uint64_t off = (uint64_t)p << 32;  // a multiple of p above 2^31
        uint64_t s_lo = 0, s_hi = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint32_t r = mod_u64((uint64_t)((int64_t)x[s] + (int64_t)off), m);
            const uint64_t prod = (uint64_t)r * w[s];  // < 2^62
            if (s < 4) s_lo += prod; else s_hi += prod;
        }
        return addmod(mod_u64(s_lo, m), mod_u64(s_hi, m), p);
    }
}

}  // namespace pgw
