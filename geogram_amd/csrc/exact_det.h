// Exact sign of a 4x4 determinant of doubles, by fixed-point integer
// arithmetic: every double is m * 2^e with a 53-bit integer m, each of the 24
// terms of the Leibniz expansion is a 212-bit integer times a power of two,
// and the terms are accumulated exactly in a 512-bit two's-complement
// integer aligned on the smallest exponent.  Used by the exact re-clip of
// cells whose floating-point sign tests were not certified; this is the role
// PCK::det_4d's expansion fallback plays for VBW::ConvexCell
// (numerics/predicates.cpp:2073-2110).
#pragma once

#include <stdint.h>

namespace vgc {

struct DSplit {
    uint64_t m;  // mantissa, < 2^53 (0 for a zero)
    int e;       // value = +/- m * 2^e
    int neg;
};

VGC_SD DSplit dsplit(double v) {
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    DSplit r;
    int bexp = int((b >> 52) & 0x7FF);
    uint64_t frac = b & ((uint64_t(1) << 52) - 1);
    if (bexp == 0) {
        r.m = frac;
        r.e = -1074;
    } else {
        r.m = frac | (uint64_t(1) << 52);
        r.e = bexp - 1075;
    }
    r.neg = int(b >> 63);
    return r;
}

constexpr int kExactLimbs = 8;

// 64x64 -> 128 bit product
VGC_SD void mul64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    unsigned __int128 p = (unsigned __int128)a * b;
    lo = (uint64_t)p;
    hi = (uint64_t)(p >> 64);
}

// out[0..na] = a[0..na-1] * b
VGC_SD void mul_limbs(const uint64_t* a, int na, uint64_t b, uint64_t* out) {
    uint64_t carry = 0;
    for (int i = 0; i < na; ++i) {
        uint64_t lo, hi;
        mul64(a[i], b, lo, hi);
        uint64_t s = lo + carry;
        hi += (s < lo) ? 1 : 0;
        out[i] = s;
        carry = hi;
    }
    out[na] = carry;
}

// acc += sign * (p << sh), p has 4 limbs, acc kExactLimbs limbs (two's complement)
VGC_SD void acc_add_shifted(uint64_t* acc, const uint64_t* p, int sh, bool negate) {
    uint64_t t[kExactLimbs];
    for (int i = 0; i < kExactLimbs; ++i) t[i] = 0;
    const int L = sh >> 6, r = sh & 63;
    for (int i = 0; i < 4; ++i) {
        if (i + L < kExactLimbs) t[i + L] |= p[i] << r;
        if (r && i + L + 1 < kExactLimbs) t[i + L + 1] |= p[i] >> (64 - r);
    }
    if (negate) {
        // two's complement of t
        uint64_t c = 1;
        for (int i = 0; i < kExactLimbs; ++i) {
            uint64_t v = ~t[i] + c;
            c = (c && v == 0) ? 1 : 0;
            t[i] = v;
        }
    }
    uint64_t c = 0;
    for (int i = 0; i < kExactLimbs; ++i) {
        uint64_t s = acc[i] + t[i];
        uint64_t c1 = (s < acc[i]) ? 1 : 0;
        uint64_t s2 = s + c;
        uint64_t c2 = (s2 < s) ? 1 : 0;
        acc[i] = s2;
        c = c1 | c2;
    }
}

// Sign (-1, 0, 1) of det[r0; r1; r2; r3] (rows of 4), or 2 when the exponent
// spread of the terms exceeds the accumulator.
VGC_SD int det4_sign_exact(const double* r0, const double* r1, const double* r2,
                           const double* r3) {
    DSplit M[4][4];
    for (int c = 0; c < 4; ++c) {
        M[0][c] = dsplit(r0[c]);
        M[1][c] = dsplit(r1[c]);
        M[2][c] = dsplit(r2[c]);
        M[3][c] = dsplit(r3[c]);
    }
    // the 24 permutations of (0,1,2,3) in lexicographic order, with parity
    const uint8_t perm[24][4] = {
        {0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {0, 3, 2, 1},
        {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 0, 2}, {1, 3, 2, 0},
        {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 3, 0, 1}, {2, 3, 1, 0},
        {3, 0, 1, 2}, {3, 0, 2, 1}, {3, 1, 0, 2}, {3, 1, 2, 0}, {3, 2, 0, 1}, {3, 2, 1, 0}};
    const int8_t par[24] = {1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1,
                            1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1};
    int emin = 1 << 30, emax = -(1 << 30);
    for (int k = 0; k < 24; ++k) {
        const DSplit &a = M[0][perm[k][0]], &b = M[1][perm[k][1]], &c = M[2][perm[k][2]],
                     &d = M[3][perm[k][3]];
        if (!a.m || !b.m || !c.m || !d.m) continue;
        int e = a.e + b.e + c.e + d.e;
        emin = e < emin ? e : emin;
        emax = e > emax ? e : emax;
    }
    if (emin > emax) return 0;  // every term vanishes
    if (emax - emin > 64 * kExactLimbs - 212 - 6) return 2;
    uint64_t acc[kExactLimbs];
    for (int i = 0; i < kExactLimbs; ++i) acc[i] = 0;
    for (int k = 0; k < 24; ++k) {
        const DSplit &a = M[0][perm[k][0]], &b = M[1][perm[k][1]], &c = M[2][perm[k][2]],
                     &d = M[3][perm[k][3]];
        if (!a.m || !b.m || !c.m || !d.m) continue;
        uint64_t p1[2], p2[3], p3[4];
        uint64_t lo, hi;
        mul64(a.m, b.m, lo, hi);
        p1[0] = lo;
        p1[1] = hi;
        mul_limbs(p1, 2, c.m, p2);
        mul_limbs(p2, 3, d.m, p3);
        bool neg = ((a.neg ^ b.neg ^ c.neg ^ d.neg) != 0) != (par[k] < 0);
        acc_add_shifted(acc, p3, a.e + b.e + c.e + d.e - emin, neg);
    }
    bool zero = true;
    for (int i = 0; i < kExactLimbs; ++i) zero = zero && acc[i] == 0;
    if (zero) return 0;
    return (acc[kExactLimbs - 1] >> 63) ? -1 : 1;
}

}  // namespace vgc
