// exact.h -- exact signs of the orient3d and insphere determinants for FP64
// inputs, for the opt-in robust predicate mode (the reference's predicates
// are the plain FP64 determinants, predicates.c:1687-1710 and 3301-3347, and
// the adaptive ones it ships, predicates.c:3940-4265, are never called).
//
// Every double is an integer times a power of two.  The 12 or 15 inputs of a
// test are put on the common grid of the smallest power present, as
// two's-complement integers of kBigLimbs 32-bit limbs, and the determinant is
// evaluated with the same expression as the FP64 one -- no rounding, so its
// sign is the sign of the exact determinant.  The width bounds the span of
// exponents a test may have (about 200 bits for insphere, 330 for orient3d);
// a wider test is reported as unresolved and keeps its FP64 sign.
//
// The limb loops are kept rolled (#pragma unroll 1) so that the numbers stay
// in private memory: the evaluation is rare, and unrolled it would hold
// hundreds of registers in every kernel that may call it.
//
// Static filters (Shewchuk, "Adaptive precision floating-point arithmetic and
// fast robust geometric predicates", 1997, the bounds behind predicates.c's
// o3derrboundA and isperrboundA): the FP64 sign is certain when |det| exceeds
// (7 + 56 eps) eps (orient3d) or (16 + 224 eps) eps (insphere) times the
// permanent, eps = 2^-53; only the other tests are evaluated exactly.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "geom.h"

namespace nat {

constexpr int kBigLimbs = 32;  // 1024-bit two's complement

struct Big {
    uint32_t w[kBigLimbs];
};

__host__ __device__ inline void big_zero(Big &r) {
    #pragma unroll 1
    for (int i = 0; i < kBigLimbs; i++) r.w[i] = 0u;
}

__host__ __device__ inline void big_add(Big &r, const Big &a, const Big &b) {
    uint64_t c = 0;
    #pragma unroll 1
    for (int i = 0; i < kBigLimbs; i++) {
        const uint64_t s = static_cast<uint64_t>(a.w[i]) + b.w[i] + c;
        r.w[i] = static_cast<uint32_t>(s);
        c = s >> 32;
    }
}

__host__ __device__ inline void big_sub(Big &r, const Big &a, const Big &b) {
    uint64_t c = 1;  // a + ~b + 1
    #pragma unroll 1
    for (int i = 0; i < kBigLimbs; i++) {
        const uint64_t s = static_cast<uint64_t>(a.w[i]) + static_cast<uint32_t>(~b.w[i]) + c;
        r.w[i] = static_cast<uint32_t>(s);
        c = s >> 32;
    }
}

// r = a * b modulo 2^(32 kBigLimbs); exact when the product fits (the callers
// check the magnitudes first).  r may not alias a or b.
__host__ __device__ inline void big_mul(Big &r, const Big &a, const Big &b) {
    big_zero(r);
    #pragma unroll 1
    for (int i = 0; i < kBigLimbs; i++) {
        if (a.w[i] == 0u) continue;
        uint64_t c = 0;
        #pragma unroll 1
        for (int j = 0; i + j < kBigLimbs; j++) {
            const uint64_t t =
                static_cast<uint64_t>(a.w[i]) * b.w[j] + r.w[i + j] + c;
            r.w[i + j] = static_cast<uint32_t>(t);
            c = t >> 32;
        }
    }
}

__host__ __device__ inline int big_sign(const Big &a) {
    if (a.w[kBigLimbs - 1] & 0x80000000u) return -1;
    #pragma unroll 1
    for (int i = 0; i < kBigLimbs; i++)
        if (a.w[i]) return 1;
    return 0;
}

// x = m 2^e with m an integer below 2^53 (e of the last mantissa bit);
// returns false for zero
__host__ __device__ inline bool dbl_split(double x, uint64_t &m, int &e, bool &neg) {
    uint64_t bits;
    memcpy(&bits, &x, sizeof(bits));
    neg = (bits >> 63) != 0;
    const int ex = static_cast<int>((bits >> 52) & 0x7ffu);
    m = bits & ((1ull << 52) - 1ull);
    if (ex == 0) {
        e = -1074;
    } else {
        m |= 1ull << 52;
        e = ex - 1075;
    }
    return m != 0;
}

// r = x / 2^E (an integer, E at most x's last-bit exponent)
__host__ __device__ inline void big_from_double(Big &r, double x, int E) {
    big_zero(r);
    uint64_t m;
    int e;
    bool neg;
    if (!dbl_split(x, m, e, neg)) return;
    const int sh = e - E;
    const int limb = sh >> 5, bit = sh & 31;
    const uint64_t lo = m << bit;                            // bits 0..63 of the shifted m
    const uint64_t hi = bit ? (m >> (64 - bit)) : 0ull;      // bits 64.. (m < 2^53)
    if (limb < kBigLimbs) r.w[limb] = static_cast<uint32_t>(lo);
    if (limb + 1 < kBigLimbs) r.w[limb + 1] = static_cast<uint32_t>(lo >> 32);
    if (limb + 2 < kBigLimbs) r.w[limb + 2] = static_cast<uint32_t>(hi);
    if (neg) {
        Big z;
        big_zero(z);
        Big t = r;
        big_sub(r, z, t);
    }
}

// Common grid of n doubles: E = the least last-bit exponent, and the bits
// their magnitudes need above it; false when all are zero.
__host__ __device__ inline bool grid_of(const double *v, int n, int &E, int &bits) {
    int lo = 1 << 20, hi = -(1 << 20);
    bool any = false;
    #pragma unroll 1
    for (int i = 0; i < n; i++) {
        uint64_t m;
        int e;
        bool neg;
        if (!dbl_split(v[i], m, e, neg)) continue;
        any = true;
        lo = e < lo ? e : lo;
        hi = e > hi ? e : hi;
    }
    E = lo;
    bits = any ? 53 + (hi - lo) : 0;
    return any;
}

// Exact sign of orient3d(a, b, c, d) (the expression of geom.h), or 2 when
// the exponents span too wide a range for kBigLimbs limbs.
__host__ __device__ inline int orient3d_exact_sign(const d3 &a, const d3 &b, const d3 &c,
                                                   const d3 &d) {
    const double v[12] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z};
    int E, bits;
    if (!grid_of(v, 12, E, bits)) return 0;
    const int L = bits + 1;  // differences
    if (3 * L + 4 > 32 * kBigLimbs - 1) return 2;
    Big p[12];
    #pragma unroll 1
    for (int i = 0; i < 12; i++) big_from_double(p[i], v[i], E);
    Big ad[3], bd[3], cd[3];
    #pragma unroll 1
    for (int k = 0; k < 3; k++) {
        big_sub(ad[k], p[k], p[9 + k]);
        big_sub(bd[k], p[3 + k], p[9 + k]);
        big_sub(cd[k], p[6 + k], p[9 + k]);
    }
    // adx (bdy cdz - bdz cdy) + bdx (cdy adz - cdz ady) + cdx (ady bdz - adz bdy)
    Big t1, t2, m, s, r;
    big_mul(t1, bd[1], cd[2]);
    big_mul(t2, bd[2], cd[1]);
    big_sub(m, t1, t2);
    big_mul(r, ad[0], m);
    big_mul(t1, cd[1], ad[2]);
    big_mul(t2, cd[2], ad[1]);
    big_sub(m, t1, t2);
    big_mul(t1, bd[0], m);
    big_add(s, r, t1);
    big_mul(t1, ad[1], bd[2]);
    big_mul(t2, ad[2], bd[1]);
    big_sub(m, t1, t2);
    big_mul(t1, cd[0], m);
    big_add(r, s, t1);
    return big_sign(r);
}

// Exact sign of insphere(a, b, c, d, e) (the expression of geom.h), or 2 when
// the exponents span too wide a range.
__host__ __device__ inline int insphere_exact_sign(const d3 &a, const d3 &b, const d3 &c,
                                                   const d3 &d, const d3 &e) {
    const double v[15] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z,
                          d.x, d.y, d.z, e.x, e.y, e.z};
    int E, bits;
    if (!grid_of(v, 15, E, bits)) return 0;
    const int L = bits + 1;
    if (5 * L + 8 > 32 * kBigLimbs - 1) return 2;
    // differences from e: [point][axis]
    Big x[4][3];
    {
        Big pe[3];
        #pragma unroll 1
        for (int k = 0; k < 3; k++) big_from_double(pe[k], v[12 + k], E);
        #pragma unroll 1
        for (int i = 0; i < 4; i++)
            #pragma unroll 1
            for (int k = 0; k < 3; k++) {
                Big p;
                big_from_double(p, v[3 * i + k], E);
                big_sub(x[i][k], p, pe[k]);
            }
    }
    Big t1, t2;
    // 2x2 minors in x, y: ab bc cd da ac bd
    auto minor = [&](Big &r, int i, int j) {  // ix jy - jx iy
        big_mul(t1, x[i][0], x[j][1]);
        big_mul(t2, x[j][0], x[i][1]);
        big_sub(r, t1, t2);
    };
    Big ab, bc, cd, da, ac, bd;
    minor(ab, 0, 1);
    minor(bc, 1, 2);
    minor(cd, 2, 3);
    minor(da, 3, 0);
    minor(ac, 0, 2);
    minor(bd, 1, 3);
    // abc = aez bc - bez ac + cez ab, bcd = bez cd - cez bd + dez bc,
    // cda = cez da + dez ac + aez cd, dab = dez ab + aez bd + bez da
    Big abc, bcd, cda, dab, s;
    big_mul(t1, x[0][2], bc);
    big_mul(t2, x[1][2], ac);
    big_sub(s, t1, t2);
    big_mul(t1, x[2][2], ab);
    big_add(abc, s, t1);
    big_mul(t1, x[1][2], cd);
    big_mul(t2, x[2][2], bd);
    big_sub(s, t1, t2);
    big_mul(t1, x[3][2], bc);
    big_add(bcd, s, t1);
    big_mul(t1, x[2][2], da);
    big_mul(t2, x[3][2], ac);
    big_add(s, t1, t2);
    big_mul(t1, x[0][2], cd);
    big_add(cda, s, t1);
    big_mul(t1, x[3][2], ab);
    big_mul(t2, x[0][2], bd);
    big_add(s, t1, t2);
    big_mul(t1, x[1][2], da);
    big_add(dab, s, t1);
    // lifts
    Big lift[4];
    #pragma unroll 1
    for (int i = 0; i < 4; i++) {
        big_mul(t1, x[i][0], x[i][0]);
        big_mul(t2, x[i][1], x[i][1]);
        big_add(s, t1, t2);
        big_mul(t1, x[i][2], x[i][2]);
        big_add(lift[i], s, t1);
    }
    // (dlift abc - clift dab) + (blift cda - alift bcd)
    Big u, w;
    big_mul(t1, lift[3], abc);
    big_mul(t2, lift[2], dab);
    big_sub(u, t1, t2);
    big_mul(t1, lift[1], cda);
    big_mul(t2, lift[0], bcd);
    big_sub(w, t1, t2);
    big_add(s, u, w);
    return big_sign(s);
}

// Static filter of orient3d: true when the FP64 sign is not certified.
__host__ __device__ __forceinline__ bool orient3d_uncertain(const d3 &a, const d3 &b,
                                                            const d3 &c, const d3 &d) {
    const double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    const double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    const double det = adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) +
                       cdx * (ady * bdz - adz * bdy);
    const double perm = (fabs(bdy * cdz) + fabs(bdz * cdy)) * fabs(adx) +
                        (fabs(cdy * adz) + fabs(cdz * ady)) * fabs(bdx) +
                        (fabs(ady * bdz) + fabs(adz * bdy)) * fabs(cdx);
    const double eps = 1.1102230246251565e-16;  // 2^-53
    return fabs(det) <= (7.0 + 56.0 * eps) * eps * perm;
}

}  // namespace nat
