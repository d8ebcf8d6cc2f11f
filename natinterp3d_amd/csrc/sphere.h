// sphere.h -- a certified shortcut for the FP64 in-sphere decision of the
// cavity search (natural_insertionfree.c:498-530 tests each popped tet with
// the plain insphere determinant, predicates.c:3301-3347).
//
// For a positively oriented tet (a, b, c, d) with circumcentre c* and
// circumradius r, the exact lifted determinant factors as
//     insphere(a, b, c, d, e) = orient3d(a, b, c, d) * (r^2 - |e - c*|^2)
// (column operations on the lifted matrix; orient3d as in geom.h).  Per tet
// the mesh stores an FP64 circumcentre c~, R2 = |a - c~|^2 and two
// constants; a query then computes m = R2 - |e - c~|^2 and s = R2 +
// |e - c~|^2 (about 20 operations instead of 77) and the sign of m is the
// sign of the FP64 determinant whenever
//   (i)   |m| > 3 x 8 eps s                   rounding of m and s,
//   (ii)  |m| > 3 x 2 sqrt(4.01 s) delta      delta >= |c* - c~|, from
//         r^2 - |e - c*|^2 = R2 - |e - c~|^2 - 2 (a - e).(c* - c~) and
//         |a - e| <= r + |e - c*| <= sqrt(4.01 s),
//   (iii) |m| > 3 x 2 (16 + 224 eps) eps x 72 (4.01 s)^2.5 / O_lb:
//         Shewchuk's bound on the FP64 determinant's error is
//         (16 + 224 eps) eps x its permanent (doubled here), the permanent
//         is at most 72 L^5 with L = max |v - e| <= r + |e - c*|, and
//         O_lb <= orient3d: where |r^2 - |e - c*|^2| exceeds that bound
//         divided by the orientation, the FP64 determinant has the exact
//         sign, which is the sign of m.
// Each condition takes a third of |m|, so together they cover the sum of the
// three error terms; they are tested squared (no square root):
//   m^2 > 1e-28 s^2,  m^2 > d1 s  (d1 = 160 delta^2),  m^2 w > s^5
//   (w = 1e21 O_lb^2 >= O_lb^2 / (9 x 6.55e-26 x 4.01^5)).
// delta is bounded a posteriori: with u_i = v_i - a and w~ = c~ - a, the
// exact residuals rho_i = |u_i|^2 - 2 u_i.w~ satisfy 2 U (c* - c~) = rho, so
// |c* - c~| <= ||adj U||_F |rho| / (2 |det U|) <= 6 L_u^2 |rho|_1 / (2 O_lb),
// with |rho_i| bounded by its FP64 value plus 16 eps (3 L_u^2 + 6 L_u W).
// A tet whose orientation the static filter cannot certify positive
// (orient3d's bound, (7 + 56 eps) eps x its permanent, doubled), or whose
// delta exceeds 1e-6 of its radius, gets w = 0 and is never certified; so
// does any test with s outside [1e-50, 1e50] (no overflow or underflow of
// s^5 and m^2 w).  A certified test gives the sign of the exact determinant,
// which is also the decision of the robust predicates.
#pragma once
#include <hip/hip_runtime.h>

#include "geom.h"

namespace nat {

struct alignas(64) SphereTet {
    double cx, cy, cz;  // FP64 circumcentre c~
    double r2;          // |a - c~|^2, a the first vertex of the oriented tet
    double d1;          // 160 delta^2, rounded up; low 4 bits: sphere_dm
    double w;           // 1e21 O_lb^2; 0: never certified
    int4 n;             // neighbour across face 0..3 (tet_n)
};
static_assert(sizeof(SphereTet) == 64, "two tets per 128-byte line");

// (a, b, c, d): the tet's vertices in the order the search passes them to
// insphere (p0 and p1 swapped when the orientation bit is set)
__host__ __device__ inline SphereTet sphere_of(const d3 &a, const d3 &b, const d3 &c,
                                               const d3 &d, const int4 &n) {
    constexpr double eps = 1.1102230246251565e-16;  // 2^-53
    SphereTet s;
    s.n = n;
    s.d1 = 0.0;
    s.w = 0.0;
    const d3 cc = cc_tet(a, b, c, d);
    s.cx = cc.x;
    s.cy = cc.y;
    s.cz = cc.z;
    const d3 wv = sub3(cc, a);
    s.r2 = dot3(wv, wv);
    // the orientation, certified positive
    const double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    const double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    const double o = orient3d(a, b, c, d);
    const double perm = fabs(adx) * (fabs(bdy * cdz) + fabs(bdz * cdy)) +
                        fabs(bdx) * (fabs(cdy * adz) + fabs(cdz * ady)) +
                        fabs(cdx) * (fabs(ady * bdz) + fabs(adz * bdy));
    const double olb = o - 2.0 * (7.0 + 56.0 * eps) * eps * perm;
    if (!(olb > 0.0)) return s;
    // a posteriori bound on |c* - c~|
    const d3 u[3] = {sub3(b, a), sub3(c, a), sub3(d, a)};
    double lu = 0.0;
    for (int i = 0; i < 3; i++) lu = fmax(lu, fmax(fabs(u[i].x), fmax(fabs(u[i].y), fabs(u[i].z))));
    const double wm = fmax(fabs(wv.x), fmax(fabs(wv.y), fabs(wv.z)));
    const double err = 16.0 * eps * (3.0 * lu * lu + 6.0 * lu * wm);
    double rho = 0.0;
    for (int i = 0; i < 3; i++) rho += fabs(dot3(u[i], u[i]) - 2.0 * dot3(u[i], wv)) + err;
    const double delta = 1.01 * 6.0 * lu * lu * rho / (2.0 * olb);
    if (!(delta * delta <= 1e-12 * s.r2)) return s;
    const double w = 1e21 * olb * olb;
    if (!(w < 1e300)) return s;
    // rounded up to a multiple of 16 units in the last place: its low four
    // bits are free for the step-3 table's face mask (sphere_dm)
    const double d1 = 160.0 * delta * delta;
    unsigned long long bits;
    __builtin_memcpy(&bits, &d1, sizeof(bits));
    bits = (bits + 15ull) & ~15ull;
    __builtin_memcpy(&s.d1, &bits, sizeof(bits));
    s.w = w;
    return s;
}

// the face mask a tet's step-3 table was built with (DevMesh::old_vol), kept
// in the low bits of d1 (0 before the table is built); raising d1 by at most
// 15 units in the last place only makes the certificate stricter
__host__ __device__ __forceinline__ unsigned sphere_dm(const SphereTet &s) {
    unsigned long long bits;
    __builtin_memcpy(&bits, &s.d1, sizeof(bits));
    return static_cast<unsigned>(bits & 15ull);
}

// 1: e inside the tet's circumsphere, 0: outside (both certified to be the
// sign of the FP64 insphere determinant), -1: not certified
__host__ __device__ __forceinline__ int sphere_side(const SphereTet &s, const d3 &e) {
    const double dx = e.x - s.cx, dy = e.y - s.cy, dz = e.z - s.cz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double m = s.r2 - d2;
    const double sum = s.r2 + d2;
    const double m2 = m * m;
    const double s2 = sum * sum;
    // '&' rather than '&&': no short-circuit branches, so a caller's loads of
    // the record are all issued before the arithmetic
    const bool ok = (sum > 1e-50) & (sum < 1e50) & (m2 > 1e-28 * s2) & (m2 > s.d1 * sum) &
                    (m2 * s.w > s2 * s2 * sum);
    return ok ? (m > 0.0 ? 1 : 0) : -1;
}

}  // namespace nat
