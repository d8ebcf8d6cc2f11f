// geom32.h -- FP32 geometry of the mixed-precision row path.  Points are
// taken relative to the query (differences formed in FP64, then rounded), so
// the FP32 values are of the size of the local point spacing.  Reciprocals use
// the hardware v_rcp_f32 (1 ulp).  Each tet or virtual tet first gets a
// condition estimate; the callers recompute the ill-conditioned ones with the
// FP64 functions of geom.h.
#pragma once
#include <hip/hip_runtime.h>

#include "geom.h"

namespace nat {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 rel32(const d3 &p, const d3 &q) {
    return {static_cast<float>(p.x - q.x), static_cast<float>(p.y - q.y),
            static_cast<float>(p.z - q.z)};
}
__device__ __forceinline__ f3 sub3f(const f3 &a, const f3 &b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ f3 cross3f(const f3 &b, const f3 &c) {
    return {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
}
__device__ __forceinline__ float dot3f(const f3 &a, const f3 &b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// component-wise select (a conditional on whole structs can go through memory)
__device__ __forceinline__ f3 self3(bool s, const f3 &a, const f3 &b) {
    return {s ? a.x : b.x, s ? a.y : b.y, s ? a.z : b.z};
}

// Condition of a tet: |(b-a).((c-a)x(d-a))| / (|b-a| |c-a| |d-a|), in [0, 1];
// about 0.7 for a regular tet, small for slivers and flat tets, whose
// circumcentres are far and badly determined.
__device__ __forceinline__ float tet_cond(const f3 &a, const f3 &b, const f3 &c, const f3 &d) {
    const f3 ba = sub3f(b, a), ca = sub3f(c, a), da = sub3f(d, a);
    const float det = dot3f(ba, cross3f(ca, da));
    const float l = dot3f(ba, ba) * dot3f(ca, ca) * dot3f(da, da);
    return fabsf(det) * rsqrtf(l);
}

// Condition of a triangle: |(b-a)x(c-a)|^2 / (|b-a|^2 |c-a|^2) (sin^2 of the
// angle at a).
__device__ __forceinline__ float tri_cond(const f3 &a, const f3 &b, const f3 &c) {
    const f3 ab = sub3f(b, a), ac = sub3f(c, a);
    const f3 n = cross3f(ab, ac);
    return dot3f(n, n) * rcpf(dot3f(ab, ab) * dot3f(ac, ac));
}

// The circumcentre formulas of geom.h in FP32.
__device__ __forceinline__ f3 cc_tri32(const f3 &a, const f3 &b, const f3 &c) {
    const f3 ab = sub3f(b, a), ac = sub3f(c, a);
    const f3 n = cross3f(ab, ac);
    const float n2 = dot3f(n, n);
    const float ab2 = dot3f(ab, ab), ac2 = dot3f(ac, ac);
    const f3 t = {ab2 * ac.x - ac2 * ab.x, ab2 * ac.y - ac2 * ab.y, ab2 * ac.z - ac2 * ab.z};
    const f3 cr = cross3f(t, n);
    const float inv = rcpf(2.0f * n2);
    return {a.x + cr.x * inv, a.y + cr.y * inv, a.z + cr.z * inv};
}

__device__ __forceinline__ f3 cc_tet32(const f3 &a, const f3 &b, const f3 &c, const f3 &d) {
    const f3 ba = sub3f(b, a), ca = sub3f(c, a), da = sub3f(d, a);
    const f3 x1 = cross3f(ba, ca), x2 = cross3f(da, ba), x3 = cross3f(ca, da);
    const float l1 = dot3f(da, da), l2 = dot3f(ca, ca), l3 = dot3f(ba, ba);
    const f3 s = {x3.x * l3 + (x1.x * l1 + x2.x * l2), x3.y * l3 + (x1.y * l1 + x2.y * l2),
                  x3.z * l3 + (x1.z * l1 + x2.z * l2)};
    const float inv = rcpf(2.0f * dot3f(ba, x3));
    return {s.x * inv + a.x, s.y * inv + a.y, s.z * inv + a.z};
}

// subcell_volume of geom.h in FP32 (same orientation rule and formula).
__device__ __forceinline__ float subcell_volume32(const f3 &pk, const f3 &pa, const f3 &pb,
                                                  const f3 &pc, const f3 &ct, const f3 &c_fab,
                                                  const f3 &c_fac, const f3 &c_fbc) {
    const f3 ta = sub3f(pa, pk), tb = sub3f(pb, pk), tc = sub3f(pc, pk);
    const bool sw = dot3f(ta, cross3f(tb, tc)) < 0;
    const f3 a = self3(sw, pb, pa), b = self3(sw, pa, pb);
    const f3 fac = self3(sw, c_fbc, c_fac), fbc = self3(sw, c_fac, c_fbc);
    const f3 df = sub3f(c_fab, pk), dg = sub3f(fac, pk), dh = sub3f(fbc, pk),
             dt = sub3f(ct, pk);
    const f3 C1 = cross3f(df, dt), C2 = cross3f(dt, dg), C3 = cross3f(dh, dt);
    const f3 dab = sub3f(a, b), dac = sub3f(a, pc), dbc = sub3f(b, pc);
    return (dot3f(dab, C1) + dot3f(dac, C2) + dot3f(dbc, C3)) * (1.0f / 12.0f);
}

}  // namespace nat
