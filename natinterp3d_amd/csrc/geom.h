// geom.h -- FP64 geometry shared by the host mesh builder and the gfx950
// kernels.  Every expression keeps the operand order of the reference so that
// results are bit-identical (the library is built with -ffp-contract=off and
// without fast-math, and FP64 division is IEEE correctly rounded on gfx950):
//   orient3d        plain 3x3 determinant       (predicates.c:1687-1710)
//   insphere        plain lifted determinant    (predicates.c:3301-3347)
//   cc_tri          triangle circumcentre       (natural_insertionfree.c:24-49)
//   cc_tet          tetrahedron circumcentre    (natural_insertionfree.c:54-84)
//   subcell_volume  Voronoi subcell volume      (natural_insertionfree.c:92-146)
#pragma once
#include <hip/hip_runtime.h>

namespace nat {

struct d3 {
    double x, y, z;
};

__host__ __device__ __forceinline__ d3 sub3(const d3 &a, const d3 &b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__host__ __device__ __forceinline__ d3 cross3(const d3 &b, const d3 &c) {
    return {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
}
__host__ __device__ __forceinline__ double dot3(const d3 &a, const d3 &b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Sign of the orientation of (a,b,c) seen from d; > 0 for the builder's
// positive orientation.
__host__ __device__ __forceinline__ double orient3d(const d3 &a, const d3 &b, const d3 &c,
                                                    const d3 &d) {
    const double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    const double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) +
           cdx * (ady * bdz - adz * bdy);
}

// > 0 when e lies inside the circumsphere of the positively oriented (a,b,c,d).
__host__ __device__ __forceinline__ double insphere(const d3 &a, const d3 &b, const d3 &c,
                                                    const d3 &d, const d3 &e) {
    const double aex = a.x - e.x, bex = b.x - e.x, cex = c.x - e.x, dex = d.x - e.x;
    const double aey = a.y - e.y, bey = b.y - e.y, cey = c.y - e.y, dey = d.y - e.y;
    const double aez = a.z - e.z, bez = b.z - e.z, cez = c.z - e.z, dez = d.z - e.z;
    const double ab = aex * bey - bex * aey;
    const double bc = bex * cey - cex * bey;
    const double cd = cex * dey - dex * cey;
    const double da = dex * aey - aex * dey;
    const double ac = aex * cey - cex * aey;
    const double bd = bex * dey - dex * bey;
    const double abc = aez * bc - bez * ac + cez * ab;
    const double bcd = bez * cd - cez * bd + dez * bc;
    const double cda = cez * da + dez * ac + aez * cd;
    const double dab = dez * ab + aez * bd + bez * da;
    const double alift = aex * aex + aey * aey + aez * aez;
    const double blift = bex * bex + bey * bey + bez * bez;
    const double clift = cex * cex + cey * cey + cez * cez;
    const double dlift = dex * dex + dey * dey + dez * dez;
    return (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
}

// Static error bound of the insphere determinant (Shewchuk, "Adaptive
// precision floating-point arithmetic and fast robust geometric predicates",
// 1997): with eps = 2^-53, the computed sign is certain when
// |det| > (16 + 224 eps) eps * permanent, the permanent being the same
// expression with every product and difference replaced by the sum of the
// magnitudes of its terms.  Returns true when the sign is not certified.
__host__ __device__ __forceinline__ bool insphere_uncertain(const d3 &a, const d3 &b, const d3 &c,
                                                            const d3 &d, const d3 &e) {
    const double aex = a.x - e.x, bex = b.x - e.x, cex = c.x - e.x, dex = d.x - e.x;
    const double aey = a.y - e.y, bey = b.y - e.y, cey = c.y - e.y, dey = d.y - e.y;
    const double aez = a.z - e.z, bez = b.z - e.z, cez = c.z - e.z, dez = d.z - e.z;
    auto minor = [](double p, double q, double r, double s, double &perm) {
        perm = fabs(p * q) + fabs(r * s);
        return p * q - r * s;
    };
    double pab, pbc, pcd, pda, pac, pbd;
    const double ab = minor(aex, bey, bex, aey, pab);
    const double bc = minor(bex, cey, cex, bey, pbc);
    const double cd = minor(cex, dey, dex, cey, pcd);
    const double da = minor(dex, aey, aex, dey, pda);
    const double ac = minor(aex, cey, cex, aey, pac);
    const double bd = minor(bex, dey, dex, bey, pbd);
    const double abc = aez * bc - bez * ac + cez * ab;
    const double bcd = bez * cd - cez * bd + dez * bc;
    const double cda = cez * da + dez * ac + aez * cd;
    const double dab = dez * ab + aez * bd + bez * da;
    const double pabc = fabs(aez) * pbc + fabs(bez) * pac + fabs(cez) * pab;
    const double pbcd = fabs(bez) * pcd + fabs(cez) * pbd + fabs(dez) * pbc;
    const double pcda = fabs(cez) * pda + fabs(dez) * pac + fabs(aez) * pcd;
    const double pdab = fabs(dez) * pab + fabs(aez) * pbd + fabs(bez) * pda;
    const double alift = aex * aex + aey * aey + aez * aez;
    const double blift = bex * bex + bey * bey + bez * bez;
    const double clift = cex * cex + cey * cey + cez * cez;
    const double dlift = dex * dex + dey * dey + dez * dez;
    const double det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
    const double perm = (dlift * pabc + clift * pdab) + (blift * pcda + alift * pbcd);
    const double eps = 1.1102230246251565e-16;  // 2^-53
    return fabs(det) <= (16.0 + 224.0 * eps) * eps * perm;
}

// Circumcentre of triangle (a,b,c); the centroid when |ab x ac|^2 < 1e-30.
__host__ __device__ __forceinline__ d3 cc_tri(const d3 &a, const d3 &b, const d3 &c) {
    const d3 ab = sub3(b, a), ac = sub3(c, a);
    const d3 n = cross3(ab, ac);
    const double n2 = dot3(n, n);
    if (n2 < 1e-30) {
        return {(a.x + b.x + c.x) / 3.0, (a.y + b.y + c.y) / 3.0, (a.z + b.z + c.z) / 3.0};
    }
    const double ab2 = dot3(ab, ab), ac2 = dot3(ac, ac);
    const d3 t = {ab2 * ac.x - ac2 * ab.x, ab2 * ac.y - ac2 * ab.y, ab2 * ac.z - ac2 * ab.z};
    const d3 cr = cross3(t, n);
    const double inv = 1.0 / (2.0 * n2);
    return {a.x + cr.x * inv, a.y + cr.y * inv, a.z + cr.z * inv};
}

// Circumcentre of tetrahedron (a,b,c,d); the centroid when the
// denominator 2 (b-a).((c-a)x(d-a)) is below 1e-30 in magnitude.
__host__ __device__ __forceinline__ d3 cc_tet(const d3 &a, const d3 &b, const d3 &c,
                                              const d3 &d) {
    const d3 ba = sub3(b, a), ca = sub3(c, a), da = sub3(d, a);
    const d3 x1 = cross3(ba, ca), x2 = cross3(da, ba), x3 = cross3(ca, da);
    const double l1 = dot3(da, da), l2 = dot3(ca, ca), l3 = dot3(ba, ba);
    const d3 s = {x3.x * l3 + (x1.x * l1 + x2.x * l2), x3.y * l3 + (x1.y * l1 + x2.y * l2),
                  x3.z * l3 + (x1.z * l1 + x2.z * l2)};
    const double denom = 2.0 * dot3(ba, x3);
    if (fabs(denom) < 1e-30) {
        return {(a.x + b.x + c.x + d.x) * 0.25, (a.y + b.y + c.y + d.y) * 0.25,
                (a.z + b.z + c.z + d.z) * 0.25};
    }
    const double inv = 1.0 / denom;
    return {s.x * inv + a.x, s.y * inv + a.y, s.z * inv + a.z};
}

// Signed volume of the Voronoi subcell of pk inside tet (pk,pa,pb,pc), given
// the tet circumcentre ct and the circumcentres of faces (pk,pa,pb),
// (pk,pa,pc), (pk,pb,pc).
__host__ __device__ __forceinline__ double subcell_volume(const d3 &pk, const d3 &pa,
                                                          const d3 &pb, const d3 &pc,
                                                          const d3 &ct, const d3 &c_fab,
                                                          const d3 &c_fac, const d3 &c_fbc) {
    const d3 ta = sub3(pa, pk), tb = sub3(pb, pk), tc = sub3(pc, pk);
    const double ox = tb.y * tc.z - tb.z * tc.y;
    const double oy = tb.z * tc.x - tb.x * tc.z;
    const double oz = tb.x * tc.y - tb.y * tc.x;
    const double sv = ta.x * ox + ta.y * oy + ta.z * oz;
    const bool sw = sv < 0;
    const d3 a = sw ? pb : pa, b = sw ? pa : pb;
    const d3 fac = sw ? c_fbc : c_fac, fbc = sw ? c_fac : c_fbc;
    const d3 df = sub3(c_fab, pk), dg = sub3(fac, pk), dh = sub3(fbc, pk), dt = sub3(ct, pk);
    const d3 C1 = cross3(df, dt), C2 = cross3(dt, dg), C3 = cross3(dh, dt);
    const d3 dab = sub3(a, b), dac = sub3(a, pc), dbc = sub3(b, pc);
    const double vol = dab.x * C1.x + dab.y * C1.y + dab.z * C1.z + dac.x * C2.x +
                       dac.y * C2.y + dac.z * C2.z + dbc.x * C3.x + dbc.y * C3.y +
                       dbc.z * C3.z;
    return vol / 12.0;
}

// Vertex slots of face i of a tet (the face shared with neighbour slot i):
// (p0,p1,p2), (p3,p1,p0), (p0,p2,p3), (p3,p2,p1)   (delaunay.c:444-467)
__host__ __device__ __forceinline__ int face_slot(int face, int k) {
    // packed 2-bit table, face-major
    constexpr unsigned kTab = (0u << 0) | (1u << 2) | (2u << 4) |     // face 0
                              (3u << 6) | (1u << 8) | (0u << 10) |    // face 1
                              (0u << 12) | (2u << 14) | (3u << 16) |  // face 2
                              (3u << 18) | (2u << 20) | (1u << 22);   // face 3
    return (kTab >> (2 * (face * 3 + k))) & 3u;
}

}  // namespace nat
