// sibson_kernels.hip -- gfx950 kernels of the Sibson row computation.
//
// Level 0 (every query) runs as two kernels:
//   locate_kernel      one thread per query: start tet (exact nearest
//                      centroid, then the visibility walk)  delaunay.c:383-407
//   topo_group_kernel  16 lanes per query, state in LDS: cavity in the
//                      reference's LIFO order, boundary faces, neighbour
//                      first-appearance order; a compact record per query
//                                             natural_insertionfree.c:486-584
//   geom_rows_kernel  two lanes per query: Voronoi subcell volumes from the
//                 record and the step-3 table (steps 3-5), row into the pool;
//                 geom_kernel, one thread per query, where the table is
//                 absent or the row policy asks for one lane
//                                             natural_insertionfree.c:628-737
// Queries whose cavity or row exceeds the level-0 capacities are finished by
// sibson_rows_kernel, which does all of it in one thread with larger scratch.
//
// Every FP64 decision and value is bit-identical to the reference: the same
// expressions (geom.h), the same visiting and accumulation orders, built
// with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "exact.h"
#include "geom.h"
#include "geom32.h"
#include "nat_internal.h"
#include "sphere.h"
#include "step3_table.h"

namespace nat {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ d3 ld_vert(const DevMesh &m, int slot) {
    const double4 v = m.verts[slot];
    return {v.x, v.y, v.z};
}
__device__ __forceinline__ int sel4(const int4 &v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ int4 ld_tet_v(const DevMesh &m, int t, bool &neg) {
    int4 v = m.tet_v[t];
    neg = v.w < 0;
    v.w &= 0x7fffffff;
    return v;
}
// the four vertices of a fat tet
__device__ __forceinline__ void fat_points(const FatTet &f, d3 &p0, d3 &p1, d3 &p2, d3 &p3) {
    const double4 a = f.c0, b = f.c1, c = f.c2;
    p0 = {a.x, a.y, a.z};
    p1 = {a.w, b.x, b.y};
    p2 = {b.z, b.w, c.x};
    p3 = {c.y, c.z, c.w};
}

__device__ __forceinline__ double seld(const d3 &v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}
__device__ __forceinline__ unsigned hash_key(int k, int log2) {
    return (static_cast<unsigned>(k) * 2654435761u) >> (32 - log2);
}
__device__ __forceinline__ size_t il(int k, size_t S, int tid) {  // interleaved index
    return static_cast<size_t>(k) * S + tid;
}

// ------------------------------------------------------------- k-d nearest

// kd_nearest3_data / kd_nearest_i (kdtree.c:377-434, 543-564) with an
// explicit stack: frame d holds the node at depth d (split axis d mod 3),
// the phase reached (0 = enter, 1 = nearer subtree done, 2 = farther subtree
// done) and the rectangle bound it saved before slicing.
__device__ int kd_nearest(const DevMesh &m, const LocParams &L, const d3 &q, const Scratch &sc,
                          int tid, int cap, bool &ovf) {
    const int root = L.kd_root;
    if (root < 0) return -1;
    double rmin[3] = {L.rmin[0], L.rmin[1], L.rmin[2]};
    double rmax[3] = {L.rmax[0], L.rmax[1], L.rmax[2]};
    const size_t S = static_cast<size_t>(sc.nthr);
    const double4 r0 = m.kd[root];
    double best_d;
    {
        const double dx = r0.x - q.x, dy = r0.y - q.y, dz = r0.z - q.z;
        best_d = dx * dx + dy * dy + dz * dz;
    }
    int best = root;
    int sp = 1;
    sc.kd_node[tid] = root;
    while (sp > 0) {
        const size_t fi = il(sp - 1, S, tid);
        const int fr = sc.kd_node[fi];
        const int node = fr & 0x3fffffff;
        const int phase = fr >> 30;
        const int dir = (sp - 1) % 3;
        const double4 nd = m.kd[node];
        const int2 lr = *reinterpret_cast<const int2 *>(&nd.w);
        const d3 np = {nd.x, nd.y, nd.z};
        const double npd = seld(np, dir);
        const double qd = seld(q, dir);
        const bool go_left = (qd - npd) <= 0;
        const int nearer = go_left ? lr.x : lr.y;
        const int farther = go_left ? lr.y : lr.x;
        if (phase == 0) {
            if (nearer >= 0) {
                if (sp >= cap) {
                    ovf = true;
                    return -1;
                }
                double &c = go_left ? rmax[dir] : rmin[dir];
                sc.kd_saved[fi] = c;
                c = npd;
                sc.kd_node[fi] = node | (1 << 30);
                sc.kd_node[il(sp, S, tid)] = nearer;
                sp++;
                continue;
            }
        } else if (phase == 1) {
            double &c = go_left ? rmax[dir] : rmin[dir];
            c = sc.kd_saved[fi];
        } else {
            double &c = go_left ? rmin[dir] : rmax[dir];
            c = sc.kd_saved[fi];
            sp--;
            continue;
        }
        {
            double d = 0;
            const double dx = np.x - q.x, dy = np.y - q.y, dz = np.z - q.z;
            d += dx * dx;
            d += dy * dy;
            d += dz * dz;
            if (d < best_d) {
                best = node;
                best_d = d;
            }
        }
        if (farther >= 0) {
            double &c = go_left ? rmin[dir] : rmax[dir];
            const double saved = c;
            c = npd;
            double rd = 0;
            if (q.x < rmin[0]) { const double t = rmin[0] - q.x; rd += t * t; }
            else if (q.x > rmax[0]) { const double t = rmax[0] - q.x; rd += t * t; }
            if (q.y < rmin[1]) { const double t = rmin[1] - q.y; rd += t * t; }
            else if (q.y > rmax[1]) { const double t = rmax[1] - q.y; rd += t * t; }
            if (q.z < rmin[2]) { const double t = rmin[2] - q.z; rd += t * t; }
            else if (q.z > rmax[2]) { const double t = rmax[2] - q.z; rd += t * t; }
            if (rd < best_d) {
                if (sp >= cap) {
                    ovf = true;
                    return -1;
                }
                sc.kd_saved[fi] = saved;
                sc.kd_node[fi] = node | (2 << 30);
                sc.kd_node[il(sp, S, tid)] = farther;
                sp++;
                continue;
            }
            c = saved;
        }
        sp--;
    }
    return best;
}

// ------------------------------------------------- grid nearest centroid

// The k-d search above returns the centroid of least squared distance
// (dx^2 + dy^2) + dz^2 -- it never prunes a subtree holding a strictly
// closer centroid -- and among equal distances the first one it meets.  This
// search finds the same least distance over a uniform grid of the centroids
// and returns its tet when the answer is certain: no unexamined centroid can
// be as close (distance certificate) and no examined one ties.  Otherwise it
// returns -1 and the caller runs the k-d replica.
__device__ int grid_nearest(const DevMesh &m, const LocParams &L, const d3 &q) {
    const double fx = (q.x - L.g0[0]) * L.ginv;
    const double fy = (q.y - L.g0[1]) * L.ginv;
    const double fz = (q.z - L.g0[2]) * L.ginv;
    if (!(fx >= 0.0 && fx < L.gn[0] && fy >= 0.0 && fy < L.gn[1] && fz >= 0.0 &&
          fz < L.gn[2]))
        return -1;
    const int ix = static_cast<int>(fx), iy = static_cast<int>(fy), iz = static_cast<int>(fz);
    const int ii[3] = {ix, iy, iz};
    double best_d = __builtin_huge_val();
    int best = -1;
    bool tie = false;
    const double h = L.gh;
    constexpr double kSafe = 1.0 - 1e-9;
    for (int r = 0; r <= 2; r++) {
        for (int dz = -r; dz <= r; dz++) {
            const int z = iz + dz;
            if (z < 0 || z >= L.gn[2]) continue;
            for (int dy = -r; dy <= r; dy++) {
                const int y = iy + dy;
                if (y < 0 || y >= L.gn[1]) continue;
                const bool shell = (dz == -r || dz == r || dy == -r || dy == r);
                const int step = shell ? 1 : (r > 0 ? 2 * r : 1);
                for (int dx = -r; dx <= r; dx += step) {
                    const int x = ix + dx;
                    if (x < 0 || x >= L.gn[0]) continue;
                    // squared distance from q to the cell box, rounded down
                    const double lx = L.g0[0] + x * h, ly = L.g0[1] + y * h,
                                 lz = L.g0[2] + z * h;
                    const double ex = fmax(fmax(lx - q.x, q.x - (lx + h)), 0.0);
                    const double ey = fmax(fmax(ly - q.y, q.y - (ly + h)), 0.0);
                    const double ez = fmax(fmax(lz - q.z, q.z - (lz + h)), 0.0);
                    const double cd = (ex * ex + ey * ey + ez * ez) * kSafe;
                    if (cd > best_d) continue;
                    const int c = L.cell_base + (z * L.gn[1] + y) * L.gn[0] + x;
                    const int b = m.cell_start[c], e = m.cell_start[c + 1];
                    for (int j = b; j < e; j++) {
                        const double4 p = m.cell_pts[j];
                        const double ddx = p.x - q.x, ddy = p.y - q.y, ddz = p.z - q.z;
                        const double d = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (d < best_d) {
                            best_d = d;
                            best = static_cast<int>(__double_as_longlong(p.w));
                            tie = false;
                        } else if (d == best_d) {
                            tie = true;
                        }
                    }
                }
            }
        }
        // distance from q to the outside of the searched block; a side that
        // reaches the grid's edge holds every centroid beyond it (they are
        // filed into the edge cells) and does not limit the certificate
        double rho = __builtin_huge_val();
        const double qq[3] = {q.x, q.y, q.z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (ii[a] - r > 0) rho = fmin(rho, qq[a] - (L.g0[a] + (ii[a] - r) * h));
            if (ii[a] + r < L.gn[a] - 1)
                rho = fmin(rho, (L.g0[a] + (ii[a] + r + 1) * h) - qq[a]);
        }
        if (best >= 0 && best_d < rho * rho * kSafe) return tie ? -1 : best;
    }
    return -1;
}

// The two decisions of the query path.  ROBUST = false: the reference's plain
// FP64 signs.  ROBUST = true (opt-in): the FP64 sign where the static error
// bound certifies it, else the exact sign (exact.h); m.robust_ctr counts the
// exact evaluations [0], the decisions that differ from the FP64 sign [1] and
// the tests too wide for the exact evaluator, which keep the FP64 sign [2].
__device__ __noinline__ bool robust_inside(const DevMesh &m, const d3 &a, const d3 &b,
                                           const d3 &c, const d3 &d, const d3 &e, double fast) {
    const int sg = insphere_exact_sign(a, b, c, d, e);
    unsigned long long *ctr = m.robust_ctr;
    if (ctr) atomicAdd(&ctr[0], 1ull);
    if (sg == 2) {
        if (ctr) atomicAdd(&ctr[2], 1ull);
        return fast > 0;
    }
    if (ctr && (sg > 0) != (fast > 0)) atomicAdd(&ctr[1], 1ull);
    return sg > 0;
}
__device__ __noinline__ bool robust_separated(const DevMesh &m, const d3 &a, const d3 &b,
                                              const d3 &c, const d3 &q, double fast) {
    const int sg = orient3d_exact_sign(a, b, c, q);
    unsigned long long *ctr = m.robust_ctr;
    if (ctr) atomicAdd(&ctr[0], 1ull);
    if (sg == 2) {
        if (ctr) atomicAdd(&ctr[2], 1ull);
        return fast < 0;
    }
    if (ctr && (sg < 0) != (fast < 0)) atomicAdd(&ctr[1], 1ull);
    return sg < 0;
}
// e inside the circumsphere of the positively oriented (a, b, c, d)
template <bool ROBUST>
__device__ __forceinline__ bool inside_sphere(const DevMesh &m, const d3 &a, const d3 &b,
                                              const d3 &c, const d3 &d, const d3 &e) {
    const double ins = insphere(a, b, c, d, e);
    if constexpr (ROBUST) {
        if (insphere_uncertain(a, b, c, d, e)) return robust_inside(m, a, b, c, d, e, ins);
    }
    return ins > 0;
}
// The search's decision for tet t (record sp = m.sph[t]): the certified
// circumsphere test, the tet's determinant where that cannot decide.
template <bool ROBUST>
__device__ __forceinline__ bool tet_inside(const DevMesh &m, const SphereTet &sp, int t,
                                           const d3 &q) {
    const int sd = sphere_side(sp, q);
    if (sd >= 0) return sd > 0;
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const d3 p0 = ld_vert(m, tv.x), p1 = ld_vert(m, tv.y);
    const d3 p2 = ld_vert(m, tv.z), p3 = ld_vert(m, tv.w);
    return neg ? inside_sphere<ROBUST>(m, p1, p0, p2, p3, q)
               : inside_sphere<ROBUST>(m, p0, p1, p2, p3, q);
}

// q on the outer side of face (a, b, c)
template <bool ROBUST>
__device__ __forceinline__ bool separated(const DevMesh &m, const d3 &a, const d3 &b,
                                          const d3 &c, const d3 &q) {
    const double o = orient3d(a, b, c, q);
    if constexpr (ROBUST) {
        if (orient3d_uncertain(a, b, c, q)) return robust_separated(m, a, b, c, q, o);
    }
    return o < 0;
}

// findContainingSimplex (delaunay.c:383-407): leave through the first face
// whose plane separates the query, restart the face loop after each move.
template <bool ROBUST>
__device__ int walk(const DevMesh &m, int s, const d3 &q, bool &ovf) {
    bool neg;
    int4 tv = ld_tet_v(m, s, neg);
    int steps = 0;
    int i = 0;
    while (i < 4) {
        const d3 a = ld_vert(m, sel4(tv, face_slot(i, 0)));
        const d3 b = ld_vert(m, sel4(tv, face_slot(i, 1)));
        const d3 c = ld_vert(m, sel4(tv, face_slot(i, 2)));
        if (separated<ROBUST>(m, a, b, c, q)) {
            const int nb = sel4(m.tet_n[s], i);
            if (nb < 0) return -1;
            s = nb;
            tv = ld_tet_v(m, s, neg);
            i = 0;
            if (++steps > (1 << 24)) {
                ovf = true;
                return -1;
            }
            continue;
        }
        i++;
    }
    return s;
}

// Start tet of the cavity search (findContainingSimplex with the k-d seed).
template <bool ROBUST>
__device__ int locate(const DevMesh &m, const LocParams &L, const d3 &q, const Scratch &sc,
                      int tid, const Caps &caps, bool &ovf) {
    int s = grid_nearest(m, L, q);
    if (s < 0) s = kd_nearest(m, L, q, sc, tid, caps.kd, ovf);
    if (ovf || s < 0) return -1;
    return walk<ROBUST>(m, s, q, ovf);
}

// ----------------------------------------------------------- hash tables

// Open-addressing table of one thread, entries (key, stamp) plus an int
// value; entries of other stamps read as empty, so a new query only needs a
// new stamp.  Set semantics of natural_insertionfree.c:152-211 and of the
// neighbour map natural_insertionfree.c:373-382.
struct HashTab {
    int2 *tab;
    int *val;
    size_t base;
    unsigned mask;
    int log2, stamp;
    __device__ HashTab(int2 *t, int *v, int tid, int lg, int st)
        : tab(t), val(v), base(static_cast<size_t>(tid) << lg), mask((1u << lg) - 1),
          log2(lg), stamp(st) {}
    // slot of key, inserting it (value -1) when absent; found = was present
    __device__ unsigned insert(int key, bool &found) const {
        unsigned h = hash_key(key, log2);
        for (;;) {
            const int2 e = tab[base + h];
            if (e.y != stamp) {
                tab[base + h] = make_int2(key, stamp);
                val[base + h] = -1;
                found = false;
                return h;
            }
            if (e.x == key) {
                found = true;
                return h;
            }
            h = (h + 1) & mask;
        }
    }
    __device__ bool contains(int key) const {
        unsigned h = hash_key(key, log2);
        for (;;) {
            const int2 e = tab[base + h];
            if (e.y != stamp) return false;
            if (e.x == key) return true;
            h = (h + 1) & mask;
        }
    }
    // value of key, -1 when absent
    __device__ int get(int key) const {
        unsigned h = hash_key(key, log2);
        for (;;) {
            const int2 e = tab[base + h];
            if (e.y != stamp) return -1;
            if (e.x == key) return val[base + h];
            h = (h + 1) & mask;
        }
    }
    __device__ void set(unsigned h, int v) const { val[base + h] = v; }
};

// findCavity (natural_insertionfree.c:498-530): LIFO search from `start`;
// a tet is marked visited when popped, pushed neighbours are those not yet
// visited at push time.  Returns the cavity size (tets in sc.cav in cavity
// order, cavity positions in the visited table) or -1 on overflow.
template <bool ROBUST>
__device__ int cavity_search(const DevMesh &m, const Caps &caps, const Scratch &sc, int tid,
                             const HashTab &vis, int start, const d3 &q) {
    const size_t S = static_cast<size_t>(sc.nthr);
    const int vis_max = (3 << caps.vis_log2) / 4;
    int nvis = 0, ncav = 0, top = 1;
    sc.stack[tid] = start;
    while (top > 0) {
        top--;
        const int cur = sc.stack[il(top, S, tid)];
        bool found;
        const unsigned h = vis.insert(cur, found);
        if (found) continue;
        if (++nvis >= vis_max) return -1;
        const SphereTet sp = m.sph[cur];
        if (tet_inside<ROBUST>(m, sp, cur, q)) {
            if (ncav >= caps.cav) return -1;
            sc.cav[il(ncav, S, tid)] = cur;
            vis.set(h, ncav);
            ncav++;
            const int4 tn = sp.n;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int nb = sel4(tn, i);
                if (nb >= 0 && !vis.contains(nb)) {
                    if (top >= caps.stack) return -1;
                    sc.stack[il(top, S, tid)] = nb;
                    top++;
                }
            }
        }
    }
    return ncav;
}

// Coincident-vertex path (natural_insertionfree.c:596-622): vertices of the
// start tet equal to the query (the reference's re-walk ends in the same
// tet).  Writes their indices to out[0..n) and returns n.
__device__ int coincident(const DevMesh &m, int start, const d3 &q, int out[4]) {
    bool neg;
    const int4 tv = ld_tet_v(m, start, neg);
    int n = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int p = sel4(tv, i);
        const d3 v = ld_vert(m, p);
        if (p < m.n_data && v.x == q.x && v.y == q.y && v.z == q.z) out[n++] = p;
    }
    return n;
}

// The same in an int4 (unused entries 0), built with fixed indices so that
// nothing goes to private memory.
__device__ __forceinline__ int coincident4(const DevMesh &m, int start, const d3 &q, int4 &out) {
    bool neg;
    const int4 tv = ld_tet_v(m, start, neg);
    int n = 0;
    int o0 = 0, o1 = 0, o2 = 0, o3 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int p = sel4(tv, i);
        const d3 v = ld_vert(m, p);
        if (p < m.n_data && v.x == q.x && v.y == q.y && v.z == q.z) {
            o0 = n == 0 ? p : o0;
            o1 = n == 1 ? p : o1;
            o2 = n == 2 ? p : o2;
            o3 = n == 3 ? p : o3;
            n++;
        }
    }
    out = make_int4(o0, o1, o2, o3);
    return n;
}

// Slot of vertex v in a tet's vertex record (v is one of the four).
__device__ __forceinline__ int slot_in(const int4 &tv, int v) {
    return tv.x == v ? 0 : (tv.y == v ? 1 : (tv.z == v ? 2 : 3));
}

// Component-wise: a conditional on whole structs would select addresses and
// send the points through private memory.
__device__ __forceinline__ void cswap(d3 &x, d3 &y, int &rx, int &ry) {
    const bool s = rx > ry;
    const double x0 = x.x, x1 = x.y, x2 = x.z;
    x.x = s ? y.x : x0;
    x.y = s ? y.y : x1;
    x.z = s ? y.z : x2;
    y.x = s ? x0 : y.x;
    y.y = s ? x1 : y.y;
    y.z = s ? x2 : y.z;
    const int r = rx;
    rx = s ? ry : rx;
    ry = s ? r : ry;
}

// Orders three points by ascending rank (ranks distinct), with selects only.
__device__ __forceinline__ void sort3_by_rank(d3 &x0, d3 &x1, d3 &x2, int r0, int r1, int r2) {
    cswap(x0, x1, r0, r1);
    cswap(x1, x2, r1, r2);
    cswap(x0, x1, r0, r1);
}

// Vertex slots of the face of a tet opposite its vertex K, ascending.
template <int K>
struct FaceOf {
    static constexpr int a = K == 0 ? 1 : 0, b = K <= 1 ? 2 : 1, c = K <= 2 ? 3 : 2;
};

// The four points of a tet as separate values (not an array: selects among
// them must stay selects rather than become indexed private memory).
struct Quad {
    d3 p0, p1, p2, p3;
    template <int I>
    __device__ __forceinline__ const d3 &get() const {
        if constexpr (I == 0) return p0;
        else if constexpr (I == 1) return p1;
        else if constexpr (I == 2) return p2;
        else return p3;
    }
};

// Circumcentre of face K of a cavity tet (natural_insertionfree.c:649-652):
// its three vertices in ascending slot order of this tet, or, when fcmask
// bit K says an earlier cavity tet holds the face (the value was computed
// first there), in that tet's ascending slot order.  Both orders are of the
// same three points, so only the order is looked up.
template <int K>
__device__ __forceinline__ d3 step3_face_cc(const DevMesh &m, const Quad &P, const int4 &tv,
                                            const int4 &tn, unsigned fcmask) {
    constexpr int a = FaceOf<K>::a, b = FaceOf<K>::b, c = FaceOf<K>::c;
    d3 x0 = P.get<a>(), x1 = P.get<b>(), x2 = P.get<c>();
    int r0 = 0, r1 = 1, r2 = 2;
    if ((fcmask >> K) & 1u) {
        bool nneg;
        const int4 nv = ld_tet_v(m, sel4(tn, 3 - K), nneg);
        r0 = slot_in(nv, sel4(tv, a));
        r1 = slot_in(nv, sel4(tv, b));
        r2 = slot_in(nv, sel4(tv, c));
    }
    sort3_by_rank(x0, x1, x2, r0, r1, r2);
    return cc_tri(x0, x1, x2);
}

// Old subcell of vertex VI: the other three vertices ascending, the face
// circumcentres opposite them in descending order (natural_insertionfree.c:
// 660-672).
template <int VI, typename SlotOf, typename Add>
__device__ __forceinline__ void step3_vertex(const Quad &P, const Quad &F, const d3 &ct,
                                             const int4 &tv, SlotOf slot_of, Add add) {
    constexpr int a = FaceOf<VI>::a, b = FaceOf<VI>::b, c = FaceOf<VI>::c;
    const int slot = slot_of(VI, sel4(tv, VI));
    if (slot < 0) return;
    add(slot, subcell_volume(P.get<VI>(), P.get<a>(), P.get<b>(), P.get<c>(), ct, F.get<c>(),
                             F.get<b>(), F.get<a>()));
}

// Step 3 for one cavity tet (natural_insertionfree.c:642-676): calls
// add(slot, vol) for each vertex with a neighbour slot, in vertex order.
// fcmask bit k: face opposite p[k] is held by an earlier cavity tet.
template <typename SlotOf, typename Add>
__device__ __forceinline__ void old_contributions(const DevMesh &m, int t, unsigned fcmask,
                                                  SlotOf slot_of, Add add) {
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const Quad P = {ld_vert(m, tv.x), ld_vert(m, tv.y), ld_vert(m, tv.z), ld_vert(m, tv.w)};
    const d3 ct = cc_tet(P.p0, P.p1, P.p2, P.p3);
    int4 tn = make_int4(-1, -1, -1, -1);
    if (fcmask) tn = m.tet_n[t];
    const Quad F = {step3_face_cc<0>(m, P, tv, tn, fcmask), step3_face_cc<1>(m, P, tv, tn, fcmask),
                    step3_face_cc<2>(m, P, tv, tn, fcmask), step3_face_cc<3>(m, P, tv, tn, fcmask)};
    step3_vertex<0>(P, F, ct, tv, slot_of, add);
    step3_vertex<1>(P, F, ct, tv, slot_of, add);
    step3_vertex<2>(P, F, ct, tv, slot_of, add);
    step3_vertex<3>(P, F, ct, tv, slot_of, add);
}

// The step-3 table (DevMesh::old_vol, addressed as step3_table.h says):
// per tet, the old subcell volume of each vertex for each reduced face mask.
// Same functions, same operands: each entry has the bits old_contributions
// would compute.
__device__ __forceinline__ d3 sel3d(bool s, const d3 &a, const d3 &b) {
    return {s ? a.x : b.x, s ? a.y : b.y, s ? a.z : b.z};
}

// vertex VI's value for the reduced mask m (its bits of faces a, b, c choose
// the circumcentres' vertex orders)
template <int VI>
__device__ __forceinline__ double old_vol_value(const Quad &P, const Quad &Fo, const Quad &Fn,
                                                const d3 &ct, unsigned m) {
    constexpr int a = FaceOf<VI>::a, b = FaceOf<VI>::b, c = FaceOf<VI>::c;
    const d3 fa = sel3d((m >> a) & 1u, Fn.get<a>(), Fo.get<a>());
    const d3 fb = sel3d((m >> b) & 1u, Fn.get<b>(), Fo.get<b>());
    const d3 fc = sel3d((m >> c) & 1u, Fn.get<c>(), Fo.get<c>());
    return subcell_volume(P.get<VI>(), P.get<a>(), P.get<b>(), P.get<c>(), ct, fc, fb, fa);
}

__device__ __forceinline__ bool same_bits(const d3 &a, const d3 &b) {
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
           __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
           __double_as_longlong(a.z) == __double_as_longlong(b.z);
}

// The cavity search's circumsphere of each tet (sphere.h), from its fat
// record, with the vertex order the search passes to insphere.
__global__ void __launch_bounds__(256) sphere_kernel(DevMesh m, SphereTet *__restrict__ out) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= m.n_tets) return;
    const FatTet &ft = m.fat[t];
    d3 p0, p1, p2, p3;
    fat_points(ft, p0, p1, p2, p3);
    out[t] = ft.v.w < 0 ? sphere_of(p1, p0, p2, p3, ft.n) : sphere_of(p0, p1, p2, p3, ft.n);
}

__global__ void __launch_bounds__(256) old_vol_kernel(DevMesh m, double *__restrict__ out,
                                                      SphereTet *__restrict__ sph) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.n_tets) return;
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const Quad P = {ld_vert(m, tv.x), ld_vert(m, tv.y), ld_vert(m, tv.z), ld_vert(m, tv.w)};
    const d3 ct = cc_tet(P.p0, P.p1, P.p2, P.p3);
    const int4 tn = m.tet_n[t];
    const Quad Fo = {step3_face_cc<0>(m, P, tv, tn, 0u), step3_face_cc<1>(m, P, tv, tn, 0u),
                     step3_face_cc<2>(m, P, tv, tn, 0u), step3_face_cc<3>(m, P, tv, tn, 0u)};
    // face K's neighbour is tn[3 - K]; without one the variant is never used
    const Quad Fn = {step3_face_cc<0>(m, P, tv, tn, tn.w >= 0 ? 1u : 0u),
                     step3_face_cc<1>(m, P, tv, tn, tn.z >= 0 ? 2u : 0u),
                     step3_face_cc<2>(m, P, tv, tn, tn.y >= 0 ? 4u : 0u),
                     step3_face_cc<3>(m, P, tv, tn, tn.x >= 0 ? 8u : 0u)};
    const unsigned dm = (same_bits(Fo.p0, Fn.p0) ? 0u : 1u) | (same_bits(Fo.p1, Fn.p1) ? 0u : 2u) |
                        (same_bits(Fo.p2, Fn.p2) ? 0u : 4u) | (same_bits(Fo.p3, Fn.p3) ? 0u : 8u);
    {  // the mask into the low bits of the circumsphere record's d1 (sphere_dm)
        unsigned long long b = __double_as_longlong(sph[t].d1);
        b = (b & ~15ull) | dm;
        sph[t].d1 = __longlong_as_double(static_cast<long long>(b));
    }
    double4 *o = reinterpret_cast<double4 *>(out + static_cast<size_t>(t) * 64);
    for (unsigned mk = 0; mk < 16; mk++) {
        if (mk & ~dm) continue;  // not a reduced mask of this tet
        o[table_pos(mk, dm)] = make_double4(old_vol_value<0>(P, Fo, Fn, ct, mk),
                                            old_vol_value<1>(P, Fo, Fn, ct, mk),
                                            old_vol_value<2>(P, Fo, Fn, ct, mk),
                                            old_vol_value<3>(P, Fo, Fn, ct, mk));
    }
    for (unsigned p = 1u << __builtin_popcount(dm); p < 16; p++)
        o[p] = make_double4(0.0, 0.0, 0.0, 0.0);  // unused positions
}

// Step 3 for one cavity tet from the table; the same calls of add as
// old_contributions (vertex order, slots from slot_of(VI, .)).
// pos: table_pos(fcmask, sphere_dm(m.sph[t])).
template <typename SlotOf, typename Add>
__device__ __forceinline__ void old_contributions_table(const DevMesh &m, int t, unsigned pos,
                                                        SlotOf slot_of, Add add) {
#pragma unroll
    for (int vi = 0; vi < 4; vi++) {
        const int s = slot_of(vi, -1);
        if (s >= 0) add(s, m.old_vol[old_vol_index(t, pos, vi)]);
    }
}

// Derivative of a stolen volume: d sigma_k / dq = (1 / |pk - q|) times the
// moment, integral of (x - q) dA, of the new Voronoi facet between q and pk.
// Inside the virtual tet (pk, q, A, B) that facet is the quadrilateral
// (m, ca, cv, cb): m the midpoint of q and pk, ca and cb the circumcentres of
// faces (pk, q, A) and (pk, q, B), cv the tet's circumcentre.  Returns its
// moment divided by |pk - q|^2, triangle areas signed along pk - q (the
// boundary faces' consistent orientation makes the pieces add up).
// Everything relative to q: with h = m - q = w / 2 and A, V, B the
// circumcentres minus q, twice the triangle areas are a1 = ((A-h) x (V-h)).w
// and a2 = ((V-h) x (B-h)).w, their centroids (h+A+V)/3 and (h+V+B)/3, and
// the result (a1 (h+A+V) + a2 (h+V+B)) / (6 |w|^2): one division.
// facet_moment6 leaves the division to the caller (1 / (6 |w|^2) depends on
// the neighbour alone: facet_scale).
__device__ __forceinline__ d3 facet_moment6(const d3 &pk, const d3 &q, const d3 &ca,
                                            const d3 &cb, const d3 &cv) {
    const d3 w = sub3(pk, q);
    const d3 h = {w.x * 0.5, w.y * 0.5, w.z * 0.5};
    const d3 A = sub3(ca, q), V = sub3(cv, q), B = sub3(cb, q);
    const d3 u = sub3(A, h), v = sub3(V, h), x = sub3(B, h);
    const double a1 = dot3(cross3(u, v), w);
    const double a2 = dot3(cross3(v, x), w);
    return {a1 * (h.x + A.x + V.x) + a2 * (h.x + V.x + B.x),
            a1 * (h.y + A.y + V.y) + a2 * (h.y + V.y + B.y),
            a1 * (h.z + A.z + V.z) + a2 * (h.z + V.z + B.z)};
}
__device__ __forceinline__ double facet_scale(const d3 &pk, const d3 &q) {
    const d3 w = sub3(pk, q);
    return 1.0 / (6.0 * dot3(w, w));
}
__device__ __forceinline__ d3 facet_moment(const d3 &pk, const d3 &q, const d3 &ca,
                                           const d3 &cb, const d3 &cv) {
    const d3 m6 = facet_moment6(pk, q, ca, cb, cv);
    const double s = facet_scale(pk, q);
    return {m6.x * s, m6.y * s, m6.z * s};
}

struct NoMoment {
    __device__ void operator()(int, const d3 &) const {}
    static constexpr bool active = false;
    static constexpr bool raw = false;
};

// Step 4 for one boundary face, face fi of cavity tet t
// (natural_insertionfree.c:681-719): calls sub(slot, vol) per face vertex,
// and mom(slot, facet moment) when the gradient is wanted.  WithVol = false
// computes the moments only.
template <bool WithVol = true, typename SlotOf, typename Sub, typename Mom = NoMoment>
__device__ __forceinline__ void new_contributions(const DevMesh &m, int t, int fi, const d3 &q,
                                                  SlotOf slot_of, Sub sub, Mom mom = Mom()) {
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const int s[3] = {face_slot(fi, 0), face_slot(fi, 1), face_slot(fi, 2)};
    const d3 F[3] = {ld_vert(m, sel4(tv, s[0])), ld_vert(m, sel4(tv, s[1])),
                     ld_vert(m, sel4(tv, s[2]))};
    const d3 cv = cc_tet(q, F[0], F[1], F[2]);
    // the face's circumcentre in ascending slot order (the order of
    // faceCCLookup's first call, natural_insertionfree.c:649-652): faces 0
    // and 2 list their slots ascending, faces 1 and 3 descending
    const bool desc = fi & 1;
    d3 cf = {0.0, 0.0, 0.0};
    if constexpr (WithVol) cf = cc_tri(desc ? F[2] : F[0], F[1], desc ? F[0] : F[2]);
    if constexpr (!WithVol) {
        // moments only (no bit-exact order to keep): the triangle (F[k], q,
        // F[k+1]) is c_a of vertex k and c_b of vertex k+1, computed once
        const d3 t0 = cc_tri(F[0], q, F[1]), t1 = cc_tri(F[1], q, F[2]);
        const d3 t2 = cc_tri(F[2], q, F[0]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int slot = slot_of(s[k], sel4(tv, s[k]));
            if (slot < 0) continue;
            const d3 &ca = k == 0 ? t0 : (k == 1 ? t1 : t2);
            const d3 &cb = k == 0 ? t2 : (k == 1 ? t0 : t1);
            if constexpr (Mom::active) {
                if constexpr (Mom::raw)
                    mom(slot, facet_moment6(F[k], q, ca, cb, cv));
                else
                    mom(slot, facet_moment(F[k], q, ca, cb, cv));
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int slot = slot_of(s[k], sel4(tv, s[k]));
        if (slot < 0) continue;
        const d3 &A = F[(k + 1) % 3];
        const d3 &B = F[(k + 2) % 3];
        const d3 ca = cc_tri(F[k], q, A);
        const d3 cb = cc_tri(F[k], q, B);
        if constexpr (WithVol) sub(slot, subcell_volume(F[k], q, A, B, cv, ca, cb, cf));
        if constexpr (Mom::active) mom(slot, facet_moment(F[k], q, ca, cb, cv));
    }
}

// 1 / x for the gradient (no prescribed rounding): the hardware estimate and
// one Newton step, r (2 - x r), which squares its relative error (about
// 2^-52 after it) -- 3 instructions where the correctly rounded division
// takes about 10.  x is a denominator already checked against 1e-30.
__device__ __forceinline__ double grad_rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// The facet moments of boundary face fi of cavity tet t -- the gradient's
// part of step 4 -- in the query's frame.  With P_k = F_k - q, l_k = |P_k|^2
// and X_k = P_k x P_k+1, the virtual tet's circumcentre and the three edge
// triangles' share the cross products and squares:
//   cv - q = (l0 X1 + l1 X2 + l2 X0) / (2 P0.X1),
//   circumcentre of (F_k, q, F_k+1) - q = ((l_k P_k+1 - l_k+1 P_k) x X_k) / (2 |X_k|^2)
// (the centroid below the same 1e-30 denominators as cc_tet and cc_tri).  The
// gradient has no prescribed rounding, so nothing here has to match the
// forward's circumcentres bit for bit, and the divisions are grad_rcp.  f(slot, m6) receives the unscaled
// moment of each face vertex that has a neighbour slot (facet_moment6 with
// every point taken relative to q).
template <typename SlotOf, typename F>
__device__ __forceinline__ void face_moments_rel(const DevMesh &m, int t, int fi, const d3 &q,
                                                 SlotOf slot_of, F f) {
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const int s[3] = {face_slot(fi, 0), face_slot(fi, 1), face_slot(fi, 2)};
    d3 P[3];
    double l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        P[k] = sub3(ld_vert(m, sel4(tv, s[k])), q);
        l[k] = dot3(P[k], P[k]);
    }
    const d3 X[3] = {cross3(P[0], P[1]), cross3(P[1], P[2]), cross3(P[2], P[0])};
    d3 V;
    const double den = 2.0 * dot3(P[0], X[1]);
    if (fabs(den) < 1e-30) {
        V = {(P[0].x + P[1].x + P[2].x) * 0.25, (P[0].y + P[1].y + P[2].y) * 0.25,
             (P[0].z + P[1].z + P[2].z) * 0.25};
    } else {
        const double inv = grad_rcp(den);
        V = {(l[0] * X[1].x + l[1] * X[2].x + l[2] * X[0].x) * inv,
             (l[0] * X[1].y + l[1] * X[2].y + l[2] * X[0].y) * inv,
             (l[0] * X[1].z + l[1] * X[2].z + l[2] * X[0].z) * inv};
    }
    d3 T[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const d3 &a = P[k], &b = P[(k + 1) % 3];
        const double x2 = dot3(X[k], X[k]);
        if (x2 < 1e-30) {
            T[k] = {(a.x + b.x) / 3.0, (a.y + b.y) / 3.0, (a.z + b.z) / 3.0};
        } else {
            const double la = l[k], lb = l[(k + 1) % 3];
            const d3 c = cross3({la * b.x - lb * a.x, la * b.y - lb * a.y, la * b.z - lb * a.z},
                                X[k]);
            const double inv = 0.5 * grad_rcp(x2);
            T[k] = {c.x * inv, c.y * inv, c.z * inv};
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int slot = slot_of(s[k], sel4(tv, s[k]));
        if (slot < 0) continue;
        const d3 &w = P[k], &A = T[k], &B = T[(k + 2) % 3];
        const d3 h = {w.x * 0.5, w.y * 0.5, w.z * 0.5};
        const d3 u = sub3(A, h), v = sub3(V, h), x = sub3(B, h);
        const double a1 = dot3(cross3(u, v), w);
        const double a2 = dot3(cross3(v, x), w);
        f(slot, d3{a1 * (h.x + A.x + V.x) + a2 * (h.x + V.x + B.x),
                   a1 * (h.y + A.y + V.y) + a2 * (h.y + V.y + B.y),
                   a1 * (h.z + A.z + V.z) + a2 * (h.z + V.z + B.z)});
    }
}

// ---- mixed precision: steps 3 and 4 in FP32 relative to the query, with
// FP64 (the functions above, bit-identical to the plain path) for tets whose
// condition estimate is below tau and for tets with a super vertex (points
// far from the query, whose FP32 offsets would lose the local detail).
// Return false when FP64 is needed.

template <typename SlotOf, typename Add>
__device__ __forceinline__ bool old_contributions32(const DevMesh &m, int t, const d3 &q,
                                                    float tau, SlotOf slot_of, Add add) {
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const int N = m.n_data;
    if (tv.x >= N || tv.y >= N || tv.z >= N || tv.w >= N) return false;
    const f3 r0 = rel32(ld_vert(m, tv.x), q), r1 = rel32(ld_vert(m, tv.y), q);
    const f3 r2 = rel32(ld_vert(m, tv.z), q), r3 = rel32(ld_vert(m, tv.w), q);
    if (!(tet_cond(r0, r1, r2, r3) >= tau)) return false;
    const f3 ct = cc_tet32(r0, r1, r2, r3);
    // face circumcentres opposite each vertex, vertices ascending
    const f3 c0 = cc_tri32(r1, r2, r3), c1 = cc_tri32(r0, r2, r3);
    const f3 c2 = cc_tri32(r0, r1, r3), c3 = cc_tri32(r0, r1, r2);
    auto one = [&](int vi, const f3 &pk, const f3 &pa, const f3 &pb, const f3 &pc, const f3 &fa,
                   const f3 &fb, const f3 &fc) {
        const int slot = slot_of(vi, sel4(tv, vi));
        if (slot >= 0)
            add(slot, static_cast<double>(subcell_volume32(pk, pa, pb, pc, ct, fc, fb, fa)));
    };
    one(0, r0, r1, r2, r3, c1, c2, c3);
    one(1, r1, r0, r2, r3, c0, c2, c3);
    one(2, r2, r0, r1, r3, c0, c1, c3);
    one(3, r3, r0, r1, r2, c0, c1, c2);
    return true;
}

template <typename SlotOf, typename Sub>
__device__ __forceinline__ bool new_contributions32(const DevMesh &m, int t, int fi, const d3 &q,
                                                    float tau, SlotOf slot_of, Sub sub) {
    bool neg;
    const int4 tv = ld_tet_v(m, t, neg);
    const int s[3] = {face_slot(fi, 0), face_slot(fi, 1), face_slot(fi, 2)};
    const int N = m.n_data;
    if (sel4(tv, s[0]) >= N || sel4(tv, s[1]) >= N || sel4(tv, s[2]) >= N) return false;
    const f3 F[3] = {rel32(ld_vert(m, sel4(tv, s[0])), q), rel32(ld_vert(m, sel4(tv, s[1])), q),
                     rel32(ld_vert(m, sel4(tv, s[2])), q)};
    const f3 o = {0.0f, 0.0f, 0.0f};  // the query
    const float tau2 = tau * tau;
    bool good = tet_cond(o, F[0], F[1], F[2]) >= tau;
#pragma unroll
    for (int k = 0; k < 3; k++)
        good = good && tri_cond(F[k], o, F[(k + 1) % 3]) >= tau2 &&
               tri_cond(F[k], o, F[(k + 2) % 3]) >= tau2;
    if (!good) return false;
    const f3 cv = cc_tet32(o, F[0], F[1], F[2]);
    const f3 cf = cc_tri32(F[0], F[1], F[2]);  // the order is immaterial in FP32
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int slot = slot_of(s[k], sel4(tv, s[k]));
        if (slot < 0) continue;
        const f3 &A = F[(k + 1) % 3];
        const f3 &B = F[(k + 2) % 3];
        const f3 ca = cc_tri32(F[k], o, A);
        const f3 cb = cc_tri32(F[k], o, B);
        sub(slot, static_cast<double>(subcell_volume32(F[k], o, A, B, cv, ca, cb, cf)));
    }
    return true;
}

template <typename F, bool RAW = false>
struct MomentFn {
    F f;
    static constexpr bool active = true;
    static constexpr bool raw = RAW;  // f gets 6 |w|^2 times the moment
    __device__ void operator()(int slot, const d3 &gm) const { f(slot, gm); }
};
template <typename F>
__device__ __forceinline__ MomentFn<F> moment_fn(F f) {
    return MomentFn<F>{f};
}
template <typename F>
__device__ __forceinline__ MomentFn<F, true> moment6_fn(F f) {
    return MomentFn<F, true>{f};
}

// a_j = <dL/dy[qi], values[j]>
__device__ __forceinline__ double grad_dot(const GradIn &g, int qi, int j) {
    const double *gr = g.gy + static_cast<size_t>(qi) * g.D;
    const double *vr = g.values + static_cast<size_t>(j) * g.D;
    double a = 0.0;
    for (int d = 0; d < g.D; d++) a += gr[d] * vr[d];
    return a;
}

// dL/dq of one row, accumulated while its volumes are computed.  With
// sigma_j the stolen volume of neighbour j, S = sum sigma_j,
// lambda_j = sigma_j / S, G_j = d sigma_j / dq and a_j = <dL/dy, values[j]>:
//   dL/dq = sum_j a_j d lambda_j / dq = (sum_j a_j G_j - abar sum_j G_j) / S,
//   abar = sum_j a_j sigma_j / S,
// so four running sums replace per-neighbour storage.  Rows the reference
// leaves empty (S <= 1e-30) get 0.
struct GradAcc {
    double s = 0.0, as = 0.0;
    d3 ag = {0.0, 0.0, 0.0}, tg = {0.0, 0.0, 0.0};
    __device__ void vol(double a, double v) {
        s += v;
        as += a * v;
    }
    __device__ void mom(double a, const d3 &gm) {
        ag = {ag.x + a * gm.x, ag.y + a * gm.y, ag.z + a * gm.z};
        tg = {tg.x + gm.x, tg.y + gm.y, tg.z + gm.z};
    }
    __device__ d3 result() const {
        if (!(s > 1e-30)) return {0.0, 0.0, 0.0};
        const double inv = 1.0 / s;
        const double abar = as * inv;
        return {(ag.x - abar * tg.x) * inv, (ag.y - abar * tg.y) * inv,
                (ag.z - abar * tg.z) * inv};
    }
};

__device__ __forceinline__ void store_dq(const GradIn &g, int qi, const d3 &v) {
    g.dq[3 * static_cast<size_t>(qi)] = v.x;
    g.dq[3 * static_cast<size_t>(qi) + 1] = v.y;
    g.dq[3 * static_cast<size_t>(qi) + 2] = v.z;
}

__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
    const int lane = __lane_id();
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// Reserves one pool range for the wave's rows; returns this lane's offset.
__device__ __forceinline__ long long wave_reserve(unsigned long long *counter, int cnt) {
    int total;
    const int pre = wave_excl_scan(cnt, total);
    unsigned long long base = 0;
    if (__lane_id() == 0 && total > 0)
        base = atomicAdd(counter, static_cast<unsigned long long>(total));
    base = __shfl(base, 0, 64);
    return static_cast<long long>(base) + pre;
}

// Work split by XCD: items [0, n) in kParts contiguous parts, part p with its
// own counter ctr[p].  Workgroups are dispatched to the 8 XCDs round-robin,
// so the blocks of one part share an L2 and a compact run of the (Morton)
// item order, and the mesh data of the queries in flight on an XCD fits its
// L2.  A wave drains its own part, then helps the others; k counts the parts
// it found exhausted.  Returns false when all are; else the batch's first
// item and size (<= batch).
constexpr int kParts = 8;
__device__ __forceinline__ bool part_batch(unsigned int *ctr, int n, int batch, int &k,
                                           int &first, int &cnt) {
    while (k < kParts) {
        const int p = static_cast<int>((blockIdx.x + static_cast<unsigned>(k)) % kParts);
        const int lo = static_cast<int>(static_cast<long long>(n) * p / kParts);
        const int hi = static_cast<int>(static_cast<long long>(n) * (p + 1) / kParts);
        unsigned b = 0;
        if (__lane_id() == 0) b = atomicAdd(&ctr[p], static_cast<unsigned>(batch));
        b = __shfl(b, 0, 64);
        if (static_cast<long long>(lo) + b < hi) {
            first = lo + static_cast<int>(b);
            cnt = min(batch, hi - first);
            return true;
        }
        k++;
    }
    return false;
}

// part_batch in two halves, so that the claim's atomic is in flight while
// the wave works on its current batch: part_claim_issue takes the range
// (lane 0's old counter value; meaningful for the part it names),
// part_claim_finish reads it and, if that part is spent, goes on with
// part_batch over the remaining parts.
__device__ __forceinline__ unsigned part_claim_issue(unsigned int *ctr, int batch, int k) {
    unsigned b = 0;
    if (k < kParts && __lane_id() == 0) {
        const int p = static_cast<int>((blockIdx.x + static_cast<unsigned>(k)) % kParts);
        b = atomicAdd(&ctr[p], static_cast<unsigned>(batch));
    }
    return b;
}
__device__ __forceinline__ bool part_claim_finish(unsigned int *ctr, int n, int batch, unsigned b,
                                                  int &k, int &first, int &cnt) {
    if (k >= kParts) return false;
    b = __shfl(b, 0, 64);
    const int p = static_cast<int>((blockIdx.x + static_cast<unsigned>(k)) % kParts);
    const int lo = static_cast<int>(static_cast<long long>(n) * p / kParts);
    const int hi = static_cast<int>(static_cast<long long>(n) * (p + 1) / kParts);
    if (static_cast<long long>(lo) + b < hi) {
        first = lo + static_cast<int>(b);
        cnt = min(batch, hi - first);
        return true;
    }
    k++;
    return part_batch(ctr, n, batch, k, first, cnt);
}

// The wave's next 64 work items; returns false when the list is exhausted.
__device__ __forceinline__ bool next_batch(unsigned int *work, int n_items, int &item) {
    unsigned base = 0;
    if (__lane_id() == 0) base = atomicAdd(work, 64u);
    base = __shfl(base, 0, 64);
    item = static_cast<int>(base) + __lane_id();
    return base < static_cast<unsigned>(n_items);
}

// --------------------------------------------------------- fused row kernel

// One query with this level's capacities: the whole of
// getWeightsSingleQueryIF (natural_insertionfree.c:591-738).  Returns the
// row length (row in nb_idx / nb_vol) or -1 when a capacity was exceeded.
// GRAD: the same cavity and volumes, but the row is reduced to dL/dq of
// query qi (written to *dq; nb_vol then holds <dL/dy, values[j]> per slot).
template <bool GRAD, bool ROBUST>
__device__ int fused_row(const DevMesh &m, const Caps &caps, const Scratch &sc, int tid,
                         int stamp, const d3 &q, int qi, const GradIn *g = nullptr,
                         d3 *dq = nullptr) {
    const size_t S = static_cast<size_t>(sc.nthr);
    const int N = m.n_data;
    bool ovf = false;
    const int start = locate<ROBUST>(m, loc_of(m, qi), q, sc, tid, caps, ovf);
    if (ovf) return -1;
    if (start < 0) return 0;  // outside the super simplex: empty row
    const HashTab vis(sc.vis, sc.vis_pos, tid, caps.vis_log2, stamp);
    const int ncav = cavity_search<ROBUST>(m, caps, sc, tid, vis, start, q);
    if (ncav < 0) return -1;
    if (GRAD) *dq = {0.0, 0.0, 0.0};  // empty and coincident rows
    if (ncav == 0) {
        if (GRAD) return 0;
        int coin[4];
        const int nco = coincident(m, start, q, coin);
        if (nco > 0) {
            const double w = 1.0 / nco;
            for (int i = 0; i < nco; i++) {
                sc.nb_idx[il(i, S, tid)] = coin[i];
                sc.nb_vol[il(i, S, tid)] = w;
            }
        }
        return nco;
    }
    // boundary faces and neighbour order (natural_insertionfree.c:559-584)
    const HashTab nbm(sc.nbmap, sc.nbslot, tid, caps.nbmap_log2, stamp);
    int nnb = 0, nbnd = 0;
    for (int ci = 0; ci < ncav; ci++) {
        const int t = sc.cav[il(ci, S, tid)];
        const int4 tn = m.tet_n[t];
        bool neg;
        const int4 tv = ld_tet_v(m, t, neg);
        for (int fi = 0; fi < 4; fi++) {
            const int nb = sel4(tn, fi);
            if (nb >= 0 && vis.get(nb) >= 0) continue;
            if (nbnd >= caps.bnd) return -1;
            sc.bnd[il(nbnd, S, tid)] = (ci << 2) | fi;
            nbnd++;
            for (int k = 0; k < 3; k++) {
                const int v = sel4(tv, face_slot(fi, k));
                if (v >= N) continue;
                bool found;
                const unsigned h = nbm.insert(v, found);
                if (found) continue;
                if (nnb >= caps.nb) return -1;
                nbm.set(h, nnb);
                sc.nb_idx[il(nnb, S, tid)] = v;
                sc.nb_vol[il(nnb, S, tid)] = 0.0;
                nnb++;
            }
        }
    }
    if (nnb == 0) return 0;
    auto slot_of = [&](int, int v) { return v < N ? nbm.get(v) : -1; };
    auto fc_mask = [&](int ci, int t) {
        const int4 tn = m.tet_n[t];
        unsigned mask = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int nb = sel4(tn, 3 - k);
            if (nb < 0) continue;
            const int cj = vis.get(nb);
            if (cj >= 0 && cj < ci) mask |= 1u << k;
        }
        return mask;
    };
    if constexpr (GRAD) {
        for (int j = 0; j < nnb; j++)
            sc.nb_vol[il(j, S, tid)] = grad_dot(*g, qi, sc.nb_idx[il(j, S, tid)]);
        GradAcc acc;
        auto a_of = [&](int slot) { return sc.nb_vol[il(slot, S, tid)]; };
        for (int ci = 0; ci < ncav; ci++) {
            const int t = sc.cav[il(ci, S, tid)];
            old_contributions(m, t, fc_mask(ci, t), slot_of,
                              [&](int slot, double vol) { acc.vol(a_of(slot), vol); });
        }
        for (int bi = 0; bi < nbnd; bi++) {
            const int code = sc.bnd[il(bi, S, tid)];
            new_contributions(
                m, sc.cav[il(code >> 2, S, tid)], code & 3, q, slot_of,
                [&](int slot, double vol) { acc.vol(a_of(slot), -vol); },
                moment_fn([&](int slot, const d3 &gm) { acc.mom(a_of(slot), gm); }));
        }
        *dq = acc.result();
        return 0;
    }
    auto add = [&](int slot, double vol) { sc.nb_vol[il(slot, S, tid)] += vol; };
    auto sub = [&](int slot, double vol) { sc.nb_vol[il(slot, S, tid)] -= vol; };
    for (int ci = 0; ci < ncav; ci++) {
        const int t = sc.cav[il(ci, S, tid)];
        old_contributions(m, t, fc_mask(ci, t), slot_of, add);
    }
    for (int bi = 0; bi < nbnd; bi++) {
        const int code = sc.bnd[il(bi, S, tid)];
        new_contributions(m, sc.cav[il(code >> 2, S, tid)], code & 3, q, slot_of, sub);
    }
    // step 5, natural_insertionfree.c:721-737
    double total = 0.0;
    for (int j = 0; j < nnb; j++) total += sc.nb_vol[il(j, S, tid)];
    if (!(total > 1e-30)) return 0;
    const double inv = 1.0 / total;
    for (int j = 0; j < nnb; j++) sc.nb_vol[il(j, S, tid)] *= inv;
    return nnb;
}

// n_items_dev, when given, holds the item count (written by an earlier kernel
// of the same stream: the graph form of the launch sequence).  Queries whose
// problem starts its thread-per-query rows at a later level pass straight on.
template <bool ROBUST>
__global__ void __launch_bounds__(256) sibson_rows_kernel(DevMesh m, Caps caps, Scratch sc,
                                                          int level,
                                                          const double *__restrict__ q,
                                                          const int *__restrict__ items,
                                                          int n_items,
                                                          const unsigned *n_items_dev,
                                                          unsigned int *work, RowOut out) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= sc.nthr) return;  // nthr is a multiple of 64: whole waves leave
    if (n_items_dev) n_items = static_cast<int>(*n_items_dev);
    const size_t S = static_cast<size_t>(sc.nthr);
    int epoch = sc.epoch[tid];
    int item;
    while (next_batch(work, n_items, item)) {
        int qi = -1, cnt = 0;
        if (item < n_items) {
            qi = items ? items[item] : item;
            if (static_cast<int>(route_of(m, qi) >> 8) > level) {
                cnt = -1;
            } else {
                epoch = epoch == 0x7ffffffe ? 1 : epoch + 1;
                const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                               q[3 * static_cast<size_t>(qi) + 2]};
                cnt = fused_row<false, ROBUST>(m, caps, sc, tid, epoch, qq, qi);
            }
            if (cnt < 0) {
                out.ovf_list[atomicAdd(out.ovf_count, 1u)] = qi;
                cnt = 0;
            }
        }
        const long long off = wave_reserve(out.pool_used, cnt);
        if (qi >= 0) {
            out.row_off[qi] = off;
            out.row_cnt[qi] = cnt;
            if (off + cnt <= out.pool_cap) {
                for (int j = 0; j < cnt; j++) {
                    out.pool_cols[off + j] = sc.nb_idx[il(j, S, tid)];
                    out.pool_vals[off + j] = sc.nb_vol[il(j, S, tid)];
                }
            }
        }
    }
    sc.epoch[tid] = epoch;
}

// Gradient of the rows of the given queries (those without a topology
// record): the fused row with the gradient reduction.  Queries that exceed
// the level's capacities go to ovf_list for the next level.
template <bool ROBUST>
__global__ void __launch_bounds__(256) sibson_grad_rows_kernel(
    DevMesh m, Caps caps, Scratch sc, int level, const double *__restrict__ q,
    const int *__restrict__ items, int n_items, unsigned int *work, GradIn g, int *ovf_list,
    unsigned int *ovf_count) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= sc.nthr) return;
    int epoch = sc.epoch[tid];
    int item;
    while (next_batch(work, n_items, item)) {
        if (item >= n_items) continue;
        const int qi = items[item];
        if (static_cast<int>(route_of(m, qi) >> 8) > level) {
            ovf_list[atomicAdd(ovf_count, 1u)] = qi;
            continue;
        }
        epoch = epoch == 0x7ffffffe ? 1 : epoch + 1;
        const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                       q[3 * static_cast<size_t>(qi) + 2]};
        d3 dq = {0.0, 0.0, 0.0};  // rows outside the super simplex return before setting it
        if (fused_row<true, ROBUST>(m, caps, sc, tid, epoch, qq, qi, &g, &dq) < 0)
            ovf_list[atomicAdd(ovf_count, 1u)] = qi;
        else
            store_dq(g, qi, dq);
    }
    sc.epoch[tid] = epoch;
}

// ------------------------------------------------------------- start tets

// Start tet of every work item: exact nearest centroid (grid; the k-d
// replica when the grid cannot certify it), then the visibility walk.
// start[qi] = tet, -1 outside the super simplex, -2 k-d stack overflow.
// 8 waves per SIMD: the search is latency-bound and uses 32 VGPRs; without
// the hint the compiler takes 112 SGPRs, which allow only 7 (800 per SIMD).
template <bool ROBUST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
locate_kernel(DevMesh m, Caps caps, Scratch sc,
                                                     const double *__restrict__ q,
                                                     const int *__restrict__ items, int n_items,
                                                     int *__restrict__ start) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= sc.nthr) return;  // sc.nthr: the scratch stride, at least the threads used
    // the blocks of XCD p (block b runs on XCD b mod kParts) stride over part
    // p of the Morton-ordered items, so each L2 sees one compact region
    const int nthr = min(static_cast<int>(gridDim.x * blockDim.x), sc.nthr);
    const int nblk = nthr / static_cast<int>(blockDim.x);
    const int parts = nblk >= kParts ? kParts : 1;
    const int p = static_cast<int>(blockIdx.x) % parts;
    const int lo = static_cast<int>(static_cast<long long>(n_items) * p / parts);
    const int hi = static_cast<int>(static_cast<long long>(n_items) * (p + 1) / parts);
    const int pblk = (nblk - p + parts - 1) / parts;  // blocks of this part
    const int stride = pblk * static_cast<int>(blockDim.x);
    for (int i = lo + static_cast<int>(blockIdx.x) / parts * static_cast<int>(blockDim.x) +
                 static_cast<int>(threadIdx.x);
         i < hi; i += stride) {
        const int qi = items ? items[i] : i;
        const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                       q[3 * static_cast<size_t>(qi) + 2]};
        bool ovf = false;
        const int s = locate<ROBUST>(m, loc_of(m, qi), qq, sc, tid, caps, ovf);
        start[qi] = ovf ? -2 : s;
    }
}

// Topology of a query, in three kernels.
//
// topo_search_kernel: 16 lanes for a quad of four consecutive queries of the
// Morton order (4 quads per 64-thread block, tables in LDS).  The queries of
// a quad are close together and their cavities overlap, so one search serves
// all four: it tests every tet any of them pops in the reference's search --
// each start tet and every neighbour of a tet whose circumsphere holds one of
// the queries -- against each query, front by front, the tests of one front
// in parallel; a tet is expanded when it is inside any query's sphere.  A
// query's decisions are the same FP64 insphere tests the reference makes, and
// every tet its own search pops is among those tested.  Per query, the tets
// inside its sphere get dense indices (its start tet 0), its record is
// reserved, and a handoff pair per dense index (tet, in-sphere adjacency: the
// dense index of the in-sphere neighbour across each face, 0xff across a
// face to a tet outside the sphere or to none, i.e. a boundary face) goes to
// the end of the record.
//
// topo_dfs_kernel and topo_order_kernel (below) turn the adjacency into
// what the reference's order rules fix: the cavity order (the
// LIFO search, natural_insertionfree.c:498-530, is a depth-first preorder of
// the in-sphere tets that enters neighbour slots 3, 2, 1, 0 in that order:
// a tet outside the sphere is never appended or expanded, so it does not
// change the order of the others), the boundary faces in (cavity position,
// face) order, the face-circumcentre masks, and the neighbours in order of
// first appearance in the face-vertex sequence
// (natural_insertionfree.c:373-382, 559-584); then they complete the record.
// A query whose in-sphere tets are not all reachable from its start tet
// through in-sphere tets (the quad's search can find such tets; a search of
// its own would not) goes to the larger levels, which search it alone.
//
// Record of one query, in 16-byte units at topo.rec[rec_off[qi]]:
//   [0]            (C, B, NB, 0): cavity tets, boundary faces, neighbours
//   [1 .. C]       per cavity tet in cavity order: (tet, slot(p0) | slot(p1)
//                  << 16, slot(p2) | slot(p3) << 16, fcmask | table position
//                  << 4) -- slot 0xffff = none; fcmask bit k: the face
//                  opposite p[k] is held by a cavity tet earlier in cavity
//                  order (its circumcentre uses that tet's vertex order); the
//                  position is table_pos(fcmask, sphere_dm(sph[tet])) (0
//                  without the step-3 table)
//   next ceil(C/8)   the cavity order, 16 bits per position: its dense index
//                    | the fcmask of that tet << 6
//   next ceil(B/16)  boundary faces in sequence order, a byte each: cavity
//                    position << 2 | face
//   next ceil(NB/4)  neighbour vertex indices in slot order (room is reserved
//                    for min(32, B/2 + 4): a cavity bounded by one closed
//                    surface has B/2 + 2 boundary vertices)
//   next 2 ceil(C/4) handoff: the in-sphere adjacency, then the tet (and
//                    its table mask: handoff_tet, handoff_dm), per dense
//                    index (int each)
// C = 0 rows: NB = vertices coinciding with the query (weight 1/NB each, in
// unit [1]), or NB = 0 for an empty row.  Between the kernels the header is
// (C, B, neighbour room, 1) and the entries are not written yet; the order
// kernel writes them in cavity order, so the geometry reads them in
// sequence.
namespace grp {
constexpr int G = 16;             // lanes per quad
constexpr int Q = 4;              // queries per quad
constexpr int QPB = 4;            // quads per 64-thread block
constexpr int VT = 255;           // visited-table slots (0xff = none)
constexpr int VMAX = 192;         // visited tets of a quad
constexpr int VQ = 96;            // ... counted per query of the quad (budget)
constexpr int CMAX = 64;          // cavity tets of a query
constexpr int BMAX = 136;         // boundary faces (a cavity of C tets has about 2C + 2)
constexpr int NMAX = kLevel0Row;  // neighbours
constexpr int NT_LOG2 = 6;
constexpr int NT = 1 << NT_LOG2;  // vertex-table slots of the order kernel
}  // namespace grp

__device__ __forceinline__ int nb_room(int B) { return min(grp::NMAX, B / 2 + 4); }
// offsets of the sections of a record with C cavity tets, B boundary faces and
// room for `room` neighbours, in 16-byte units; every kernel that reads or
// writes a section takes its offset from here
__device__ __forceinline__ int rec_ord_off(int C) { return 1 + C; }
__device__ __forceinline__ int rec_bnd_off(int C) { return rec_ord_off(C) + (C + 7) / 8; }
__device__ __forceinline__ int rec_nbv_off(int C, int B) { return rec_bnd_off(C) + (B + 15) / 16; }
__device__ __forceinline__ int rec_x_off(int C, int B, int room) {
    return rec_nbv_off(C, B) + (room + 3) / 4;
}
__device__ __forceinline__ int rec_units(int C, int B, int room) {
    return rec_x_off(C, B, room) + 2 * ((C + 3) / 4);
}
// neighbour slot of vertex vi of a record entry cr (-1: none)
__device__ __forceinline__ int rec_slot(const int4 &cr, int vi) {
    const int w = vi < 2 ? cr.y : cr.z;
    const int s = (vi & 1) ? ((w >> 16) & 0xffff) : (w & 0xffff);
    return s == 0xffff ? -1 : s;
}

// sections of a complete record whose header is hd
__device__ __forceinline__ const unsigned short *rec_order(const int4 *r, const int4 &hd) {
    return reinterpret_cast<const unsigned short *>(r + rec_ord_off(hd.x));
}
__device__ __forceinline__ const unsigned char *rec_bnd(const int4 *r, const int4 &hd) {
    return reinterpret_cast<const unsigned char *>(r + rec_bnd_off(hd.x));
}
__device__ __forceinline__ const int *rec_nbv(const int4 *r, const int4 &hd) {
    return reinterpret_cast<const int *>(r + rec_nbv_off(hd.x, hd.y));
}
// the handoff arrays of a record in flight (header (C, B, room, 1))
__device__ __forceinline__ int *rec_xadj(int4 *r, int C, int B, int room) {
    return reinterpret_cast<int *>(r + rec_x_off(C, B, room));
}
__device__ __forceinline__ int *rec_xtet(int4 *r, int C, int B, int room) {
    return reinterpret_cast<int *>(r + rec_x_off(C, B, room) + (C + 3) / 4);
}

// The handoff's tet entries (rec_xtet) carry the tet's step-3 table mask
// (sphere_dm) in bits 27..30 when every tet index fits in 27 bits; the order
// kernel then needs no load for it.
__device__ __forceinline__ bool dm_packs(const DevMesh &m) { return m.n_tets <= (1 << 27); }
__device__ __forceinline__ int handoff_tet(const DevMesh &m, int xt) {
    return dm_packs(m) ? (xt & ((1 << 27) - 1)) : xt;
}
__device__ __forceinline__ unsigned handoff_dm(const DevMesh &m, int xt) {
    return dm_packs(m) ? static_cast<unsigned>(xt) >> 27 : sphere_dm(m.sph[xt]);
}

struct alignas(16) QuadMem {
    // the hash of the search, then (its keys no longer read) the dense maps
    union {
        int vkey[256];                     // tet index, -1 = empty slot (slot 255 unused)
        unsigned char dense[grp::Q][256];  // per query: dense index of an in-sphere slot
    };
    unsigned vnb[256];             // visited slots of the 4 neighbours, 8 bits each (0xff none)
    unsigned char vflag[256];      // bit 0 tested; bit 1 + j inside query j's circumsphere
    unsigned char vlist[grp::VMAX];    // visited slots in test order; fronts are ranges
    int vtet[grp::VMAX];           // their tets
    int sslot[grp::Q];             // slot of each query's start tet
    int cnt[4];                    // [0] first front, [3] overflow flag
};

__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned vt_hash(int t) {
    return (hash_key(t, 8) * 255u) >> 8;  // 0..254
}
// visited-table insert of tet t, probing from slot h: its slot (isnew when
// this call put it there), -1 when the table is full
__device__ __forceinline__ int vis_probe(QuadMem &g, int t, unsigned h, bool &isnew) {
    for (int n = 0; n < grp::VT; n++) {
        const int old = atomicCAS(&g.vkey[h], -1, t);
        if (old == -1 || old == t) {
            isnew = old == -1;
            return static_cast<int>(h);
        }
        h = h + 1 == grp::VT ? 0 : h + 1;
    }
    isnew = false;
    return -1;
}

// The four neighbours of visited slot s: their slots (0xff none) and flags
// (0 for none), read once for all the quad's queries.
struct SlotNbrs {
    unsigned nbs;  // vnb[s]
    unsigned fl[4];
};
__device__ __forceinline__ SlotNbrs slot_nbrs(const QuadMem &g, int s) {
    SlotNbrs n;
    n.nbs = g.vnb[s];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned ns = (n.nbs >> (8 * k)) & 0xffu;
        n.fl[k] = ns != 0xffu ? g.vflag[ns] : 0u;
    }
    return n;
}
// Query j's in-sphere adjacency word of visited slot s: byte k = the dense
// index of the neighbour across face k when that is inside query j's
// circumsphere, else 0xff.
__device__ __forceinline__ unsigned adjacency_of(const QuadMem &g, const SlotNbrs &n, int j) {
    const unsigned bit = 2u << j;
    unsigned a = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned ns = (n.nbs >> (8 * k)) & 0xffu;
        const unsigned d = (n.fl[k] & bit) ? g.dense[j][ns] : 0xffu;
        a |= d << (8 * k);
    }
    return a;
}

// 64-thread blocks of 4 quads; 13.2 KB of LDS per block allow 12 per CU, and
// at most 168 VGPRs keep 3 waves per SIMD
template <bool ROBUST>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
topo_search_kernel(DevMesh m, const double *__restrict__ q, const int *__restrict__ items,
                   int n_items, const int *__restrict__ start_of, unsigned int *work, Topo topo,
                   unsigned int *ovf_count, int *ovf_list) {
    __shared__ QuadMem gmem[grp::QPB];
    const int lane = __lane_id();
    const int gl = lane & (grp::G - 1);
    const int gw = lane >> 4;
    QuadMem &g = gmem[gw];
    constexpr int kBatch = grp::Q * grp::QPB;
    // Software pipeline over the wave's batches: the next batch is claimed
    // when the current one starts, its items are read after the current
    // search, its queries and start tets before the current handoff, so
    // their latencies pass while the current batch is worked on.  Lane j
    // of a quad holds query j's values; the quad's lanes share them.
    int parts_done = 0, nfirst = 0, ncnt = 0;
    bool have = part_batch(work, n_items, kBatch, parts_done, nfirst, ncnt);
    int nqil = -1, nstl = -1, nbl = 0;
    d3 nql = {0, 0, 0};
    auto load_items = [&]() {
        const int item = grp::Q * gw + gl;
        nqil = have && gl < grp::Q && item < ncnt ? items[nfirst + item] : -1;
    };
    auto load_queries = [&]() {
        nstl = -1;
        nbl = 0;
        nql = {0, 0, 0};
        if (nqil >= 0) {
            nstl = start_of[nqil];
            nql = {q[3 * static_cast<size_t>(nqil)], q[3 * static_cast<size_t>(nqil) + 1],
                   q[3 * static_cast<size_t>(nqil) + 2]};
            nbl = min(grp::VQ, static_cast<int>(route_of(m, nqil) & 0xffu));
        }
    };
    load_items();
    load_queries();
    while (have) {
        // query j of this quad: item first + 4 gw + j
        const int first = nfirst;
        const int qil = nqil, stl = nstl, bl = nbl;
        const d3 ql = nql;
        const unsigned claim = part_claim_issue(work, kBatch, parts_done);
        int qi[grp::Q], st[grp::Q];
        d3 qq[grp::Q];
        bool act[grp::Q], ok[grp::Q];
        int budget = 0;
#pragma unroll
        for (int j = 0; j < grp::Q; j++) {
            qi[j] = __shfl(qil, j, grp::G);
            st[j] = __shfl(stl, j, grp::G);
            qq[j] = {__shfl(ql.x, j, grp::G), __shfl(ql.y, j, grp::G), __shfl(ql.z, j, grp::G)};
            const int b = __shfl(bl, j, grp::G);
            // a query outside the super simplex (st -1) has an empty row; k-d
            // overflow (st -2) and budget 0 go to the larger levels
            act[j] = qi[j] >= 0 && st[j] >= 0 && b > 0;
            ok[j] = qi[j] >= 0 && st[j] != -2 && b > 0;
            if (act[j]) budget += b;
        }
        const int vcap = min(grp::VMAX, budget);
        const bool any = act[0] || act[1] || act[2] || act[3];
        int ncav[grp::Q] = {0, 0, 0, 0}, nbnd[grp::Q] = {0, 0, 0, 0}, nnb[grp::Q] = {0, 0, 0, 0};
        int nvis = 0;
        if (any) {
            for (int i = gl; i < 64; i += grp::G) {
                reinterpret_cast<int4 *>(g.vkey)[i] = make_int4(-1, -1, -1, -1);
                reinterpret_cast<unsigned *>(g.vflag)[i] = 0u;
            }
            if (gl == 0) g.cnt[3] = 0;
            group_sync();
            // the first front: the start tets (lanes 0-3), then their
            // neighbours (lane 4 j + k: neighbour k of query j's start tet),
            // each group placed in lane order
            {
                const int jq = gl >> 2, kq = gl & 3;
                int stj = -1;
#pragma unroll
                for (int j = 0; j < grp::Q; j++)
                    if (jq == j && act[j]) stj = st[j];
                const int nbk = stj >= 0 ? sel4(m.tet_n[stj], kq) : -1;
                bool isnew = false;
                int s0 = -1;
                if (gl < grp::Q) {
                    int stl = -1;
#pragma unroll
                    for (int j = 0; j < grp::Q; j++)
                        if (gl == j && act[j]) stl = st[j];
                    if (stl >= 0) s0 = vis_probe(g, stl, vt_hash(stl), isnew);
                    g.sslot[gl] = s0;
                    const unsigned mk = static_cast<unsigned>(__ballot(isnew) >> (16 * gw)) &
                                        0xfu;
                    if (isnew) {
                        const int p = __popc(mk & ((1u << gl) - 1u));
                        g.vlist[p] = static_cast<unsigned char>(s0);
                        g.vtet[p] = stl;
                    }
                }
                const int n0 = __popc(static_cast<unsigned>(__ballot(gl < grp::Q && isnew) >>
                                                            (16 * gw)) & 0xffffu);
                group_sync();
                bool isnew2 = false;
                int s2 = -1;
                if (nbk >= 0) s2 = vis_probe(g, nbk, vt_hash(nbk), isnew2);
                const bool add = nbk >= 0 && s2 >= 0 && isnew2;
                const unsigned mk2 = static_cast<unsigned>(__ballot(add) >> (16 * gw)) & 0xffffu;
                if (add) {
                    const int p = n0 + __popc(mk2 & ((1u << gl) - 1u));
                    g.vlist[p] = static_cast<unsigned char>(s2);
                    g.vtet[p] = nbk;
                }
                if (gl == 0) g.cnt[0] = n0 + __popc(mk2);
            }
            group_sync();
            int f0 = 0, f1 = g.cnt[0];
            while (f0 < f1) {
                int nnext = 0;  // size of the next front (uniform over the 16 lanes)
                for (int i0 = f0; i0 < f1; i0 += grp::G) {
                    const int i = i0 + gl;
                    unsigned newslots = 0;
                    int nnew = 0;
                    int nt0 = 0, nt1 = 0, nt2 = 0, nt3 = 0;  // tets of the new slots
                    if (i < f1) {
                        const int slot = g.vlist[i];
                        const int t = g.vtet[i];
                        // the certified circumsphere test first; the
                        // determinant only where it cannot decide
                        // (one copy of the record: every query's test
                        // reads it, so its loads are issued once, together)
                        const SphereTet sp = m.sph[t];
                        const int4 tn = sp.n;
                        unsigned in4 = 0, unc = 0;
#pragma unroll
                        for (int j = 0; j < grp::Q; j++) {
                            const int sd = sphere_side(sp, qq[j]);
                            in4 |= (act[j] && sd > 0 ? 1u : 0u) << j;
                            unc |= (act[j] && sd < 0 ? 1u : 0u) << j;
                        }
                        if (unc) {
                            const FatTet &ft = m.fat[t];
                            d3 p0, p1, p2, p3;
                            fat_points(ft, p0, p1, p2, p3);
                            if (ft.v.w < 0) {
                                const d3 tmp = p0;
                                p0 = p1;
                                p1 = tmp;
                            }
#pragma unroll
                            for (int j = 0; j < grp::Q; j++)
                                if (((unc >> j) & 1u) &&
                                    inside_sphere<ROBUST>(m, p0, p1, p2, p3, qq[j]))
                                    in4 |= 1u << j;
                        }
                        unsigned packed = 0xffffffffu;
                        if (in4) {
                            packed = 0;
                            // first probes of the four neighbours issued together
                            int nb[4], old[4];
                            unsigned hh[4];
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                nb[k] = sel4(tn, k);
                                hh[k] = vt_hash(nb[k]);
                                old[k] = nb[k] >= 0 ? atomicCAS(&g.vkey[hh[k]], -1, nb[k]) : -2;
                            }
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                unsigned ns = 0xff;
                                if (nb[k] >= 0) {
                                    bool isnew = old[k] == -1;
                                    int s2 = static_cast<int>(hh[k]);
                                    if (old[k] != -1 && old[k] != nb[k])
                                        s2 = vis_probe(g, nb[k],
                                                       hh[k] + 1 == grp::VT ? 0 : hh[k] + 1, isnew);
                                    if (s2 < 0) {
                                        g.cnt[3] = 1;
                                    } else {
                                        ns = static_cast<unsigned>(s2);
                                        if (isnew) {
                                            nt0 = nnew == 0 ? nb[k] : nt0;
                                            nt1 = nnew == 1 ? nb[k] : nt1;
                                            nt2 = nnew == 2 ? nb[k] : nt2;
                                            nt3 = nnew == 3 ? nb[k] : nt3;
                                            newslots |= ns << (8 * nnew++);
                                        }
                                    }
                                }
                                packed |= ns << (8 * k);
                            }
                        }
                        g.vnb[slot] = packed;
                        // an in-sphere tet carries its step-3 table mask to
                        // the handoff
                        if (in4 && dm_packs(m)) g.vtet[i] = t | static_cast<int>(sphere_dm(sp) << 27);
                        g.vflag[slot] = static_cast<unsigned char>(1u | (in4 << 1));
                    }
                    // positions in the next front: exclusive scan over the 16 lanes
                    int incl = nnew;
#pragma unroll
                    for (int d = 1; d < grp::G; d <<= 1) {
                        const int y = __shfl_up(incl, d, grp::G);
                        if (gl >= d) incl += y;
                    }
                    const int p = f1 + nnext + incl - nnew;
                    for (int j = 0; j < nnew; j++) {
                        if (p + j < grp::VMAX) {
                            g.vlist[p + j] =
                                static_cast<unsigned char>((newslots >> (8 * j)) & 0xffu);
                            g.vtet[p + j] = j == 0 ? nt0 : (j == 1 ? nt1 : (j == 2 ? nt2 : nt3));
                        } else {
                            g.cnt[3] = 1;
                        }
                    }
                    nnext += __shfl(incl, grp::G - 1, grp::G);
                }
                group_sync();
                f0 = f1;
                f1 += nnext;
                if (f1 > vcap || g.cnt[3]) break;
            }
            const bool qok = g.cnt[3] == 0 && f1 <= vcap;
            nvis = qok ? f1 : 0;  // the list is valid up to f1 only when nothing overflowed
            // per query: its start tet inside its sphere, else the
            // coincident-vertex path (natural_insertionfree.c:596-622)
            unsigned cav_q = 0;  // queries with a cavity
#pragma unroll
            for (int j = 0; j < grp::Q; j++) {
                if (!act[j]) continue;
                if (!qok) {
                    ok[j] = false;
                    continue;
                }
                if (g.vflag[g.sslot[j]] & (2u << j)) {
                    cav_q |= 1u << j;
                } else {
                    int4 coin;
                    nnb[j] = coincident4(m, st[j], qq[j], coin);
                }
            }
            if (cav_q) {
                // dense indices in test order, the start tet 0
#pragma unroll
                for (int j = 0; j < grp::Q; j++) ncav[j] = (cav_q >> j) & 1u;
                for (int i0 = 0; i0 < nvis; i0 += grp::G) {
                    const int i = i0 + gl;
                    const int s = i < nvis ? g.vlist[i] : -1;
                    const unsigned fl = s >= 0 ? g.vflag[s] : 0u;
#pragma unroll
                    for (int j = 0; j < grp::Q; j++) {
                        if (!((cav_q >> j) & 1u)) continue;
                        const bool in = (fl & (2u << j)) && s != g.sslot[j];
                        const unsigned mine =
                            static_cast<unsigned>(__ballot(in) >> (16 * gw)) & 0xffffu;
                        if (in) {
                            const int d = ncav[j] + __popc(mine & ((1u << gl) - 1u));
                            g.dense[j][s] = static_cast<unsigned char>(min(d, 255));
                        }
                        ncav[j] += __popc(mine);
                    }
                }
                if (gl == 0)
#pragma unroll
                    for (int j = 0; j < grp::Q; j++)
                        if ((cav_q >> j) & 1u) g.dense[j][g.sslot[j]] = 0;
                group_sync();
                // boundary faces: the 0xff bytes of the in-sphere adjacency
                int b[grp::Q] = {0, 0, 0, 0};
                for (int i = gl; i < nvis; i += grp::G) {
                    const int s = g.vlist[i];
                    const unsigned fl = g.vflag[s];
                    if (!(fl & (cav_q << 1))) continue;
                    const SlotNbrs nb = slot_nbrs(g, s);
#pragma unroll
                    for (int j = 0; j < grp::Q; j++) {
                        const unsigned bit = 2u << j;
                        if (!((cav_q >> j) & 1u) || !(fl & bit) || ncav[j] > grp::CMAX) continue;
                        // faces to a neighbour outside the circumsphere (or none)
#pragma unroll
                        for (int k = 0; k < 4; k++) b[j] += !(nb.fl[k] & bit);
                    }
                }
#pragma unroll
                for (int j = 0; j < grp::Q; j++) {
#pragma unroll
                    for (int d = 1; d < grp::G; d <<= 1) b[j] += __shfl_xor(b[j], d, grp::G);
                    nbnd[j] = b[j];
                    if ((cav_q >> j) & 1u) {
                        if (ncav[j] > grp::CMAX || nbnd[j] > grp::BMAX) {
                            ok[j] = false;
                            cav_q &= ~(1u << j);
                        }
                    } else {
                        ncav[j] = 0;
                    }
                }
            }
        }
        // reserve the records: lane j of the quad reserves query j's
        // the next batch: its claim is back by now; read its items (every
        // lane of the wave takes part: the claim is the wave's)
        have = part_claim_finish(work, n_items, kBatch, claim, parts_done, nfirst, ncnt);
        load_items();
        int units = 0;
#pragma unroll
        for (int j = 0; j < grp::Q; j++) {
            int u = 0;
            if (qi[j] >= 0 && ok[j])
                u = ncav[j] > 0 ? rec_units(ncav[j], nbnd[j], nb_room(nbnd[j]))
                                : 1 + (nnb[j] + 3) / 4;
            if (gl == j) units = u;
        }
        const long long off_l = wave_reserve(topo.rec_used, units);
        long long off[grp::Q];
        bool fits[grp::Q];
#pragma unroll
        for (int j = 0; j < grp::Q; j++) {
            off[j] = __shfl(off_l, j, grp::G);
            int u = 0;
            if (qi[j] >= 0 && ok[j])
                u = ncav[j] > 0 ? rec_units(ncav[j], nbnd[j], nb_room(nbnd[j]))
                                : 1 + (nnb[j] + 3) / 4;
            fits[j] = off[j] + u <= topo.rec_cap;
            if (gl == j && qi[j] >= 0) {
                const bool pending = ok[j] && fits[j] && ncav[j] > 0;
                if (topo.itab)
                    topo.itab[first + grp::Q * gw + j] = make_int4(
                        qi[j], static_cast<int>(off[j] & 0xffffffffLL),
                        static_cast<int>(off[j] >> 32),
                        pending ? ncav[j] | (nbnd[j] << 8) | (nb_room(nbnd[j]) << 16) : -1);
                if (!ok[j] || !fits[j]) {
                    topo.rec_off[qi[j]] = -1;
                    if (!ok[j]) ovf_list[atomicAdd(ovf_count, 1u)] = qi[j];
                } else {
                    topo.rec_off[qi[j]] = off[j];
                    int4 *r = topo.rec + off[j];
                    if (ncav[j] == 0) {
                        int4 coin = make_int4(0, 0, 0, 0);
                        if (nnb[j] > 0) coincident4(m, st[j], qq[j], coin);
                        r[0] = make_int4(0, 0, nnb[j], 0);
                        if (nnb[j] > 0) r[1] = coin;
                    } else {
                        r[0] = make_int4(ncav[j], nbnd[j], nb_room(nbnd[j]), 1);
                    }
                }
            }
        }
        load_queries();  // of the next batch
        // handoff pairs of the queries with a cavity (the order kernel reads
        // the tets' vertices itself)
        for (int i = gl; i < nvis; i += grp::G) {
            const int s = g.vlist[i];
            const unsigned fl = g.vflag[s];
            if (!(fl & 0x1eu)) continue;
            const int t = g.vtet[i];  // with the table mask (handoff_tet, handoff_dm)
            const SlotNbrs nb = slot_nbrs(g, s);
#pragma unroll
            for (int j = 0; j < grp::Q; j++) {
                if (qi[j] < 0 || !ok[j] || !fits[j] || ncav[j] == 0 || !(fl & (2u << j)))
                    continue;
                int4 *r = topo.rec + off[j];
                const int d = g.dense[j][s];
                const int room = nb_room(nbnd[j]);
                rec_xadj(r, ncav[j], nbnd[j], room)[d] = static_cast<int>(adjacency_of(g, nb, j));
                rec_xtet(r, ncav[j], nbnd[j], room)[d] = t;
            }
        }
        group_sync();
    }
}

// The order rules, in two kernels.  Block b of either takes a run of part
// b mod 8 of the items (the parts of the search kernel), so a query's record
// is read on the XCD whose L2 the search kernel wrote it through.
//
// topo_dfs_kernel, one thread per query (the walk is serial, and 64 of them
// share a wave): the depth-first preorder from the start tet (dense 0),
// entering neighbour slots 3, 2, 1, 0.  Appending the tet at position p also
// settles its part of the boundary sequence (its faces whose adjacency byte
// is 0xff, in face order: positions are appended in order) and its
// face-circumcentre mask (bit k: the neighbour across face 3 - k was appended
// before it).  Both go out as packed streams, a 32-bit store per two
// positions or four boundary faces.
constexpr int kDfsThreads = 64;
constexpr int kDfsPool = 2048;  // adjacency words of a wave kept in LDS (mean 64 x 20)

__global__ void __launch_bounds__(kDfsThreads)
topo_dfs_kernel(const int *__restrict__ items, int n_items, Topo topo, unsigned int *ovf_count,
                int *ovf_list, int pool_words) {
    // the wave's adjacency words: each query's C words at its offset in one
    // pool (a prefix sum over the lanes); a query past the pool's end reads
    // and writes its words in the record instead
    __shared__ unsigned pool[kDfsPool];
    const int t = static_cast<int>(threadIdx.x);
    const int part = static_cast<int>(blockIdx.x) % kParts;
    const int lo = static_cast<int>(static_cast<long long>(n_items) * part / kParts);
    const int hi = static_cast<int>(static_cast<long long>(n_items) * (part + 1) / kParts);
    const int i = lo + static_cast<int>(blockIdx.x) / kParts * kDfsThreads + t;
    int qi = -1, C = 0, B = 0, room = 0;
    int4 *r = nullptr;
    if (i < hi) {
        if (topo.itab) {  // the search's item entry: no offset or header loads first
            const int4 e = topo.itab[i];
            if (e.w >= 0) {
                qi = e.x;
                r = topo.rec + (static_cast<long long>(static_cast<unsigned>(e.y)) |
                                (static_cast<long long>(e.z) << 32));
                C = e.w & 0xff;
                B = (e.w >> 8) & 0xff;
                room = (e.w >> 16) & 0xff;
            }
        } else {
            qi = items ? items[i] : i;
            const long long roff = topo.rec_off[qi];
            if (roff >= 0) {
                r = topo.rec + roff;
                const int4 hd = r[0];
                if (hd.w == 1) {  // else a coincident or empty row: complete
                    C = hd.x;
                    B = hd.y;
                    room = hd.z;
                } else {
                    r = nullptr;
                }
            }
        }
    }
    int total;
    const int off = wave_excl_scan(r ? C : 0, total);  // every lane of the wave takes part
    if (!r) return;
    int *xadj = rec_xadj(r, C, B, room);
    const bool in_pool = off + C <= pool_words;
    if (in_pool) {
        const int4 *xa4 = reinterpret_cast<const int4 *>(xadj);
        for (int d0 = 0; d0 < C; d0 += 4) {
            const int4 v = xa4[d0 >> 2];
            pool[off + d0] = static_cast<unsigned>(v.x);
            if (d0 + 1 < C) pool[off + d0 + 1] = static_cast<unsigned>(v.y);
            if (d0 + 2 < C) pool[off + d0 + 2] = static_cast<unsigned>(v.z);
            if (d0 + 3 < C) pool[off + d0 + 3] = static_cast<unsigned>(v.w);
        }
    }
    // The depth-first parent of a tet is one of its neighbours: once the tet
    // is appended, the byte of its adjacency word that holds the parent gets
    // bit 6 (a dense index is below 64 and 0xff marks a boundary face, so
    // the marked byte is neither; the child scan skips it as >= C)
    auto parent_of = [&](unsigned a) {
        int p = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned b = (a >> (8 * k)) & 0xffu;
            if ((b & 0xc0u) == 0x40u) p = static_cast<int>(b & 0x3fu);
        }
        return p;
    };
    unsigned *ro = reinterpret_cast<unsigned *>(r + rec_ord_off(C));
    unsigned *rb = reinterpret_cast<unsigned *>(r + rec_bnd_off(C));
    unsigned long long vis = 1ull;
    int n = 0, nb = 0;
    unsigned ow = 0, bw = 0;
    // append dense index c at position n: its order entry, boundary faces, mask
    auto append = [&](bool on, int c, unsigned a) {
        // bytes 0xff (boundary faces): bit 7 of each byte of bm (an exact
        // zero-byte test of ~a, whose other bytes are >= 0xc0)
        const unsigned na = ~a;
        unsigned bm = on ? (na - 0x01010101u) & ~na & 0x80808080u : 0u;
        while (bm) {  // in face order
            const int k = __builtin_ctz(bm) >> 3;
            bm &= bm - 1u;
            if (nb < B) {
                bw |= static_cast<unsigned>((n << 2) | k) << (8 * (nb & 3));
                if ((nb & 3) == 3) {
                    rb[nb >> 2] = bw;
                    bw = 0;
                }
            }
            nb++;
        }
        unsigned fc = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned ns = (a >> (8 * (3 - k))) & 0xffu;
            if (ns != 0xffu && ((vis >> ns) & 1ull)) fc |= 1u << k;
        }
        if (on && n < C) {
            ow |= (static_cast<unsigned>(c) | (fc << 6)) << (16 * (n & 1));
            if (n & 1) {
                ro[n >> 1] = ow;
                ow = 0;
            }
        }
        n += on ? 1 : 0;
    };
    // the walk with the words in the pool or (a query past its end) in the
    // record: two copies of the loop, so the pool's accesses stay LDS
    // accesses (one loop over a pointer to either would make them generic)
    auto walk = [&](auto adj_of, auto set_adj) {
        int u = 0;
        unsigned a = adj_of(0);
        append(true, 0, a);
        // one step per iteration, down to the highest free slot or back up to
        // the parent; a holds u's word with its parent mark
        for (;;) {
            int c = -1;
#pragma unroll
            for (int k = 0; k < 4; k++) {  // the highest free slot wins
                const unsigned ck = (a >> (8 * k)) & 0xffu;
                if (ck < static_cast<unsigned>(C) && !((vis >> ck) & 1ull)) c = static_cast<int>(ck);
            }
            // both kinds of step in one straight path (predicated, not
            // branched: a branch lets the compiler split the up-steps into a
            // loop of their own, which the other lanes of the wave wait out)
            const bool down = c >= 0;
            if (!down && u == 0) break;
            const int nx = down ? c : parent_of(a);
            unsigned an = adj_of(nx);
            append(down, c, an);  // the mask sees the tets appended before c
            unsigned mark = 0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (down && ((an >> (8 * k)) & 0xffu) == static_cast<unsigned>(u))
                    mark = 0x40u << (8 * k);
            an |= mark;
            if (down) set_adj(c, an);
            vis |= down ? 1ull << c : 0ull;
            u = nx;
            a = an;
        }
    };
    if (in_pool)
        walk([&](int d) { return pool[off + d]; }, [&](int d, unsigned w) { pool[off + d] = w; });
    else
        walk([&](int d) { return static_cast<unsigned>(xadj[d]); },
             [&](int d, unsigned w) { xadj[d] = static_cast<int>(w); });
    if (n != C || nb != B) {
        topo.rec_off[qi] = -1;
        if (topo.itab) topo.itab[i].w = -1;
        ovf_list[atomicAdd(ovf_count, 1u)] = qi;
        return;
    }
    if (n & 1) ro[n >> 1] = ow;
    if (nb & 3) rb[nb >> 2] = bw;
}

// topo_order_kernel, 16 lanes per query, 4 queries per 64-thread block (one
// wave): the neighbours in order of first appearance in the face-vertex
// sequence (least sequence number by atomicMin, then ballot ranks), the
// cavity entries' vertex slots, the header.
constexpr int kOrderLanes = 16;
constexpr int kOrderQueries = 4;

struct alignas(16) OrderMem {
    int4 cv[grp::CMAX];            // vertex slots per position
    int dtet[grp::CMAX];           // tet per position
    int nkey[grp::NT];             // vertex table: keys, -1 = empty
    int nmin[grp::NT];             // first appearance (face-vertex sequence number)
    unsigned char fcm[grp::CMAX];  // face-circumcentre mask | table position << 4, per position
    unsigned char nslot[grp::NT];  // neighbour slot per vertex-table slot
    unsigned char bnd[grp::BMAX];  // boundary faces in sequence order, position << 2 | face
    unsigned char posof[grp::CMAX];  // position of each dense index
};

__device__ __forceinline__ int order_vertex(const OrderMem &g, int e) {
    const int code = g.bnd[e / 3];
    return sel4(g.cv[code >> 2], face_slot(code & 3, e % 3)) & 0x7fffffff;
}

__global__ void __launch_bounds__(64)
topo_order_kernel(DevMesh m, const int *__restrict__ items, int n_items, Topo topo,
                  unsigned int *ovf_count, int *ovf_list) {
    __shared__ OrderMem omem[kOrderQueries];
    const int lane = __lane_id();
    const int gl = lane & (kOrderLanes - 1);
    const int gw = lane >> 4;
    OrderMem &g = omem[gw];
    const int part = static_cast<int>(blockIdx.x) % kParts;
    const int lo = static_cast<int>(static_cast<long long>(n_items) * part / kParts);
    const int hi = static_cast<int>(static_cast<long long>(n_items) * (part + 1) / kParts);
    const int i = lo + static_cast<int>(blockIdx.x) / kParts * kOrderQueries + gw;
    if (i >= hi) return;  // whole groups leave: nothing below spans groups
    int qi, C, B, room;
    int4 *r;
    if (topo.itab) {  // the search's item entry: no offset or header loads first
        const int4 e = topo.itab[i];
        if (e.w < 0) return;
        qi = e.x;
        r = topo.rec + (static_cast<long long>(static_cast<unsigned>(e.y)) |
                        (static_cast<long long>(e.z) << 32));
        C = e.w & 0xff;
        B = (e.w >> 8) & 0xff;
        room = (e.w >> 16) & 0xff;
    } else {
        qi = items ? items[i] : i;
        const long long roff = topo.rec_off[qi];
        if (roff < 0) return;
        r = topo.rec + roff;
        const int4 hd0 = r[0];
        if (hd0.w != 1) return;  // coincident or empty row: complete
        C = hd0.x;
        B = hd0.y;
        room = hd0.z;
    }
    const int4 hd = make_int4(C, B, room, 1);
    const int N = m.n_data;
    const unsigned gmask = 0xffffu;
    const int *xtet = rec_xtet(r, C, B, room);
    const unsigned short *ro = rec_order(r, hd);
    const unsigned char *rb = rec_bnd(r, hd);
    // the cavity tets: the tets by dense index (with their vertex slots, read
    // as soon as the tets arrive) and the cavity order (position -> dense
    // index and face mask) are read side by side, then put together in
    // position order: two dependent loads after the item entry, not three
    static_assert(grp::CMAX <= 4 * kOrderLanes, "four dense indices per lane");
    // four named sets, not arrays: they stay in registers
    int t0 = -1, t1 = -1, t2 = -1, t3 = -1;  // -1: no such dense index
    unsigned dm0 = 0, dm1 = 0, dm2 = 0, dm3 = 0;
    int4 v0 = make_int4(0, 0, 0, 0), v1 = v0, v2 = v0, v3 = v0;
#define NAT_LOAD_TET(K)                                  \
    if (gl + (K) * kOrderLanes < C) {                    \
        const int xt = xtet[gl + (K) * kOrderLanes];     \
        t##K = handoff_tet(m, xt);                       \
        dm##K = handoff_dm(m, xt);                       \
        v##K = m.tet_v[t##K];                            \
    }
    NAT_LOAD_TET(0)
    NAT_LOAD_TET(1)
    NAT_LOAD_TET(2)
    NAT_LOAD_TET(3)
#undef NAT_LOAD_TET
    for (int p = gl; p < C; p += kOrderLanes) {
        const unsigned o = ro[p];  // position p: its dense index and mask
        g.posof[o & 63u] = static_cast<unsigned char>(p);
        g.fcm[p] = static_cast<unsigned char>((o >> 6) & 15u);
    }
    for (int e = gl; e < B; e += kOrderLanes) g.bnd[e] = rb[e];
    for (int h = gl; h < grp::NT; h += kOrderLanes) {
        g.nkey[h] = -1;
        g.nmin[h] = 0x7fffffff;
    }
    group_sync();
#define NAT_PLACE_TET(K)                                                         \
    if (t##K >= 0) {                                                             \
        const int p = g.posof[gl + (K) * kOrderLanes];                           \
        g.dtet[p] = t##K;                                                        \
        g.cv[p] = v##K;                                                          \
        const unsigned fm = g.fcm[p];                                            \
        g.fcm[p] = static_cast<unsigned char>(fm | (table_pos(fm, dm##K) << 4)); \
    }
    NAT_PLACE_TET(0)
    NAT_PLACE_TET(1)
    NAT_PLACE_TET(2)
    NAT_PLACE_TET(3)
#undef NAT_PLACE_TET
    group_sync();
    // first appearance of each data vertex in the face-vertex sequence
    const int nent = 3 * B;
    bool full = false;
    for (int e = gl; e < nent; e += kOrderLanes) {
        const int v = order_vertex(g, e);
        if (v >= N) continue;
        unsigned h = hash_key(v, grp::NT_LOG2);
        int k = 0;
        for (; k < grp::NT; k++) {
            const int old = atomicCAS(&g.nkey[h], -1, v);
            if (old == -1 || old == v) break;
            h = (h + 1) & (grp::NT - 1);
        }
        if (k == grp::NT) {
            full = true;
            continue;
        }
        atomicMin(&g.nmin[h], e);
    }
    bool ok = !((__ballot(full) >> (16 * gw)) & gmask);
    group_sync();
    // slots in order of first appearance
    int nnb = 0;
    int *rn = reinterpret_cast<int *>(r + rec_nbv_off(C, B));
    for (int e0 = 0; ok && e0 < nent; e0 += kOrderLanes) {
        const int e = e0 + gl;
        bool first = false;
        int v = -1;
        unsigned h = 0;
        if (e < nent) {
            v = order_vertex(g, e);
            if (v < N) {
                h = hash_key(v, grp::NT_LOG2);
                while (g.nkey[h] != v) h = (h + 1) & (grp::NT - 1);
                first = g.nmin[h] == e;
            }
        }
        const unsigned mine = static_cast<unsigned>(__ballot(first) >> (16 * gw)) & gmask;
        if (first) {
            const int slot = nnb + __popc(mine & ((1u << gl) - 1u));
            g.nslot[h] = static_cast<unsigned char>(min(slot, 255));
            if (slot < room) rn[slot] = v;
        }
        nnb += __popc(mine);
    }
    ok = ok && nnb <= room;
    group_sync();
    if (!ok) {
        if (gl == 0) {
            topo.rec_off[qi] = -1;
            ovf_list[atomicAdd(ovf_count, 1u)] = qi;
        }
        return;
    }
    if (nnb == 0) {  // no data vertex on the boundary: an empty row
        if (gl == 0) r[0] = make_int4(0, 0, 0, 0);
        return;
    }
    // the cavity entries in cavity order: tet, vertex slots, face-circumcentre mask
    for (int d = gl; d < C; d += kOrderLanes) {
        const int4 tv = g.cv[d];
        int sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int v = sel4(tv, k) & 0x7fffffff;
            sl[k] = 0xffff;
            if (v < N) {
                unsigned h = hash_key(v, grp::NT_LOG2);
                int key;
                while ((key = g.nkey[h]) != -1 && key != v) h = (h + 1) & (grp::NT - 1);
                if (key == v) sl[k] = g.nslot[h];
            }
        }
        r[1 + d] = make_int4(g.dtet[d], sl[0] | (sl[1] << 16), sl[2] | (sl[3] << 16),
                             static_cast<int>(g.fcm[d]));
    }
    if (gl == 0) r[0] = make_int4(C, B, nnb, 0);
}

// ------------------------------------------------------- wave-per-query rows

// Rows too big for the group kernel (hundreds to thousands of cavity tets,
// e.g. queries inside hollow shells): one wave per query.  The search is the
// group kernel's -- every tet the reference pops is tested once, a front at
// a time, across the 64 lanes -- with its tables in per-wave global scratch.
// The reference's LIFO order (natural_insertionfree.c:498-530) is replayed by
// one lane over a compact LDS graph of the in-sphere tets only (out-of-sphere
// tets are leaves of that search, so leaving them out keeps the order).
// Boundary faces and neighbour slots come in the reference's order by ballot
// compaction (natural_insertionfree.c:559-584).  The forward pass computes
// every subcell volume in parallel, groups the terms by neighbour with a
// stable counting sort, and sums each neighbour's terms in the reference's
// order (cavity tets, then boundary faces), so the weights are bit-identical.
// The gradient pass reduces over the wave instead (order-free sums).  A row
// beyond these capacities goes on to the thread-per-query levels.
namespace wv {
constexpr int VT_LOG2 = 13;
constexpr int VT = 1 << VT_LOG2;  // visited-table slots
constexpr int VMAX = 6144;        // visited tets
constexpr int CMAX = 2048;        // cavity tets
constexpr int SMAX = 1984;        // replay stack
constexpr int BMAX = 8192;        // boundary faces
constexpr int NT_LOG2 = 12;
constexpr int NT = 1 << NT_LOG2;  // neighbour-table slots
constexpr int NMAX = 2048;        // neighbours
constexpr int IMAX = 4 * CMAX + 3 * BMAX;  // subcell-volume terms
constexpr unsigned NONE = 0xffffu;
}  // namespace wv

// Per-wave global scratch.  Tables written with atomics are read with
// device-coherent loads; the rest are written and read by plain accesses
// separated by barriers (one wave per block).
struct WaveScr {
    int *vkey;             // [VT] tet, -1 empty
    uint2 *vnb;            // [VT] visited slots of the 4 neighbours, 16 bits each
    unsigned char *vflag;  // [VT] 1 tested, 3 tested and in the circumsphere
    int *vlist;            // [VMAX] visited slots, in test order
    int *vtet;             // [VMAX] their tets
    int *dense;            // [VT] in-sphere index of a visited slot
    int *dslot;            // [CMAX] visited slot of an in-sphere index
    int *dtet;             // [CMAX] tet of an in-sphere index
    int *cav;              // [CMAX] in-sphere index at each cavity position
    int *pos;              // [CMAX] cavity position of each in-sphere index
    uint2 *cslot;          // [CMAX] neighbour slots of the cavity tet's 4 vertices
    int *cmask;            // [CMAX] its face-circumcentre mask
    int *bnd;              // [BMAX] cavity position << 2 | face
    int *nkey;             // [NT] vertex, -1 empty
    int *nmin;             // [NT] first appearance (face-vertex sequence number)
    int *nslot;            // [NT] neighbour slot
    int *nbv;              // [NMAX] neighbours in slot order
    int *off;              // [NMAX] first term of each neighbour in val
    int *cnt;              // [NMAX] terms of each neighbour
    int *tpos;             // [IMAX] position of each term in val
    double *val;           // [IMAX] term volumes (step 4 negated) grouped by neighbour,
                           //        each group in sequence order
    double *acc;           // [NT] stolen volumes; <dL/dy, values[j]> in the gradient pass
};

__host__ __device__ constexpr size_t wv_round(size_t b) { return (b + 255) & ~size_t(255); }

__host__ __device__ inline size_t wave_scratch_bytes() {
    using namespace wv;
    return wv_round(4 * VT) + wv_round(8 * VT) + wv_round(VT) + 2 * wv_round(4 * VMAX) +
           wv_round(4 * VT) + 4 * wv_round(4 * CMAX) + wv_round(8 * CMAX) + wv_round(4 * CMAX) +
           wv_round(4 * BMAX) + 3 * wv_round(4 * NT) + 3 * wv_round(4 * NMAX) +
           wv_round(4 * IMAX) + wv_round(8 * IMAX) + wv_round(8 * NT);
}

__device__ inline WaveScr wave_carve(char *p) {
    using namespace wv;
    WaveScr w;
    auto take = [&](size_t b) {
        char *r = p;
        p += wv_round(b);
        return r;
    };
    w.vkey = reinterpret_cast<int *>(take(4 * VT));
    w.vnb = reinterpret_cast<uint2 *>(take(8 * VT));
    w.vflag = reinterpret_cast<unsigned char *>(take(VT));
    w.vlist = reinterpret_cast<int *>(take(4 * VMAX));
    w.vtet = reinterpret_cast<int *>(take(4 * VMAX));
    w.dense = reinterpret_cast<int *>(take(4 * VT));
    w.dslot = reinterpret_cast<int *>(take(4 * CMAX));
    w.dtet = reinterpret_cast<int *>(take(4 * CMAX));
    w.cav = reinterpret_cast<int *>(take(4 * CMAX));
    w.pos = reinterpret_cast<int *>(take(4 * CMAX));
    w.cslot = reinterpret_cast<uint2 *>(take(8 * CMAX));
    w.cmask = reinterpret_cast<int *>(take(4 * CMAX));
    w.bnd = reinterpret_cast<int *>(take(4 * BMAX));
    w.nkey = reinterpret_cast<int *>(take(4 * NT));
    w.nmin = reinterpret_cast<int *>(take(4 * NT));
    w.nslot = reinterpret_cast<int *>(take(4 * NT));
    w.nbv = reinterpret_cast<int *>(take(4 * NMAX));
    w.off = reinterpret_cast<int *>(take(4 * NMAX));
    w.cnt = reinterpret_cast<int *>(take(4 * NMAX));
    w.tpos = reinterpret_cast<int *>(take(4 * IMAX));
    w.val = reinterpret_cast<double *>(take(8 * IMAX));
    w.acc = reinterpret_cast<double *>(take(8 * NT));
    return w;
}

__device__ __forceinline__ int ld_coh(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// all of the wave's memory accesses so far are complete and visible to it
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// 16-bit field k of v (a shift of the 64-bit value: a select between the two
// halves with a run-time k would go through private memory)
__device__ __forceinline__ unsigned u16_of(const uint2 &v, int k) {
    const unsigned long long x = (static_cast<unsigned long long>(v.y) << 32) | v.x;
    return static_cast<unsigned>(x >> (16 * k)) & 0xffffu;
}
__device__ __forceinline__ uint2 pack_u16(unsigned a, unsigned b, unsigned c, unsigned d) {
    return make_uint2(a | (b << 16), c | (d << 16));
}

// visited-table insert of tet t: its slot (isnew when this call put it
// there), -1 when the table is full
__device__ __forceinline__ int wv_probe(int *vkey, int t, bool &isnew) {
    unsigned h = hash_key(t, wv::VT_LOG2);
    for (int n = 0; n < wv::VT; n++) {
        const int old = atomicCAS(&vkey[h], -1, t);
        if (old == -1 || old == t) {
            isnew = old == -1;
            return static_cast<int>(h);
        }
        h = (h + 1) & (wv::VT - 1);
    }
    isnew = false;
    return -1;
}

__device__ __forceinline__ int wv_nb_insert(int *nkey, int v) {
    unsigned h = hash_key(v, wv::NT_LOG2);
    for (int n = 0; n < wv::NT; n++) {
        const int old = atomicCAS(&nkey[h], -1, v);
        if (old == -1 || old == v) return static_cast<int>(h);
        h = (h + 1) & (wv::NT - 1);
    }
    return -1;
}

__device__ __forceinline__ int wv_nb_find(const int *nkey, int v) {
    unsigned h = hash_key(v, wv::NT_LOG2);
    for (int n = 0; n < wv::NT; n++) {
        const int k = ld_coh(&nkey[h]);
        if (k == v) return static_cast<int>(h);
        if (k == -1) return -1;
        h = (h + 1) & (wv::NT - 1);
    }
    return -1;
}

__device__ __forceinline__ double readlane_f64(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// 20 KB: two waves per SIMD
struct WaveLds {
    uint2 adj[wv::CMAX];           // in-sphere index of the 4 neighbours (NONE: outside);
                                   // later the neighbours' fill positions
    unsigned short stk[wv::SMAX];  // replay stack
};
static_assert(sizeof(uint2) * wv::CMAX >= sizeof(int) * wv::NMAX,
              "the fill positions reuse the adjacency");

// Topology of one query (phases B-D): on success returns true with C cavity
// tets (W.dtet[W.cav[ci]], W.cslot, W.cmask), B boundary faces (W.bnd) and NB
// neighbours (W.nbv), all in the reference's order.  ORDERED = false (the
// gradient, whose sums do not depend on order): the cavity in search order,
// and neighbour slots are the neighbour table's own slots (NB = its size).
template <bool ORDERED, bool ROBUST>
__device__ bool wave_topology(const DevMesh &m, const WaveScr &W, WaveLds &L, int st,
                              const d3 &qq, int &C, int &B, int &NB) {
    using namespace wv;
    const int lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int N = m.n_data;
    for (int i = lane; i < VT; i += 64) st_coh(&W.vkey[i], -1);
    for (int i = lane; i < NT; i += 64) {
        st_coh(&W.nkey[i], -1);
        st_coh(&W.nmin[i], 0x7fffffff);
    }
    wave_sync();
    // ---- phase B: fronts of tested tets; the start tet and its neighbours first
    int n0 = 0, s0 = -1;
    if (lane == 0) {
        bool isnew;
        s0 = wv_probe(W.vkey, st, isnew);
        W.vlist[0] = s0;
        W.vtet[0] = st;
        n0 = s0 < 0 ? -1 : 1;
        const int4 tn0 = m.tet_n[st];
        for (int k = 0; k < 4 && n0 > 0; k++) {
            const int nb = sel4(tn0, k);
            if (nb < 0) continue;
            const int s2 = wv_probe(W.vkey, nb, isnew);
            if (s2 < 0) {
                n0 = -1;
                break;
            }
            if (isnew) {
                W.vlist[n0] = s2;
                W.vtet[n0] = nb;
                n0++;
            }
        }
    }
    n0 = __shfl(n0, 0, 64);
    s0 = __shfl(s0, 0, 64);
    if (n0 < 0) return false;
    wave_sync();
    int f0 = 0, f1 = n0;
    bool fail = false;
    while (f0 < f1) {
        int nnext = 0;
        for (int i0 = f0; i0 < f1; i0 += 64) {
            const int i = i0 + lane;
            int nnew = 0;
            int nbt[4] = {-1, -1, -1, -1};                 // neighbour tets
            unsigned ns[4] = {NONE, NONE, NONE, NONE};     // their visited slots
            bool isn[4] = {false, false, false, false};    // inserted by this lane
            bool full = false;
            if (i < f1) {
                const int slot = W.vlist[i];
                const int t = W.vtet[i];
                const SphereTet sp = m.sph[t];
                const int4 tn = sp.n;
                const bool ins = tet_inside<ROBUST>(m, sp, t, qq);
                if (ins) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int nb = sel4(tn, k);
                        if (nb < 0) continue;
                        bool isnew;
                        const int s2 = wv_probe(W.vkey, nb, isnew);
                        if (s2 < 0) {
                            full = true;
                            continue;
                        }
                        ns[k] = static_cast<unsigned>(s2);
                        nbt[k] = nb;
                        isn[k] = isnew;
                        nnew += isnew ? 1 : 0;
                    }
                }
                W.vnb[slot] = pack_u16(ns[0], ns[1], ns[2], ns[3]);
                W.vflag[slot] = ins > 0 ? 3 : 1;
            }
            int total;
            const int pre = wave_excl_scan(nnew, total);
            int p = f1 + nnext + pre;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (!isn[k]) continue;
                if (p < VMAX) {
                    W.vlist[p] = static_cast<int>(ns[k]);
                    W.vtet[p] = nbt[k];
                } else {
                    full = true;
                }
                p++;
            }
            if (__ballot(full)) fail = true;
            nnext += total;
        }
        wave_sync();
        if (fail || f1 + nnext > VMAX) return false;
        f0 = f1;
        f1 += nnext;
    }
    const int V = f1;
    // ---- in-sphere tets: dense indices and their adjacency
    if (W.vflag[s0] != 3) return false;  // the start tet tested outside
    int D = 0;
    for (int i0 = 0; i0 < V; i0 += 64) {
        const int i = i0 + lane;
        int slot = -1;
        bool in = false;
        if (i < V) {
            slot = W.vlist[i];
            in = W.vflag[slot] == 3;
        }
        const unsigned long long mk = __ballot(in);
        if (in) {
            const int d = D + __popcll(mk & lt);
            if (d < CMAX) {
                W.dense[slot] = d;
                W.dslot[d] = slot;
                W.dtet[d] = W.vtet[i];
            }
        }
        D += __popcll(mk);
    }
    if (D > CMAX) return false;
    wave_sync();
    for (int d = lane; d < D; d += 64) {
        const uint2 nb = W.vnb[W.dslot[d]];
        unsigned a[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned ns = u16_of(nb, k);
            a[k] = (ns != NONE && W.vflag[ns] == 3) ? static_cast<unsigned>(W.dense[ns]) : NONE;
        }
        L.adj[d] = pack_u16(a[0], a[1], a[2], a[3]);
    }
    const int d0 = __builtin_amdgcn_readfirstlane(W.dense[s0]);
    __syncthreads();
    if constexpr (!ORDERED) {
        for (int d = lane; d < D; d += 64) {
            W.cav[d] = d;
            W.pos[d] = d;
        }
        (void)d0;
    } else {
    // ---- phase C: the reference's cavity order, run by the whole wave on
    // uniform values.  Popped bits live in registers (lane l holds indices
    // 32 l .. 32 l + 31).  Of the children a tet pushes, the last is popped
    // next, so it is taken directly instead of through the stack.
    unsigned popped = 0u;
    int nc = 0, top = 0, d = d0;
    bool ok = true;
    for (;;) {
        const unsigned w = __builtin_amdgcn_readlane(popped, d >> 5);
        if (!((w >> (d & 31)) & 1u)) {
            if (lane == (d >> 5)) popped |= 1u << (d & 31);
            if (lane == 0) {
                W.cav[nc] = d;
                W.pos[d] = nc;
            }
            nc++;
            const uint2 a = L.adj[d];
            const unsigned ax = __builtin_amdgcn_readfirstlane(a.x);
            const unsigned ay = __builtin_amdgcn_readfirstlane(a.y);
            int last = -1;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const unsigned n = ((k < 2 ? ax : ay) >> (16 * (k & 1))) & 0xffffu;
                if (n == NONE) continue;
                const unsigned wn = __builtin_amdgcn_readlane(popped, static_cast<int>(n >> 5));
                if ((wn >> (n & 31)) & 1u) continue;
                if (last >= 0) {
                    if (top >= SMAX) {
                        ok = false;
                        break;
                    }
                    if (lane == 0) L.stk[top] = static_cast<unsigned short>(last);
                    top++;
                }
                last = static_cast<int>(n);
            }
            if (!ok) break;
            if (last >= 0) {
                d = last;
                continue;
            }
        }
        if (top == 0) break;
        top--;
        d = __builtin_amdgcn_readfirstlane(static_cast<int>(L.stk[top]));
    }
    __syncthreads();
    // every in-sphere tet is reachable through in-sphere tets: nc < D only
    // after a stack overflow
    if (!ok || nc != D) return false;
    }
    C = D;
    // ---- phase D: boundary faces in (cavity position, face) order
    int nb_ = 0;
    for (int j0 = 0; j0 < 4 * C; j0 += 64) {
        const int j = j0 + lane;
        bool isb = false;
        if (j < 4 * C) isb = u16_of(L.adj[W.cav[j >> 2]], j & 3) == NONE;
        const unsigned long long mk = __ballot(isb);
        if (isb) {
            const int p = nb_ + __popcll(mk & lt);
            if (p < BMAX) W.bnd[p] = j;
        }
        nb_ += __popcll(mk);
    }
    if (nb_ > BMAX) return false;
    B = nb_;
    wave_sync();
    // neighbours: least face-vertex sequence number of each data vertex
    const int nent = 3 * B;
    auto face_v = [&](int e) {
        const int code = W.bnd[e / 3];
        bool neg;
        const int4 tv = ld_tet_v(m, W.dtet[W.cav[code >> 2]], neg);
        return sel4(tv, face_slot(code & 3, e % 3));
    };
    bool full = false;
    for (int e = lane; e < nent; e += 64) {
        const int v = face_v(e);
        if (v >= N) continue;
        const int h = wv_nb_insert(W.nkey, v);
        if (h < 0) {
            full = true;
            continue;
        }
        if constexpr (ORDERED) atomicMin(&W.nmin[h], e);
    }
    if (__ballot(full)) return false;
    wave_sync();
    if constexpr (!ORDERED) {
        NB = NT;
    } else {
    // slots in order of first appearance
    int nn = 0;
    for (int e0 = 0; e0 < nent; e0 += 64) {
        const int e = e0 + lane;
        bool first = false;
        int v = -1, h = -1;
        if (e < nent) {
            v = face_v(e);
            if (v < N) {
                h = wv_nb_find(W.nkey, v);
                first = h >= 0 && ld_coh(&W.nmin[h]) == e;
            }
        }
        const unsigned long long mk = __ballot(first);
        if (first) {
            const int slot = nn + __popcll(mk & lt);
            W.nslot[h] = slot;
            if (slot < NMAX) W.nbv[slot] = v;
        }
        nn += __popcll(mk);
    }
    if (nn > NMAX || nn == 0) return false;
    NB = nn;
    wave_sync();
    }
    // per cavity tet: neighbour slots of its vertices and the circumcentre mask
    // (face opposite p[k] held by an earlier cavity tet)
    for (int ci = lane; ci < C; ci += 64) {
        const int d = W.cav[ci];
        bool neg;
        const int4 tv = ld_tet_v(m, W.dtet[d], neg);
        unsigned sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int v = sel4(tv, k);
            const int h = v < N ? wv_nb_find(W.nkey, v) : -1;
            sl[k] = h < 0 ? NONE : ORDERED ? static_cast<unsigned>(W.nslot[h])
                                           : static_cast<unsigned>(h);
        }
        W.cslot[ci] = pack_u16(sl[0], sl[1], sl[2], sl[3]);
        const uint2 a = L.adj[d];
        int mask = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned n = u16_of(a, 3 - k);
            if (n != NONE && W.pos[n] < ci) mask |= 1 << k;
        }
        W.cmask[ci] = mask;
    }
    wave_sync();
    return true;
}

// Big record of a wave-per-query row (header .w = 2), kept only when the
// forward's topology is saved: the backward's gradient then reads the cavity
// instead of searching again.  In 16-byte units: [0] (C, B, NB, 2); [1..C]
// the cavity entries as in a level-0 record (tet, vertex slots, face mask);
// ceil(B/8) units of boundary faces, 16 bits each (cavity position << 2 |
// face); ceil(NB/4) units of neighbour vertex indices in slot order.
__device__ __forceinline__ int big_units(int C, int B, int NB) {
    return 1 + C + (B + 7) / 8 + (NB + 3) / 4;
}
__device__ __forceinline__ const unsigned short *big_bnd(const int4 *r, int C) {
    return reinterpret_cast<const unsigned short *>(r + 1 + C);
}
__device__ __forceinline__ const int *big_nbv(const int4 *r, int C, int B) {
    return reinterpret_cast<const int *>(r + 1 + C + (B + 7) / 8);
}

// Forward (GRAD = false): the row into the pool, bit-identical to the
// reference.  Gradient: dL/dq from the same cavity (rows left empty by the
// reference get 0).  Queries that exceed a capacity go to ovf_list.
template <bool GRAD, bool ROBUST>
__global__ void __launch_bounds__(64) wave_rows_kernel(DevMesh m, char *wscr,
                                                       const double *__restrict__ q,
                                                       const int *__restrict__ items, int n_items,
                                                       const unsigned *n_items_dev,
                                                       const int *__restrict__ start_of,
                                                       unsigned int *work, RowOut out, GradIn g,
                                                       int *ovf_list, unsigned int *ovf_count) {
    using namespace wv;
    __shared__ WaveLds L;
    const int lane = __lane_id();
    const WaveScr W = wave_carve(wscr + static_cast<size_t>(blockIdx.x) * wave_scratch_bytes());
    if (n_items_dev) n_items = static_cast<int>(*n_items_dev);
    for (;;) {
        unsigned it = 0;
        if (lane == 0) it = atomicAdd(work, 1u);
        it = __shfl(it, 0, 64);
        if (it >= static_cast<unsigned>(n_items)) break;
        const int qi = items[it];
        const int st = start_of[qi];
        const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                       q[3 * static_cast<size_t>(qi) + 2]};
        if constexpr (GRAD) {
            // a big record of the forward: dL/dq from its boundary faces, as
            // grad_geom_kernel does from a level-0 record, reduced over the wave
            const long long roff = g.rec_off ? g.rec_off[qi] : -1;
            const int4 *r = roff >= 0 ? g.rec + roff : nullptr;
            const int4 hd = r ? r[0] : make_int4(0, 0, 0, 0);
            if (hd.w == 2) {
                const int Cr = hd.x, Br = hd.y, NBr = hd.z;
                const int e0 = g.rowptr[qi];
                if (g.rowptr[qi + 1] - e0 != NBr) {
                    if (lane == 0) store_dq(g, qi, {0.0, 0.0, 0.0});
                    continue;
                }
                const int *rn = big_nbv(r, Cr, Br);
                double abar = 0.0;
                for (int j = lane; j < NBr; j += 64) {
                    const double a = grad_dot(g, qi, rn[j]);
                    W.acc[j] = a;
                    W.val[j] = facet_scale(ld_vert(m, rn[j]), qq);
                    abar += a * g.w[e0 + j];
                }
                wave_sync();
                d3 ag = {0.0, 0.0, 0.0}, tg = {0.0, 0.0, 0.0};
                const unsigned short *bb = big_bnd(r, Cr);
                for (int b = lane; b < Br; b += 64) {
                    const int code = bb[b];
                    const int4 cr = r[1 + (code >> 2)];
                    face_moments_rel(
                        m, cr.x, code & 3, qq, [&](int vi, int) { return rec_slot(cr, vi); },
                        [&](int slot, const d3 &g6) {
                            const double sc = W.val[slot];
                            const d3 gm = {g6.x * sc, g6.y * sc, g6.z * sc};
                            const double a = W.acc[slot];
                            ag = {ag.x + a * gm.x, ag.y + a * gm.y, ag.z + a * gm.z};
                            tg = {tg.x + gm.x, tg.y + gm.y, tg.z + gm.z};
                        });
                }
                abar = wave_sum_f64(abar);
                ag = {wave_sum_f64(ag.x), wave_sum_f64(ag.y), wave_sum_f64(ag.z)};
                tg = {wave_sum_f64(tg.x), wave_sum_f64(tg.y), wave_sum_f64(tg.z)};
                if (lane == 0) {
                    const double inv = 1.0 / g.row_total[qi];
                    store_dq(g, qi, {(ag.x - abar * tg.x) * inv, (ag.y - abar * tg.y) * inv,
                                     (ag.z - abar * tg.z) * inv});
                }
                __syncthreads();  // W.acc and W.val are rewritten by the next row
                continue;
            }
        }
        int C = 0, B = 0, NB = 0;
        if (st < 0 || !wave_topology<!GRAD, ROBUST>(m, W, L, st, qq, C, B, NB)) {
            if (lane == 0) ovf_list[atomicAdd(ovf_count, 1u)] = qi;
            __syncthreads();
            continue;
        }
        auto slot_of_ci = [&](int ci) {
            const uint2 cs = W.cslot[ci];
            return [cs](int vi, int) {
                const unsigned s = u16_of(cs, vi);
                return s == NONE ? -1 : static_cast<int>(s);
            };
        };
        if constexpr (GRAD) {
            // a_j per neighbour-table slot
            for (int h = lane; h < NT; h += 64) {
                const int v = ld_coh(&W.nkey[h]);
                if (v >= 0) W.acc[h] = grad_dot(g, qi, v);
            }
            wave_sync();
            GradAcc A;
            for (int ci = lane; ci < C; ci += 64)
                old_contributions(m, W.dtet[W.cav[ci]], static_cast<unsigned>(W.cmask[ci]),
                                  slot_of_ci(ci),
                                  [&](int slot, double vol) { A.vol(W.acc[slot], vol); });
            for (int b = lane; b < B; b += 64) {
                const int code = W.bnd[b];
                const int ci = code >> 2;
                new_contributions(
                    m, W.dtet[W.cav[ci]], code & 3, qq, slot_of_ci(ci),
                    [&](int slot, double vol) { A.vol(W.acc[slot], -vol); },
                    moment_fn([&](int slot, const d3 &gm) { A.mom(W.acc[slot], gm); }));
            }
            GradAcc T;
            T.s = wave_sum_f64(A.s);
            T.as = wave_sum_f64(A.as);
            T.ag = {wave_sum_f64(A.ag.x), wave_sum_f64(A.ag.y), wave_sum_f64(A.ag.z)};
            T.tg = {wave_sum_f64(A.tg.x), wave_sum_f64(A.tg.y), wave_sum_f64(A.tg.z)};
            if (lane == 0) store_dq(g, qi, T.result());
            __syncthreads();
        } else {
            const int I3 = 4 * C;  // step-3 terms first, then 3 per boundary face
            const int nterm = I3 + 3 * B;
            // slot of term i (-1: a vertex without one)
            auto term_slot = [&](int i) {
                unsigned u = NONE;
                if (i < I3) {
                    u = u16_of(W.cslot[i >> 2], i & 3);
                } else if (i < nterm) {
                    const int b = (i - I3) / 3, k = (i - I3) % 3;
                    const int code = W.bnd[b];
                    u = u16_of(W.cslot[code >> 2], face_slot(code & 3, k));
                }
                return u == NONE ? -1 : static_cast<int>(u);
            };
            // terms per neighbour (LDS counters reuse the adjacency)
            int *pos_l = reinterpret_cast<int *>(L.adj);
            for (int s = lane; s < NB; s += 64) pos_l[s] = 0;
            __syncthreads();
            for (int i = lane; i < nterm; i += 64) {
                const int sl = term_slot(i);
                if (sl >= 0) atomicAdd(&pos_l[sl], 1);
            }
            __syncthreads();
            // exclusive scan: first position of each neighbour's terms
            int carry = 0;
            for (int s0 = 0; s0 < NB; s0 += 64) {
                const int s = s0 + lane;
                const int c = s < NB ? pos_l[s] : 0;
                int total;
                const int pre = wave_excl_scan(c, total);
                if (s < NB) {
                    W.cnt[s] = c;
                    W.off[s] = carry + pre;
                    pos_l[s] = carry + pre;
                }
                carry += total;
            }
            __syncthreads();
            // stable placement: chunks of 64 terms in sequence order; the lanes
            // of one neighbour (matched on the 11 bits of its slot) take
            // consecutive positions in lane order
            for (int i0 = 0; i0 < nterm; i0 += 64) {
                const int i = i0 + lane;
                const int sl = i < nterm ? term_slot(i) : -1;
                unsigned long long same = __ballot(sl >= 0);
#pragma unroll
                for (int b = 0; b < 11; b++) {
                    const bool bit = (sl >> b) & 1;
                    const unsigned long long bb = __ballot(bit);
                    same &= bit ? bb : ~bb;
                }
                int base = 0;
                const int leader = sl >= 0 ? static_cast<int>(__ffsll(same)) - 1 : 0;
                if (sl >= 0 && lane == leader) base = atomicAdd(&pos_l[sl], __popcll(same));
                base = __shfl(base, leader, 64);
                if (sl >= 0) W.tpos[i] = base + __popcll(same & ((1ull << lane) - 1ull));
            }
            wave_sync();
            // the terms: step 3 per cavity tet, step 4 per boundary face
            // (negated: stolen -= vol is stolen + -vol), each at its position
            for (int ci = lane; ci < C; ci += 64) {
                const uint2 cs = W.cslot[ci];
                const int ct = W.dtet[W.cav[ci]];
                const unsigned cm = static_cast<unsigned>(W.cmask[ci]);
                auto old_fn = [&](auto... a) {
                    if (m.old_vol)
                        old_contributions_table(m, ct, table_pos(cm, sphere_dm(m.sph[ct])), a...);
                    else
                        old_contributions(m, ct, cm, a...);
                };
                old_fn(slot_of_ci(ci), [&](int slot, double vol) {
                                      const int k = slot == static_cast<int>(u16_of(cs, 0))
                                                        ? 0
                                                        : slot == static_cast<int>(u16_of(cs, 1))
                                                              ? 1
                                                              : slot == static_cast<int>(
                                                                            u16_of(cs, 2))
                                                                    ? 2
                                                                    : 3;
                                      W.val[W.tpos[4 * ci + k]] = vol;
                                  });
            }
            for (int b = lane; b < B; b += 64) {
                const int code = W.bnd[b];
                const int ci = code >> 2, fi = code & 3;
                const uint2 cs = W.cslot[ci];
                new_contributions(m, W.dtet[W.cav[ci]], fi, qq, slot_of_ci(ci),
                                  [&](int slot, double vol) {
                                      const int k =
                                          slot == static_cast<int>(u16_of(cs, face_slot(fi, 0)))
                                              ? 0
                                              : slot == static_cast<int>(
                                                            u16_of(cs, face_slot(fi, 1)))
                                                    ? 1
                                                    : 2;
                                      W.val[W.tpos[I3 + 3 * b + k]] = -vol;
                                  });
            }
            wave_sync();
            // each neighbour's terms in the reference's order
            for (int s = lane; s < NB; s += 64) {
                const double *vp = W.val + W.off[s];
                const int n = W.cnt[s];
                double v = 0.0;
                for (int j = 0; j < n; j++) v += vp[j];
                W.acc[s] = v;
            }
            wave_sync();
            // step 5 (natural_insertionfree.c:721-737): the total in slot order
            double total = 0.0;
            for (int s0 = 0; s0 < NB; s0 += 64) {
                const double x = s0 + lane < NB ? W.acc[s0 + lane] : 0.0;
                const int n = min(64, NB - s0);
                for (int l = 0; l < n; l++) total += readlane_f64(x, l);
            }
            const int cnt = total > 1e-30 ? NB : 0;
            const double inv = cnt ? 1.0 / total : 0.0;
            unsigned long long base = 0;
            if (lane == 0 && cnt > 0)
                base = atomicAdd(out.pool_used, static_cast<unsigned long long>(cnt));
            base = __shfl(base, 0, 64);
            const long long off = static_cast<long long>(base);
            if (off + cnt <= out.pool_cap) {
                for (int s = lane; s < cnt; s += 64) {
                    out.pool_cols[off + s] = W.nbv[s];
                    out.pool_vals[off + s] = W.acc[s] * inv;
                }
            }
            if (lane == 0) {
                out.row_off[qi] = off;
                out.row_cnt[qi] = cnt;
                out.row_total[qi] = total;
            }
            // the big record, when the topology will be saved
            if (out.save.rec && cnt > 0 && off + cnt <= out.pool_cap) {
                const int units = big_units(C, B, NB);
                unsigned long long rb = 0;
                if (lane == 0)
                    rb = atomicAdd(out.save.rec_used, static_cast<unsigned long long>(units));
                rb = __shfl(rb, 0, 64);
                const long long ro = static_cast<long long>(rb);
                if (ro + units <= out.save.rec_cap) {
                    int4 *r = out.save.rec + ro;
                    for (int ci = lane; ci < C; ci += 64) {
                        const uint2 cs = W.cslot[ci];
                        r[1 + ci] = make_int4(W.dtet[W.cav[ci]], static_cast<int>(cs.x),
                                              static_cast<int>(cs.y), W.cmask[ci]);
                    }
                    unsigned short *bb = reinterpret_cast<unsigned short *>(r + 1 + C);
                    for (int b = lane; b < B; b += 64) bb[b] = static_cast<unsigned short>(W.bnd[b]);
                    int *nv = reinterpret_cast<int *>(r + 1 + C + (B + 7) / 8);
                    for (int j = lane; j < NB; j += 64) nv[j] = W.nbv[j];
                    if (lane == 0) {
                        r[0] = make_int4(C, B, NB, 2);
                        out.save.rec_off[qi] = ro;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ----------------------------------------------------------- geometry kernel

// Work order of the geometry kernels: the Morton-ordered items regrouped,
// within windows of kSortWindow consecutive items (a compact region of
// space), by the size of their cavity + boundary, so that the 64 rows of a
// wave take about the same number of loop trips.  Rows without a record sort
// first.
constexpr int kSortWindow = 4096;
constexpr int kSortKeys = 128;

__global__ void __launch_bounds__(256) work_sort_kernel(const int *__restrict__ items, int n,
                                                        Topo topo, int *__restrict__ out) {
    __shared__ unsigned char key[kSortWindow];
    __shared__ int cnt[kSortKeys];
    const int base = blockIdx.x * kSortWindow;
    const int len = min(kSortWindow, n - base);
    for (int i = threadIdx.x; i < kSortKeys; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const int qi = items[base + i];
        const long long roff = topo.rec_off[qi];
        int k = 0;
        if (roff >= 0) {
            const int4 hd = topo.rec[roff];
            k = min(hd.x + hd.y, kSortKeys - 1);
        }
        key[i] = static_cast<unsigned char>(k);
        atomicAdd(&cnt[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the counts, one wave
        const int lane = threadIdx.x;
        int a = cnt[2 * lane], b = cnt[2 * lane + 1];
        int total;
        const int pre = wave_excl_scan(a + b, total);
        cnt[2 * lane] = pre;
        cnt[2 * lane + 1] = pre + a;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const int pos = atomicAdd(&cnt[key[i]], 1);
        out[base + pos] = items[base + i];
    }
}


// Steps 3-5 from the topology record, one thread per query; the row's
// per-neighbour accumulators live in LDS (slot j of thread t at
// [j * kGeomThreads + t]: the lanes of a wave always hit distinct banks).
// MIXED: steps 3-4 in FP32 where the condition estimate (>= tau) allows.
// A row whose subcell volumes cancel, sum |vol| > kappa |sum vol|, would
// carry the FP32 rounding of its largest terms into its weights: it is
// recomputed in FP64.  Counts of FP32 and FP64 tets and faces and of
// recomputed rows are added to mix[0..2].
template <bool MIXED>
__global__ void __launch_bounds__(kGeomThreads) geom_kernel(DevMesh m,
                                                            const double *__restrict__ q,
                                                            const int *__restrict__ items,
                                                            int n_items, unsigned int *work,
                                                            Topo topo, RowOut out, float tau,
                                                            float kappa,
                                                            unsigned long long *mix) {
    __shared__ double acc_lds[kLevel0Row * kGeomThreads];
    double *acc = acc_lds + threadIdx.x;
    unsigned n32 = 0, n64 = 0, nredo = 0;
    int parts_done = 0, first = 0, nbatch = 0;
    while (part_batch(work, n_items, 64, parts_done, first, nbatch)) {
        const int item = first + __lane_id();
        int qi = -1, cnt = 0, nnb = 0;
        long long roff = -1;
        if (__lane_id() < nbatch) {
            qi = items ? items[item] : item;
            roff = topo.rec_off[qi];
        }
        const int4 *r = roff >= 0 ? topo.rec + roff : nullptr;
        if (r && r[0].w == 2) r = nullptr;  // a wave-per-query row's big record
        if (r) {
            const int4 hd = r[0];
            const int ncav = hd.x, nbnd = hd.y;
            nnb = hd.z;
            if (ncav == 0) {
                cnt = nnb;  // coincident vertices (or an empty row)
            } else {
                const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                               q[3 * static_cast<size_t>(qi) + 2]};
                const unsigned char *rf = rec_bnd(r, hd);
                double absum = 0.0;
                auto run = [&](bool fp32) {
                    for (int j = 0; j < nnb; j++) acc[j * kGeomThreads] = 0.0;
                    for (int ci = 0; ci < ncav; ci++) {
                        const int4 cr = r[1 + ci];
                        auto slot_of = [&](int vi, int) { return rec_slot(cr, vi); };
                        auto add = [&](int slot, double vol) {
                            acc[slot * kGeomThreads] += vol;
                            if constexpr (MIXED) absum += fabs(vol);
                        };
                        if (m.old_vol) {  // exact FP64 values in both modes
                            old_contributions_table(m, cr.x, static_cast<unsigned>(cr.w) >> 4,
                                                    slot_of, add);
                            continue;
                        }
                        if constexpr (MIXED) {
                            if (fp32) {
                                if (old_contributions32(m, cr.x, qq, tau, slot_of, add)) {
                                    n32++;
                                    continue;
                                }
                                n64++;
                            }
                        }
                        old_contributions(m, cr.x, static_cast<unsigned>(cr.w) & 15u, slot_of,
                                          add);
                    }
                    for (int bi = 0; bi < nbnd; bi++) {
                        const int code = rf[bi];
                        const int4 cr = r[1 + (code >> 2)];
                        auto slot_of = [&](int vi, int) { return rec_slot(cr, vi); };
                        auto sub = [&](int slot, double vol) {
                            acc[slot * kGeomThreads] -= vol;
                            if constexpr (MIXED) absum += fabs(vol);
                        };
                        if constexpr (MIXED) {
                            if (fp32) {
                                if (new_contributions32(m, cr.x, code & 3, qq, tau, slot_of,
                                                        sub)) {
                                    n32++;
                                    continue;
                                }
                                n64++;
                            }
                        }
                        new_contributions(m, cr.x, code & 3, qq, slot_of, sub);
                    }
                };
                run(MIXED);
                double total = 0.0;
                for (int j = 0; j < nnb; j++) total += acc[j * kGeomThreads];
                if constexpr (MIXED) {
                    if (!(absum <= kappa * fabs(total))) {  // cancelling row: FP64
                        run(false);
                        total = 0.0;
                        for (int j = 0; j < nnb; j++) total += acc[j * kGeomThreads];
                        nredo++;
                    }
                }
                out.row_total[qi] = total;
                if (total > 1e-30) {
                    const double inv = 1.0 / total;
                    for (int j = 0; j < nnb; j++) acc[j * kGeomThreads] *= inv;
                    cnt = nnb;
                } else {
                    cnt = 0;
                }
            }
        }
        const long long off = wave_reserve(out.pool_used, cnt);
        if (!r) continue;  // another level makes this row
        out.row_off[qi] = off;
        out.row_cnt[qi] = cnt;
        if (cnt == 0 || off + cnt > out.pool_cap) continue;
        const int4 hd = r[0];
        if (hd.x == 0) {
            const int4 co = r[1];
            const double w = 1.0 / cnt;
            for (int j = 0; j < cnt; j++) {
                out.pool_cols[off + j] = sel4(co, j);
                out.pool_vals[off + j] = w;
            }
        } else {
            const int *rn = rec_nbv(r, hd);
            for (int j = 0; j < cnt; j++) {
                out.pool_cols[off + j] = rn[j];
                out.pool_vals[off + j] = acc[j * kGeomThreads];
            }
        }
    }
    if constexpr (MIXED) {
        int t32, t64, tre;
        wave_excl_scan(static_cast<int>(n32), t32);
        wave_excl_scan(static_cast<int>(n64), t64);
        wave_excl_scan(static_cast<int>(nredo), tre);
        if (__lane_id() == 0) {
            atomicAdd(mix, static_cast<unsigned long long>(t32));
            atomicAdd(mix + 1, static_cast<unsigned long long>(t64));
            atomicAdd(mix + 2, static_cast<unsigned long long>(tre));
        }
    }
}

// The same rows (step-3 table present; MIXED as above, each lane deciding
// FP32 or FP64 for its own boundary faces, the lanes' sum |vol| shares added
// across the row before the cancellation test) with G lanes per row and
// 64 / G rows per wave.  At G = 2 a CU holds 512 rows in flight (4 waves per SIMD at
// 128 VGPRs) against 640 of one thread per row, and a wave works on 32
// consecutive rows of the Morton order: the mesh data of the rows in flight
// fits L2 better, and each row's tet and vertex loads are spread over two
// lanes.  Every
// accumulator sees the additions of the loop above in the same order:
//   step 3  the cavity tets in order; lane g adds the table values of
//           vertices g, g + G, ... (the vertices of a tet have distinct
//           slots);
//   step 4  the lanes compute G boundary faces at a time, then subtract
//           their contributions face by face in sequence order (the three
//           slots of a face are distinct);
//   step 5  one lane sums the accumulators in slot order.
// The lanes of a row share a wave and its LDS operations execute in issue
// order; each lane's accesses to acc[slot][row] take data-dependent slots, so
// the compiler keeps them in program order.  Only the face-by-face
// subtraction needs a compiler barrier (per lane, "if (g == t)" in a loop
// over t is one subtraction that could be hoisted out of the loop); no
// memory fence is placed, so the table and vertex loads of later tets and
// faces can be issued early.
constexpr int kGeomRowsThreads = 256;
constexpr int kGeomRowsBatch = 64;
constexpr int kGeomRowLanes = 2;  // 14.9 ms at 1M keys / 10M queries; 4 lanes 15.1, 8 lanes 18.8

template <int G, bool MIXED>
__global__ void __launch_bounds__(kGeomRowsThreads) geom_rows_kernel(
    DevMesh m, const double *__restrict__ q, const int *__restrict__ items, int n_items,
    unsigned int *work, Topo topo, RowOut out, float tau, float kappa,
    unsigned long long *mix) {
    constexpr int RPW = 64 / G;  // rows per wave
    unsigned n32 = 0, n64 = 0, nredo = 0;
    __shared__ double acc[kLevel0Row][kGeomRowsThreads / G];
    const int lane = __lane_id();
    const int g = lane & (G - 1);
    const int lead = lane & ~(G - 1);
    const int rr = static_cast<int>(threadIdx.x) / G;
    int parts_done = 0, first = 0, nbatch = 0, sub = kGeomRowsBatch;
    for (;;) {
        if (sub >= nbatch) {
            if (!part_batch(work, n_items, kGeomRowsBatch, parts_done, first, nbatch)) break;
            sub = 0;
        }
        const int item = sub + lane / G;
        sub += RPW;
        int qi = -1, cnt = 0;
        long long roff = -1;
        if (item < nbatch) {
            qi = items ? items[first + item] : first + item;
            roff = topo.rec_off[qi];
        }
        const int4 *r = roff >= 0 ? topo.rec + roff : nullptr;
        if (r && r[0].w == 2) r = nullptr;  // a wave-per-query row's big record
        if (r) {
            const int4 hd = r[0];
            const int ncav = hd.x, nbnd = hd.y, nnb = hd.z;
            if (ncav == 0) {
                cnt = nnb;  // coincident vertices (or an empty row)
            } else {
                const d3 qq = {q[3 * static_cast<size_t>(qi)],
                               q[3 * static_cast<size_t>(qi) + 1],
                               q[3 * static_cast<size_t>(qi) + 2]};
                const unsigned char *rf = rec_bnd(r, hd);
                double absum = 0.0;  // this lane's share of sum |vol| (MIXED)
                auto run = [&](bool fp32) {
                    for (int j = g; j < nnb; j += G) acc[j][rr] = 0.0;
                    // step 3 (the exact table in both modes), cavity order
                    if (g < 4) {
                        for (int ci = 0; ci < ncav; ci++) {
                            const int4 cr = r[1 + ci];
#pragma unroll
                            for (int vi = g; vi < 4; vi += G) {
                                const double v = m.old_vol[old_vol_index(
                                    cr.x, static_cast<unsigned>(cr.w) >> 4, vi)];
                                const int slot = rec_slot(cr, vi);
                                if (slot >= 0) {
                                    acc[slot][rr] += v;
                                    if constexpr (MIXED) absum += fabs(v);
                                }
                            }
                        }
                    }
                    // step 4
                    for (int c0 = 0; c0 < nbnd; c0 += G) {
                        const int bi = c0 + g;
                        // up to three (slot, volume) pairs, in call order;
                        // filled with value selects (stores merged through a
                        // selected address would put the pairs in scratch)
                        int s0 = -1, s1 = -1, s2 = -1;
                        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
                        if (bi < nbnd) {
                            const int code = rf[bi];
                            const int4 cr = r[1 + (code >> 2)];
                            auto slot_of = [&](int vi, int) { return rec_slot(cr, vi); };
                            auto take = [&](int slot, double vol) {
                                const bool f0 = s0 < 0;
                                const bool f1 = !f0 && s1 < 0;
                                const bool f2 = !f0 && !f1;
                                s0 = f0 ? slot : s0;
                                v0 = f0 ? vol : v0;
                                s1 = f1 ? slot : s1;
                                v1 = f1 ? vol : v1;
                                s2 = f2 ? slot : s2;
                                v2 = f2 ? vol : v2;
                            };
                            bool done = false;
                            if constexpr (MIXED) {
                                if (fp32) {
                                    done = new_contributions32(m, cr.x, code & 3, qq, tau,
                                                               slot_of, take);
                                    if (done)
                                        n32++;
                                    else
                                        n64++;
                                }
                            }
                            if (!done) new_contributions(m, cr.x, code & 3, qq, slot_of, take);
                            if constexpr (MIXED) absum += fabs(v0) + fabs(v1) + fabs(v2);
                        }
                        const int nf = min(G, nbnd - c0);
                        for (int t = 0; t < nf; t++) {
                            if (g == t) {
                                if (s0 >= 0) acc[s0][rr] -= v0;
                                if (s1 >= 0) acc[s1][rr] -= v1;
                                if (s2 >= 0) acc[s2][rr] -= v2;
                            }
                            // keeps face t's subtractions between those of
                            // faces t - 1 and t + 1 (no instruction is emitted)
                            asm volatile("" ::: "memory");
                        }
                    }
                };
                auto row_total = [&]() {
                    double t = 0.0;
                    if (g == 0)
                        for (int j = 0; j < nnb; j++) t += acc[j][rr];
                    return __shfl(t, lead, 64);
                };
                run(MIXED);
                double total = row_total();
                if constexpr (MIXED) {
                    // a row whose volumes cancel carries the FP32 rounding of
                    // its largest terms into its weights: redo it in FP64
                    double ab = absum;
#pragma unroll
                    for (int d = G / 2; d >= 1; d >>= 1) ab += __shfl_xor(ab, d, 64);
                    if (!(ab <= kappa * fabs(total))) {
                        run(false);
                        total = row_total();
                        if (g == 0) nredo++;
                    }
                }
                if (g == 0) out.row_total[qi] = total;
                if (total > 1e-30) {
                    const double inv = 1.0 / total;
                    for (int j = g; j < nnb; j += G) acc[j][rr] *= inv;
                    cnt = nnb;
                }
            }
        }
        const long long off_w = wave_reserve(out.pool_used, g == 0 ? cnt : 0);
        const long long off = __shfl(off_w, lead, 64);
        if (!r) continue;  // another level makes this row
        if (g == 0) {
            out.row_off[qi] = off;
            out.row_cnt[qi] = cnt;
        }
        if (cnt == 0 || off + cnt > out.pool_cap) continue;
        const int4 hd = r[0];
        if (hd.x == 0) {
            const int4 co = r[1];
            const double w = 1.0 / cnt;
            for (int j = g; j < cnt; j += G) {
                out.pool_cols[off + j] = sel4(co, j);
                out.pool_vals[off + j] = w;
            }
        } else {
            const int *rn = rec_nbv(r, hd);
            for (int j = g; j < cnt; j += G) {
                out.pool_cols[off + j] = rn[j];
                out.pool_vals[off + j] = acc[j][rr];
            }
        }
    }
    if constexpr (MIXED) {
        int t32, t64, tre;
        wave_excl_scan(static_cast<int>(n32), t32);
        wave_excl_scan(static_cast<int>(n64), t64);
        wave_excl_scan(static_cast<int>(nredo), tre);
        if (lane == 0) {
            atomicAdd(mix, static_cast<unsigned long long>(t32));
            atomicAdd(mix + 1, static_cast<unsigned long long>(t64));
            atomicAdd(mix + 2, static_cast<unsigned long long>(tre));
        }
    }
}

// Gradient of the rows that have a topology record (the cached cavity).
// With lambda_j the forward's weights and S its total stolen volume,
// dL/dq = (sum_j a_j G_j - abar sum_j G_j) / S, abar = sum_j a_j lambda_j,
// so only the step-4 circumcentres and facet moments G_j are recomputed: no
// old subcells, no volumes.  Queries without a record (made by a larger
// level in the forward pass) are listed in ovf_list.
//
// 8 lanes per row, 8 rows per wave: the rows in flight on an XCD (and the
// vertex and tet lines they share) stay few enough for its L2, where one
// thread per row kept 8 times as many.  The lanes take the neighbours
// (a_j = <dL/dy, values[j]>, kept in LDS) and the boundary faces in turn;
// the sums are reduced over the 8 lanes (the gradient has no prescribed
// summation order).  At 1M keys / 10M queries: 12.0 ms with 8 lanes, 12.5
// with 4, 13.5 with 16, 17.1 with 32.
constexpr int kGradLanes = 8;
constexpr int kGradThreads = 256;
constexpr int kGradBatch = 32;

__global__ void __launch_bounds__(kGradThreads) grad_geom_kernel(
    DevMesh m, const double *__restrict__ q, const int *__restrict__ items, int n_items,
    unsigned int *work, Topo topo, GradIn g, int *ovf_list, unsigned int *ovf_count) {
    // (a_j - abar) / (6 |p_j - q|^2) per neighbour slot of each row in flight
    __shared__ double c_lds[kGradThreads / kGradLanes][kLevel0Row];
    const int gl = __lane_id() & (kGradLanes - 1);
    const int gw = __lane_id() / kGradLanes;
    double *nb_c = c_lds[threadIdx.x / kGradLanes];
    int parts_done = 0, first = 0, nbatch = 0, sub = kGradBatch;
    for (;;) {
        // a wave claims kGradBatch consecutive items and works through them
        // 64 / kGradLanes at a time: one atomic per claim
        if (sub >= nbatch) {
            if (!part_batch(work, n_items, kGradBatch, parts_done, first, nbatch)) break;
            sub = 0;
        }
        const int item = sub + gw;
        sub += 64 / kGradLanes;
        if (item >= nbatch) continue;
        const int qi = items ? items[first + item] : first + item;
        const long long roff = topo.rec_off[qi];
        if (roff < 0) {
            if (gl == 0) ovf_list[atomicAdd(ovf_count, 1u)] = qi;
            continue;
        }
        const int4 *r = topo.rec + roff;
        const int4 hd = r[0];
        const int ncav = hd.x, nbnd = hd.y, nnb = hd.z;
        if (nnb > kLevel0Row || hd.w == 2) {  // a wave-per-query row: that level
            if (gl == 0) ovf_list[atomicAdd(ovf_count, 1u)] = qi;
            continue;
        }
        const int e0 = g.rowptr[qi];
        // coincident rows (constant weights nearby) and empty rows: dq = 0
        if (ncav == 0 || g.rowptr[qi + 1] - e0 != nnb) {
            if (gl == 0) store_dq(g, qi, {0.0, 0.0, 0.0});
            continue;
        }
        const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                       q[3 * static_cast<size_t>(qi) + 2]};
        // per neighbour c_j = (a_j - abar) / (6 |p_j - q|^2), abar = sum_j a_j
        // lambda_j: then dq S = sum over the faces' vertices of c_j times
        // their unscaled facet moments, one accumulator
        const int *rn = rec_nbv(r, hd);
        double abar = 0.0;
        double a_own[kLevel0Row / kGradLanes], s_own[kLevel0Row / kGradLanes];
#pragma unroll
        for (int k = 0; k < kLevel0Row / kGradLanes; k++) {
            const int j = gl + k * kGradLanes;
            a_own[k] = 0.0;
            s_own[k] = 0.0;
            if (j < nnb) {
                a_own[k] = grad_dot(g, qi, rn[j]);
                s_own[k] = facet_scale(ld_vert(m, rn[j]), qq);
                abar += a_own[k] * g.w[e0 + j];
            }
        }
#pragma unroll
        for (int d = kGradLanes / 2; d >= 1; d >>= 1) abar += __shfl_xor(abar, d, kGradLanes);
#pragma unroll
        for (int k = 0; k < kLevel0Row / kGradLanes; k++) {
            const int j = gl + k * kGradLanes;
            if (j < nnb) nb_c[j] = (a_own[k] - abar) * s_own[k];
        }
        group_sync();
        d3 gq = {0.0, 0.0, 0.0};
        const unsigned char *rf = rec_bnd(r, hd);
        for (int bi = gl; bi < nbnd; bi += kGradLanes) {
            const int code = rf[bi];
            const int4 cr = r[1 + (code >> 2)];
            face_moments_rel(
                m, cr.x, code & 3, qq, [&](int vi, int) { return rec_slot(cr, vi); },
                [&](int slot, const d3 &g6) {
                    const double c = nb_c[slot];
                    gq = {gq.x + c * g6.x, gq.y + c * g6.y, gq.z + c * g6.z};
                });
        }
#pragma unroll
        for (int d = kGradLanes / 2; d >= 1; d >>= 1) {
            gq.x += __shfl_xor(gq.x, d, kGradLanes);
            gq.y += __shfl_xor(gq.y, d, kGradLanes);
            gq.z += __shfl_xor(gq.z, d, kGradLanes);
        }
        if (gl == 0) {
            const double inv = 1.0 / g.row_total[qi];
            store_dq(g, qi, {gq.x * inv, gq.y * inv, gq.z * inv});
        }
        group_sync();  // nb_c is rewritten by the group's next row
    }
}

// Audit of the insphere decisions behind each saved row: the cavity tets
// (decided inside) and the tets across its boundary faces (decided outside)
// are tested again under the static error bound.  counts[0] += decisions,
// counts[1] += decisions whose sign the bound does not certify.
__global__ void __launch_bounds__(256) predicate_audit_kernel(DevMesh m,
                                                              const double *__restrict__ q,
                                                              int n, Topo topo,
                                                              unsigned long long *counts) {
    unsigned tested = 0, uncertain = 0;
    for (int qi = blockIdx.x * blockDim.x + threadIdx.x; qi < n; qi += gridDim.x * blockDim.x) {
        const long long roff = topo.rec_off[qi];
        if (roff < 0) continue;
        const int4 *r = topo.rec + roff;
        const int4 hd = r[0];
        if (hd.w == 2) continue;  // a wave-per-query row's big record: not audited
        const int ncav = hd.x, nbnd = hd.y;
        const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                       q[3 * static_cast<size_t>(qi) + 2]};
        auto check = [&](int t) {
            bool neg;
            const int4 tv = ld_tet_v(m, t, neg);
            const d3 p0 = ld_vert(m, tv.x), p1 = ld_vert(m, tv.y);
            const d3 p2 = ld_vert(m, tv.z), p3 = ld_vert(m, tv.w);
            tested++;
            if (neg ? insphere_uncertain(p1, p0, p2, p3, qq)
                    : insphere_uncertain(p0, p1, p2, p3, qq))
                uncertain++;
        };
        for (int ci = 0; ci < ncav; ci++) check(r[1 + ci].x);
        const unsigned char *rf = rec_bnd(r, hd);
        for (int bi = 0; bi < nbnd; bi++) {
            const int code = rf[bi];
            const int nb = sel4(m.tet_n[r[1 + (code >> 2)].x], code & 3);
            if (nb >= 0) check(nb);
        }
    }
    atomicAdd(counts, static_cast<unsigned long long>(tested));
    atomicAdd(counts + 1, static_cast<unsigned long long>(uncertain));
}

// W^T g: out[cols[e], d] += vals[e] * gy[row, d] (float64 atomics; the
// summation order across rows is not fixed).  One thread per (row, d), rows
// taken in `order` when given.
__global__ void __launch_bounds__(256) spmm_t_kernel(int M, int D, const int *__restrict__ rowptr,
                                                     const int *__restrict__ cols,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ gy,
                                                     const int *__restrict__ order,
                                                     double *__restrict__ out) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(M) * D) return;
    const int i = static_cast<int>(t / D), d = static_cast<int>(t % D);
    const int row = order ? order[i] : i;
    const double gv = gy[static_cast<size_t>(row) * D + d];
    if (gv == 0.0) return;
    for (int e = rowptr[row]; e < rowptr[row + 1]; e++)
        atomicAdd(out + static_cast<size_t>(cols[e]) * D + d, vals[e] * gv);
}

// W^T g with per-block aggregation: a block takes 256 consecutive rows of
// `order` (neighbouring queries, which share most of their neighbours), 16
// lanes per row and 16 rows at a time (a row's entries read by consecutive
// lanes; the entries of one row have distinct columns, so at most the 4 rows
// of a wave meet on a table slot at once), sums their contributions per
// column in an LDS hash table of 2^hlog2 columns x D, and adds each distinct
// column to out once, with float64 atomics.  An entry whose column finds no
// table slot within 8 probes goes to out directly.  The rows per block are
// the power of two (16 to 256) that puts about 4 table sizes of entries in a
// block, from the mean row length (rowptr[M] / M, read by each block): long
// rows -- 459 entries per row on a hollow shell -- would otherwise fill the
// table early and send most entries through failed probes.  With P > 1, block b adds
// into private copy b mod P of the [N, D] output instead (priv_sum_kernel
// then sums the copies): when few columns take many rows' entries -- 20k keys
// under 2M rows of 459 entries -- the float64 atomics of every block would
// otherwise queue on the same addresses.
constexpr int kAggRows = 256;
constexpr int kAggLanes = 16;

__global__ void __launch_bounds__(256) spmm_t_agg_kernel(int M, int D, int hlog2,
                                                         const int *__restrict__ rowptr,
                                                         const int *__restrict__ cols,
                                                         const double *__restrict__ vals,
                                                         const double *__restrict__ gy,
                                                         const int *__restrict__ order,
                                                         double *__restrict__ out_all,
                                                         int P, long long nd) {
    extern __shared__ double agg_lds[];
    double *__restrict__ out = out_all + (P > 1 ? static_cast<long long>(blockIdx.x % P) * nd : 0);
    const int H = 1 << hlog2;
    double *hv = agg_lds;                                // [H * D]
    int *hk = reinterpret_cast<int *>(agg_lds + static_cast<size_t>(H) * D);  // [H]
    // rows of this block: R (a power of two, 16..256) from the mean row length
    const long long nnz = rowptr[M];
    int R = kAggRows;
    while (R > 16 && static_cast<long long>(R) * nnz > 4LL * H * M) R >>= 1;
    const long long base = static_cast<long long>(blockIdx.x) * R;
    if (base >= M) return;
    for (int i = threadIdx.x; i < H; i += blockDim.x) hk[i] = -1;
    for (int i = threadIdx.x; i < H * D; i += blockDim.x) hv[i] = 0.0;
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x) & (kAggLanes - 1);
    for (int k = static_cast<int>(threadIdx.x) / kAggLanes; k < R; k += 256 / kAggLanes) {
        const long long ii = base + k;
        if (ii >= M) break;
        const int row = order ? order[ii] : static_cast<int>(ii);
        if (row < 0 || row >= M) continue;
        const double *g = gy + static_cast<size_t>(row) * D;
        const int e = rowptr[row + 1];
        for (int j = rowptr[row] + lane; j < e; j += kAggLanes) {
            const int c = cols[j];
            const double w = vals[j];
            unsigned h = hash_key(c, hlog2);
            int slot = -1;
            for (int n = 0; n < 8; n++) {
                const int old = atomicCAS(&hk[h], -1, c);
                if (old == -1 || old == c) {
                    slot = static_cast<int>(h);
                    break;
                }
                h = (h + 1) & (H - 1);
            }
            if (slot >= 0) {
                for (int d = 0; d < D; d++) atomicAdd(&hv[slot * D + d], w * g[d]);
            } else {
                for (int d = 0; d < D; d++) atomicAdd(out + static_cast<size_t>(c) * D + d, w * g[d]);
            }
        }
    }
    __syncthreads();
    for (int sl = threadIdx.x; sl < H; sl += blockDim.x) {
        const int c = hk[sl];
        if (c < 0) continue;
        for (int d = 0; d < D; d++)
            atomicAdd(out + static_cast<size_t>(c) * D + d, hv[sl * D + d]);
    }
}

__global__ void __launch_bounds__(256) fill_zero_kernel(double *p, long long n) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

// out[t] = sum of the P private copies (each nd values) at t
__global__ void __launch_bounds__(256) priv_sum_kernel(const double *__restrict__ priv, int P,
                                                       long long nd, double *__restrict__ out) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nd) return;
    double v = 0.0;
    for (int p = 0; p < P; p++) v += priv[static_cast<long long>(p) * nd + t];
    out[t] = v;
}

// ------------------------------------------------------------ host helpers

size_t scratch_bytes_per_thread(const Caps &c) {
    const size_t vis = static_cast<size_t>(1) << c.vis_log2;
    const size_t nbm = static_cast<size_t>(1) << c.nbmap_log2;
    return static_cast<size_t>(c.kd) * (4 + 8) + static_cast<size_t>(c.stack) * 4 +
           vis * (8 + 4) + static_cast<size_t>(c.cav) * 4 + static_cast<size_t>(c.bnd) * 4 +
           static_cast<size_t>(c.nb) * (4 + 8) + nbm * (8 + 4) + 4;
}

void scratch_carve(void *base, const Caps &c, int nthr, Scratch *s) {
    char *p = static_cast<char *>(base);
    const size_t T = static_cast<size_t>(nthr);
    auto take = [&](size_t elems, size_t elt) {
        char *r = p;
        size_t bytes = elems * elt * T;
        bytes = (bytes + 255) & ~static_cast<size_t>(255);
        p += bytes;
        return static_cast<void *>(r);
    };
    const size_t vis = static_cast<size_t>(1) << c.vis_log2;
    const size_t nbm = static_cast<size_t>(1) << c.nbmap_log2;
    s->kd_saved = static_cast<double *>(take(c.kd, 8));
    s->nb_vol = static_cast<double *>(take(c.nb, 8));
    s->vis = static_cast<int2 *>(take(vis, 8));
    s->nbmap = static_cast<int2 *>(take(nbm, 8));
    s->kd_node = static_cast<int *>(take(c.kd, 4));
    s->stack = static_cast<int *>(take(c.stack, 4));
    s->vis_pos = static_cast<int *>(take(vis, 4));
    s->nbslot = static_cast<int *>(take(nbm, 4));
    s->cav = static_cast<int *>(take(c.cav, 4));
    s->bnd = static_cast<int *>(take(c.bnd, 4));
    s->nb_idx = static_cast<int *>(take(c.nb, 4));
    s->epoch = static_cast<int *>(take(1, 4));
    s->nthr = nthr;
}

template <typename K>
static int blocks_per_cu(K kernel) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, 0) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}
int rows_blocks_per_cu() { return blocks_per_cu(sibson_rows_kernel<false>); }
int topo_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, topo_search_kernel<false>, 64, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}
int locate_blocks_per_cu() { return blocks_per_cu(locate_kernel<false>); }
int grad_geom_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grad_geom_kernel, kGradThreads, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

int geom_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, geom_kernel<false>, kGeomThreads, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}
int geom_rows_blocks_per_cu(bool mixed) {
    int n = 0;
    const hipError_t e =
        mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, geom_rows_kernel<kGeomRowLanes, true>, kGeomRowsThreads, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, geom_rows_kernel<kGeomRowLanes, false>, kGeomRowsThreads, 0);
    if (e != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_rows(const DevMesh &m, const Caps &caps, const Scratch &sc, int level,
                       int nblocks, const double *d_q, const int *d_items, int n_items,
                       const unsigned *d_n_items, unsigned int *d_work, const RowOut &out,
                       hipStream_t st) {
    if (m.robust)
        hipLaunchKernelGGL(sibson_rows_kernel<true>, dim3(nblocks), dim3(256), 0, st, m, caps, sc,
                           level, d_q, d_items, n_items, d_n_items, d_work, out);
    else
        hipLaunchKernelGGL(sibson_rows_kernel<false>, dim3(nblocks), dim3(256), 0, st, m, caps,
                           sc, level, d_q, d_items, n_items, d_n_items, d_work, out);
    return hipGetLastError();
}

hipError_t launch_sphere(const DevMesh &m, SphereTet *out, hipStream_t st) {
    if (m.n_tets > 0)
        hipLaunchKernelGGL(sphere_kernel, dim3((m.n_tets + 255) / 256), dim3(256), 0, st, m, out);
    return hipGetLastError();
}

hipError_t launch_old_vol(const DevMesh &m, double *out, SphereTet *sph, hipStream_t st) {
    if (m.n_tets > 0)
        hipLaunchKernelGGL(old_vol_kernel, dim3((m.n_tets + 255) / 256), dim3(256), 0, st, m,
                           out, sph);
    return hipGetLastError();
}

size_t wave_scratch_size() { return wave_scratch_bytes(); }

int wave_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, wave_rows_kernel<false, false>, 64, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_wave_rows(const DevMesh &m, char *wscr, int nblocks, const double *d_q,
                            const int *d_items, int n_items, const unsigned *d_n_items,
                            const int *d_start, unsigned int *d_work, const RowOut &out,
                            int *ovf_list, unsigned int *ovf_count, hipStream_t st) {
    GradIn g{};
    if (m.robust)
        hipLaunchKernelGGL((wave_rows_kernel<false, true>), dim3(nblocks), dim3(64), 0, st, m,
                           wscr, d_q, d_items, n_items, d_n_items, d_start, d_work, out, g,
                           ovf_list, ovf_count);
    else
        hipLaunchKernelGGL((wave_rows_kernel<false, false>), dim3(nblocks), dim3(64), 0, st, m,
                           wscr, d_q, d_items, n_items, d_n_items, d_start, d_work, out, g,
                           ovf_list, ovf_count);
    return hipGetLastError();
}

hipError_t launch_wave_grad(const DevMesh &m, char *wscr, int nblocks, const double *d_q,
                            const int *d_items, int n_items, const int *d_start,
                            unsigned int *d_work, const GradIn &g, int *ovf_list,
                            unsigned int *ovf_count, hipStream_t st) {
    RowOut out{};
    if (m.robust)
        hipLaunchKernelGGL((wave_rows_kernel<true, true>), dim3(nblocks), dim3(64), 0, st, m,
                           wscr, d_q, d_items, n_items, nullptr, d_start, d_work, out, g,
                           ovf_list, ovf_count);
    else
        hipLaunchKernelGGL((wave_rows_kernel<true, false>), dim3(nblocks), dim3(64), 0, st, m,
                           wscr, d_q, d_items, n_items, nullptr, d_start, d_work, out, g,
                           ovf_list, ovf_count);
    return hipGetLastError();
}

hipError_t launch_locate(const DevMesh &m, const Caps &caps, const Scratch &sc, int nblocks,
                         const double *d_q, const int *d_items, int n_items, int *d_start,
                         hipStream_t st) {
    if (m.robust)
        hipLaunchKernelGGL(locate_kernel<true>, dim3(nblocks), dim3(256), 0, st, m, caps, sc, d_q,
                           d_items, n_items, d_start);
    else
        hipLaunchKernelGGL(locate_kernel<false>, dim3(nblocks), dim3(256), 0, st, m, caps, sc,
                           d_q, d_items, n_items, d_start);
    return hipGetLastError();
}

hipError_t launch_topo_search(const DevMesh &m, int nblocks, const double *d_q,
                              const int *d_items, int n_items, const int *d_start,
                              unsigned int *d_work, const Topo &topo, unsigned int *ovf_count,
                              int *ovf_list, hipStream_t st) {
    if (m.robust)
        hipLaunchKernelGGL(topo_search_kernel<true>, dim3(nblocks), dim3(64), 0, st, m, d_q,
                           d_items, n_items, d_start, d_work, topo, ovf_count, ovf_list);
    else
        hipLaunchKernelGGL(topo_search_kernel<false>, dim3(nblocks), dim3(64), 0, st, m, d_q,
                           d_items, n_items, d_start, d_work, topo, ovf_count, ovf_list);
    return hipGetLastError();
}

// Words of the DFS kernel's LDS pool a wave may use: kDfsPool, or less under
// the test hook NATINTERP3D_TEST_DFS_POOL (a non-negative integer, read at
// each launch), so that tests can send the walks to their record-resident
// copy; the rows do not change.
static int dfs_pool_words() {
    const char *s = std::getenv("NATINTERP3D_TEST_DFS_POOL");
    if (!s || !*s) return kDfsPool;
    char *end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (!end || *end != '\0' || v < 0) return kDfsPool;
    return v < kDfsPool ? static_cast<int>(v) : kDfsPool;
}

hipError_t launch_topo_order(const DevMesh &m, const int *d_items, int n_items, const Topo &topo,
                             unsigned int *ovf_count, int *ovf_list, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    // blocks per part: the items of the largest part over the queries per block
    const int per_part = (n_items + kParts - 1) / kParts;
    const int nb1 = kParts * ((per_part + kDfsThreads - 1) / kDfsThreads);
    hipLaunchKernelGGL(topo_dfs_kernel, dim3(nb1), dim3(kDfsThreads), 0, st, d_items, n_items,
                       topo, ovf_count, ovf_list, dfs_pool_words());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nb2 = kParts * ((per_part + kOrderQueries - 1) / kOrderQueries);
    hipLaunchKernelGGL(topo_order_kernel, dim3(nb2), dim3(64), 0, st, m, d_items, n_items, topo,
                       ovf_count, ovf_list);
    return hipGetLastError();
}

hipError_t launch_topo(const DevMesh &m, int nblocks, const double *d_q, const int *d_items,
                       int n_items, const int *d_start, unsigned int *d_work, const Topo &topo,
                       unsigned int *ovf_count, int *ovf_list, hipStream_t st) {
    hipError_t e = launch_topo_search(m, nblocks, d_q, d_items, n_items, d_start, d_work, topo,
                                      ovf_count, ovf_list, st);
    if (e != hipSuccess) return e;
    return launch_topo_order(m, d_items, n_items, topo, ovf_count, ovf_list, st);
}

hipError_t launch_work_sort(const int *d_items, int n, const Topo &topo, int *d_out,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(work_sort_kernel, dim3((n + kSortWindow - 1) / kSortWindow), dim3(256), 0,
                       st, d_items, n, topo, d_out);
    return hipGetLastError();
}

hipError_t launch_geom(const DevMesh &m, int nblocks, int nblocks_rows, const double *d_q,
                       const int *d_items, int n_items, unsigned int *d_work, const Topo &topo,
                       const RowOut &out, float mixed_tau, float mixed_kappa,
                       unsigned long long *mix, hipStream_t st) {
    if (m.old_vol && nblocks_rows > 0) {
        if (mix)
            hipLaunchKernelGGL((geom_rows_kernel<kGeomRowLanes, true>), dim3(nblocks_rows),
                               dim3(kGeomRowsThreads), 0, st, m, d_q, d_items, n_items, d_work,
                               topo, out, mixed_tau, mixed_kappa, mix);
        else
            hipLaunchKernelGGL((geom_rows_kernel<kGeomRowLanes, false>), dim3(nblocks_rows),
                               dim3(kGeomRowsThreads), 0, st, m, d_q, d_items, n_items, d_work,
                               topo, out, 0.0f, 0.0f, nullptr);
        return hipGetLastError();
    }
    if (mix)
        hipLaunchKernelGGL(geom_kernel<true>, dim3(nblocks), dim3(kGeomThreads), 0, st, m, d_q,
                           d_items, n_items, d_work, topo, out, mixed_tau, mixed_kappa, mix);
    else
        hipLaunchKernelGGL(geom_kernel<false>, dim3(nblocks), dim3(kGeomThreads), 0, st, m, d_q,
                           d_items, n_items, d_work, topo, out, 0.0f, 0.0f, mix);
    return hipGetLastError();
}

hipError_t launch_grad_rows(const DevMesh &m, const Caps &caps, const Scratch &sc, int level,
                            int nblocks, const double *d_q, const int *d_items, int n_items,
                            unsigned int *d_work, const GradIn &g, int *ovf_list,
                            unsigned int *ovf_count, hipStream_t st) {
    if (m.robust)
        hipLaunchKernelGGL(sibson_grad_rows_kernel<true>, dim3(nblocks), dim3(256), 0, st, m,
                           caps, sc, level, d_q, d_items, n_items, d_work, g, ovf_list,
                           ovf_count);
    else
        hipLaunchKernelGGL(sibson_grad_rows_kernel<false>, dim3(nblocks), dim3(256), 0, st, m,
                           caps, sc, level, d_q, d_items, n_items, d_work, g, ovf_list,
                           ovf_count);
    return hipGetLastError();
}

hipError_t launch_grad_geom(const DevMesh &m, int nblocks, const double *d_q, const int *d_items,
                            int n_items, unsigned int *d_work, const Topo &topo,
                            const GradIn &g, int *ovf_list, unsigned int *ovf_count,
                            hipStream_t st) {
    hipLaunchKernelGGL(grad_geom_kernel, dim3(nblocks), dim3(kGradThreads), 0, st, m, d_q,
                       d_items, n_items, d_work, topo, g, ovf_list, ovf_count);
    return hipGetLastError();
}

hipError_t launch_predicate_audit(const DevMesh &m, const double *d_q, int n, const Topo &topo,
                                  unsigned long long *d_counts, hipStream_t st) {
    const int blocks = std::max(1, std::min((n + 255) / 256, 4096));
    hipLaunchKernelGGL(predicate_audit_kernel, dim3(blocks), dim3(256), 0, st, m, d_q, n, topo,
                       d_counts);
    return hipGetLastError();
}

hipError_t launch_spmm_t(int M, int N, int D, const int *rowptr, const int *cols,
                         const double *vals, const double *gy, const int *order, double *out,
                         double *priv, int P, hipStream_t st) {
    const long long nz = static_cast<long long>(N) * D;
    const long long nt = static_cast<long long>(M) * D;
    // the largest power-of-two table of at most 2048 columns in 48 KB of LDS
    int hlog2 = 11;
    while (hlog2 > 0 && (static_cast<size_t>(1) << hlog2) * (4 + 8 * static_cast<size_t>(D)) >
                            48 * 1024)
        hlog2--;
    const bool agg = hlog2 >= 7;
    if (P < 2 || !priv || !agg || nt <= 0) P = 1;
    // the accumulators: out itself, or P private copies summed into out at the end
    double *acc = P > 1 ? priv : out;
    const long long nacc = nz * P;
    if (nacc > 0)
        hipLaunchKernelGGL(fill_zero_kernel, dim3(static_cast<unsigned>((nacc + 255) / 256)),
                           dim3(256), 0, st, acc, nacc);
    if (nt <= 0) return hipGetLastError();
    if (agg) {
        const size_t lds = (static_cast<size_t>(1) << hlog2) * (4 + 8 * static_cast<size_t>(D));
        // enough blocks for 16 rows each; blocks beyond the rows leave at once
        hipLaunchKernelGGL(spmm_t_agg_kernel, dim3(static_cast<unsigned>((M + 15) / 16)),
                           dim3(256), lds, st, M, D, hlog2, rowptr, cols, vals, gy, order, acc,
                           P, nz);
        if (P > 1)
            hipLaunchKernelGGL(priv_sum_kernel, dim3(static_cast<unsigned>((nz + 255) / 256)),
                               dim3(256), 0, st, acc, P, nz, out);
    } else {
        hipLaunchKernelGGL(spmm_t_kernel, dim3(static_cast<unsigned>((nt + 255) / 256)),
                           dim3(256), 0, st, M, D, rowptr, cols, vals, gy, order, out);
    }
    return hipGetLastError();
}

}  // namespace nat
