// sibson_kernels.hip -- gfx950 kernels of the Sibson query path.
//
//   morton_*        spatial bucketing of the queries (order only; no result
//                   depends on it)                         natural.c:18-96
//   sibson_rows     one thread per query: k-d nearest centroid, visibility
//                   walk, cavity search, boundary/neighbour order, stolen
//                   volumes, normalisation         natural_insertionfree.c:486-738
//   scan/gather     CSR assembly (count, scan, fill)  natural_insertionfree.c:895-928
//   spmm            W @ values in stored order        natinterp3d.py:131-137
//
// Every FP64 decision and value is bit-identical to the reference: the same
// expressions (geom.h), the same visiting orders, -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "geom.h"
#include "nat_internal.h"

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
__device__ __forceinline__ double seld(const d3 &v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}
__device__ __forceinline__ unsigned hash_tet(int t, int log2) {
    return (static_cast<unsigned>(t) * 2654435761u) >> (32 - log2);
}

// ------------------------------------------------------------- k-d nearest

// kd_nearest3_data / kd_nearest_i (kdtree.c:377-434, 543-564) with an
// explicit stack: frame d holds the node at depth d (split axis d mod 3),
// the phase reached (0 = enter, 1 = nearer subtree done, 2 = farther subtree
// done) and the rectangle bound it saved before slicing.
__device__ int kd_nearest(const DevMesh &m, const d3 &q, const Scratch &sc, int tid, int cap,
                          bool &ovf) {
    if (m.n_tets <= 0) return -1;
    double rmin[3] = {m.rmin[0], m.rmin[1], m.rmin[2]};
    double rmax[3] = {m.rmax[0], m.rmax[1], m.rmax[2]};
    const size_t S = static_cast<size_t>(sc.nthr);
    const double4 r0 = m.kd[0];
    double best_d;
    {
        const double dx = r0.x - q.x, dy = r0.y - q.y, dz = r0.z - q.z;
        best_d = dx * dx + dy * dy + dz * dz;
    }
    int best = 0;
    int sp = 1;
    sc.kd_node[tid] = 0;
    while (sp > 0) {
        const size_t fi = static_cast<size_t>(sp - 1) * S + tid;
        const int fr = sc.kd_node[fi];
        const int node = fr & 0x3fffffff;
        int phase = fr >> 30;
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
                sc.kd_node[static_cast<size_t>(sp) * S + tid] = nearer;
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
        // this node
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
                sc.kd_node[static_cast<size_t>(sp) * S + tid] = farther;
                sp++;
                continue;
            }
            c = saved;
        }
        sp--;
    }
    return best;
}

// findContainingSimplex (delaunay.c:383-407): leave through the first face
// whose plane separates the query, restart the face loop after each move.
__device__ int walk(const DevMesh &m, int s, const d3 &q, bool &ovf) {
    bool neg;
    int4 tv = ld_tet_v(m, s, neg);
    int steps = 0;
    int i = 0;
    while (i < 4) {
        const d3 a = ld_vert(m, sel4(tv, face_slot(i, 0)));
        const d3 b = ld_vert(m, sel4(tv, face_slot(i, 1)));
        const d3 c = ld_vert(m, sel4(tv, face_slot(i, 2)));
        if (orient3d(a, b, c, q) < 0) {
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

// Visited-set probes (set semantics of natural_insertionfree.c:152-211).
struct Visited {
    int2 *tab;
    int *pos;
    size_t S;
    int tid, log2, stamp;
    __device__ size_t at(unsigned h) const { return static_cast<size_t>(h) * S + tid; }
    // returns slot index of t, inserting it when absent; found = already present
    __device__ unsigned insert(int t, bool &found) const {
        const unsigned mask = (1u << log2) - 1;
        unsigned h = hash_tet(t, log2);
        for (;;) {
            const int2 e = tab[at(h)];
            if (e.y != stamp) {
                tab[at(h)] = make_int2(t, stamp);
                pos[at(h)] = -1;
                found = false;
                return h;
            }
            if (e.x == t) {
                found = true;
                return h;
            }
            h = (h + 1) & mask;
        }
    }
    __device__ bool contains(int t) const {
        const unsigned mask = (1u << log2) - 1;
        unsigned h = hash_tet(t, log2);
        for (;;) {
            const int2 e = tab[at(h)];
            if (e.y != stamp) return false;
            if (e.x == t) return true;
            h = (h + 1) & mask;
        }
    }
    // cavity position of t, -1 when t is not in the cavity
    __device__ int cavity_pos(int t) const {
        const unsigned mask = (1u << log2) - 1;
        unsigned h = hash_tet(t, log2);
        for (;;) {
            const int2 e = tab[at(h)];
            if (e.y != stamp) return -1;
            if (e.x == t) return pos[at(h)];
            h = (h + 1) & mask;
        }
    }
};

__device__ __forceinline__ int find_nb(const Scratch &sc, size_t S, int tid, int n, int v) {
    for (int j = 0; j < n; j++)
        if (sc.nb_idx[static_cast<size_t>(j) * S + tid] == v) return j;
    return -1;
}

// Circumcentre of the face of tet `t` opposite its vertex slot k, vertices in
// ascending slot order -- the order faceCCLookup is first called with
// (natural_insertionfree.c:649-652).
__device__ __forceinline__ d3 face_cc_own(const d3 P[4], int k) {
    const int a = k == 0 ? 1 : 0;
    const int b = k <= 1 ? 2 : 1;
    const int c = k <= 2 ? 3 : 2;
    return cc_tri(P[a], P[b], P[c]);
}

// One query.  Returns the row length (the row is left in nb_idx/nb_vol), or
// -1 when a capacity of this level was exceeded.
__device__ int sibson_row(const DevMesh &m, const Caps &caps, const Scratch &sc, int tid,
                          int stamp, const d3 &q) {
    const size_t S = static_cast<size_t>(sc.nthr);
    const int N = m.n_data;
    bool ovf = false;

    int start = kd_nearest(m, q, sc, tid, caps.kd, ovf);
    if (ovf) return -1;
    if (start < 0) return 0;
    start = walk(m, start, q, ovf);
    if (ovf) return -1;
    if (start < 0) return 0;  // outside the super simplex: empty row

    // ---- cavity search, natural_insertionfree.c:498-530
    Visited vis{sc.vis, sc.vis_pos, S, tid, caps.vis_log2, stamp};
    const int vis_max = (3 << caps.vis_log2) / 4;
    int nvis = 0, ncav = 0, top = 0;
    sc.stack[tid] = start;
    top = 1;
    while (top > 0) {
        top--;
        const int cur = sc.stack[static_cast<size_t>(top) * S + tid];
        bool found;
        const unsigned h = vis.insert(cur, found);
        if (found) continue;
        if (++nvis >= vis_max) return -1;
        bool neg;
        const int4 tv = ld_tet_v(m, cur, neg);
        const d3 p0 = ld_vert(m, tv.x), p1 = ld_vert(m, tv.y);
        const d3 p2 = ld_vert(m, tv.z), p3 = ld_vert(m, tv.w);
        const double ins = neg ? insphere(p1, p0, p2, p3, q) : insphere(p0, p1, p2, p3, q);
        if (ins > 0) {
            if (ncav >= caps.cav) return -1;
            sc.cav[static_cast<size_t>(ncav) * S + tid] = cur;
            sc.vis_pos[vis.at(h)] = ncav;
            ncav++;
            const int4 tn = m.tet_n[cur];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int nb = sel4(tn, i);
                if (nb >= 0 && !vis.contains(nb)) {
                    if (top >= caps.stack) return -1;
                    sc.stack[static_cast<size_t>(top) * S + tid] = nb;
                    top++;
                }
            }
        }
    }

    int nnb = 0;
    if (ncav == 0) {
        // coincident-vertex path, natural_insertionfree.c:596-622: the re-walk
        // of the reference ends in the same tet `start`.
        bool neg;
        const int4 tv = ld_tet_v(m, start, neg);
        int coin[4];
        int nco = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int p = sel4(tv, i);
            const d3 v = ld_vert(m, p);
            if (p < N && v.x == q.x && v.y == q.y && v.z == q.z) coin[nco++] = p;
        }
        if (nco > 0) {
            const double w = 1.0 / nco;
            for (int i = 0; i < nco; i++) {
                sc.nb_idx[static_cast<size_t>(i) * S + tid] = coin[i];
                sc.nb_vol[static_cast<size_t>(i) * S + tid] = w;
            }
        }
        return nco;
    }

    // ---- boundary faces and neighbour order, natural_insertionfree.c:559-584
    int nbnd = 0;
    for (int ci = 0; ci < ncav; ci++) {
        const int t = sc.cav[static_cast<size_t>(ci) * S + tid];
        const int4 tn = m.tet_n[t];
        bool neg;
        const int4 tv = ld_tet_v(m, t, neg);
        for (int fi = 0; fi < 4; fi++) {
            const int nb = sel4(tn, fi);
            if (nb >= 0 && vis.cavity_pos(nb) >= 0) continue;
            if (nbnd >= caps.bnd) return -1;
            sc.bnd[static_cast<size_t>(nbnd) * S + tid] = (ci << 2) | fi;
            nbnd++;
            for (int k = 0; k < 3; k++) {
                const int v = sel4(tv, face_slot(fi, k));
                if (v >= N) continue;
                if (find_nb(sc, S, tid, nnb, v) >= 0) continue;
                if (nnb >= caps.nb) return -1;
                sc.nb_idx[static_cast<size_t>(nnb) * S + tid] = v;
                sc.nb_vol[static_cast<size_t>(nnb) * S + tid] = 0.0;
                nnb++;
            }
        }
    }
    if (nnb == 0) return 0;

    // ---- step 3: old contributions, natural_insertionfree.c:642-676
    for (int ci = 0; ci < ncav; ci++) {
        const int t = sc.cav[static_cast<size_t>(ci) * S + tid];
        bool neg;
        const int4 tv = ld_tet_v(m, t, neg);
        const int4 tn = m.tet_n[t];
        d3 P[4];
        P[0] = ld_vert(m, tv.x);
        P[1] = ld_vert(m, tv.y);
        P[2] = ld_vert(m, tv.z);
        P[3] = ld_vert(m, tv.w);
        const d3 ct = cc_tet(P[0], P[1], P[2], P[3]);
        // fcc[k]: circumcentre of the face opposite p[k], evaluated in the
        // vertex order of the first cavity tet (in cavity order) holding it.
        d3 fcc[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int nb = sel4(tn, 3 - k);
            int cj = -1;
            if (nb >= 0) cj = vis.cavity_pos(nb);
            if (cj >= 0 && cj < ci) {
                bool nneg;
                const int4 nv = ld_tet_v(m, nb, nneg);
                // the neighbour's vertex not on this face
                const int fa = sel4(tv, k == 0 ? 1 : 0), fb = sel4(tv, k <= 1 ? 2 : 1),
                          fc = sel4(tv, k <= 2 ? 3 : 2);
                int kk = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int x = sel4(nv, j);
                    if (x != fa && x != fb && x != fc) kk = j;
                }
                const int ia = sel4(nv, kk == 0 ? 1 : 0), ib = sel4(nv, kk <= 1 ? 2 : 1),
                          ic = sel4(nv, kk <= 2 ? 3 : 2);
                fcc[k] = cc_tri(ld_vert(m, ia), ld_vert(m, ib), ld_vert(m, ic));
            } else {
                fcc[k] = face_cc_own(P, k);
            }
        }
        const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int vi = 0; vi < 4; vi++) {
            const int pk = vid[vi];
            if (pk >= N) continue;
            const int slot = find_nb(sc, S, tid, nnb, pk);
            if (slot < 0) continue;
            int o0 = (vi + 1) & 3, o1 = (vi + 2) & 3, o2 = (vi + 3) & 3, tmp;
            if (o2 < o0) { tmp = o0; o0 = o2; o2 = tmp; }
            if (o1 > o2) { tmp = o1; o1 = o2; o2 = tmp; }
            if (o0 > o1) { tmp = o0; o0 = o1; o1 = tmp; }
            const double vol =
                subcell_volume(P[vi], P[o0], P[o1], P[o2], ct, fcc[o2], fcc[o1], fcc[o0]);
            double &acc = sc.nb_vol[static_cast<size_t>(slot) * S + tid];
            acc += vol;
        }
    }

    // ---- step 4: new contributions, natural_insertionfree.c:681-719
    for (int bi = 0; bi < nbnd; bi++) {
        const int code = sc.bnd[static_cast<size_t>(bi) * S + tid];
        const int ci = code >> 2, fi = code & 3;
        const int t = sc.cav[static_cast<size_t>(ci) * S + tid];
        bool neg;
        const int4 tv = ld_tet_v(m, t, neg);
        d3 P[4];
        P[0] = ld_vert(m, tv.x);
        P[1] = ld_vert(m, tv.y);
        P[2] = ld_vert(m, tv.z);
        P[3] = ld_vert(m, tv.w);
        const int s0 = face_slot(fi, 0), s1 = face_slot(fi, 1), s2 = face_slot(fi, 2);
        const int fv[3] = {sel4(tv, s0), sel4(tv, s1), sel4(tv, s2)};
        const d3 F[3] = {s0 == 0 ? P[0] : (s0 == 1 ? P[1] : (s0 == 2 ? P[2] : P[3])),
                         s1 == 0 ? P[0] : (s1 == 1 ? P[1] : (s1 == 2 ? P[2] : P[3])),
                         s2 == 0 ? P[0] : (s2 == 1 ? P[1] : (s2 == 2 ? P[2] : P[3]))};
        const d3 cv = cc_tet(q, F[0], F[1], F[2]);
        const d3 cf = face_cc_own(P, 3 - fi);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int pk = fv[k];
            if (pk >= N) continue;
            const int slot = find_nb(sc, S, tid, nnb, pk);
            if (slot < 0) continue;
            const d3 &A = F[(k + 1) % 3];
            const d3 &B = F[(k + 2) % 3];
            const d3 ca = cc_tri(F[k], q, A);
            const d3 cb = cc_tri(F[k], q, B);
            const double vol = subcell_volume(F[k], q, A, B, cv, ca, cb, cf);
            double &acc = sc.nb_vol[static_cast<size_t>(slot) * S + tid];
            acc -= vol;
        }
    }

    // ---- step 5: normalise, natural_insertionfree.c:721-737
    double total = 0.0;
    for (int j = 0; j < nnb; j++) total += sc.nb_vol[static_cast<size_t>(j) * S + tid];
    if (!(total > 1e-30)) return 0;
    const double inv = 1.0 / total;
    for (int j = 0; j < nnb; j++) sc.nb_vol[static_cast<size_t>(j) * S + tid] *= inv;
    return nnb;
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

// Persistent row kernel: each wave takes 64 consecutive work items at a time
// from a global counter, computes one row per lane, then reserves one
// contiguous pool range for the wave's rows.
__global__ void __launch_bounds__(256) sibson_rows_kernel(DevMesh m, Caps caps, Scratch sc,
                                                          const double *__restrict__ q,
                                                          const int *__restrict__ items,
                                                          int n_items, unsigned int *work,
                                                          RowOut out) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    // nthr is a multiple of 64, so whole waves leave here
    if (tid >= sc.nthr) return;
    const int lane = __lane_id();
    const size_t S = static_cast<size_t>(sc.nthr);
    int epoch = sc.epoch[tid];
    for (;;) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(work, 64u);
        base = __shfl(base, 0, 64);
        if (base >= static_cast<unsigned>(n_items)) break;
        const int item = static_cast<int>(base) + lane;
        int qi = -1;
        int cnt = 0;
        if (item < n_items) {
            qi = items ? items[item] : item;
            epoch++;
            if (epoch == 0x7fffffff) epoch = 1;
            const d3 qq = {q[3 * static_cast<size_t>(qi)], q[3 * static_cast<size_t>(qi) + 1],
                           q[3 * static_cast<size_t>(qi) + 2]};
            cnt = sibson_row(m, caps, sc, tid, epoch, qq);
            if (cnt < 0) {
                const unsigned k = atomicAdd(out.ovf_count, 1u);
                out.ovf_list[k] = qi;
                cnt = 0;
            }
        }
        int total;
        const int pre = wave_excl_scan(cnt, total);
        unsigned long long wbase = 0;
        if (lane == 0 && total > 0)
            wbase = atomicAdd(out.pool_used, static_cast<unsigned long long>(total));
        wbase = __shfl(wbase, 0, 64);
        if (qi >= 0) {
            const long long off = static_cast<long long>(wbase) + pre;
            out.row_off[qi] = off;
            out.row_cnt[qi] = cnt;
            if (off + cnt <= out.pool_cap) {
                for (int j = 0; j < cnt; j++) {
                    out.pool_cols[off + j] = sc.nb_idx[static_cast<size_t>(j) * S + tid];
                    out.pool_vals[off + j] = sc.nb_vol[static_cast<size_t>(j) * S + tid];
                }
            }
        }
    }
    sc.epoch[tid] = epoch;
}

size_t scratch_bytes_per_thread(const Caps &c) {
    return static_cast<size_t>(c.kd) * (4 + 8) + static_cast<size_t>(c.stack) * 4 +
           (static_cast<size_t>(1) << c.vis_log2) * (8 + 4) + static_cast<size_t>(c.cav) * 4 +
           static_cast<size_t>(c.bnd) * 4 + static_cast<size_t>(c.nb) * (4 + 8) + 4;
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
    s->kd_saved = static_cast<double *>(take(c.kd, 8));
    s->nb_vol = static_cast<double *>(take(c.nb, 8));
    s->vis = static_cast<int2 *>(take(static_cast<size_t>(1) << c.vis_log2, 8));
    s->kd_node = static_cast<int *>(take(c.kd, 4));
    s->stack = static_cast<int *>(take(c.stack, 4));
    s->vis_pos = static_cast<int *>(take(static_cast<size_t>(1) << c.vis_log2, 4));
    s->cav = static_cast<int *>(take(c.cav, 4));
    s->bnd = static_cast<int *>(take(c.bnd, 4));
    s->nb_idx = static_cast<int *>(take(c.nb, 4));
    s->epoch = static_cast<int *>(take(1, 4));
    s->nthr = nthr;
}

int rows_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sibson_rows_kernel, 256, 0) !=
        hipSuccess)
        return 1;
    return n;
}

hipError_t launch_rows(const DevMesh &m, const Caps &caps, const Scratch &sc, int nblocks,
                       const double *d_q, const int *d_items, int n_items,
                       unsigned int *d_work, const RowOut &out, hipStream_t st) {
    hipLaunchKernelGGL(sibson_rows_kernel, dim3(nblocks), dim3(256), 0, st, m, caps, sc, d_q,
                       d_items, n_items, d_work, out);
    return hipGetLastError();
}

// --------------------------------------------------------- Morton bucketing

__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000ff;
    v = (v | (v << 8)) & 0x0300f00f;
    v = (v | (v << 4)) & 0x030c30c3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

__global__ void morton_key_kernel(const double *__restrict__ q, int M, double x0, double y0,
                                  double z0, double sx, double sy, double sz, int bits,
                                  int *__restrict__ keys, int *__restrict__ hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int lim = (1 << bits) - 1;
    auto cell = [&](double v, double o, double s) {
        double c = (v - o) * s;
        int k = (c >= 0.0 && c < 2147483647.0) ? static_cast<int>(c) : 0;  // NaN -> 0
        if (c >= 2147483647.0) k = lim;
        return k > lim ? lim : k;
    };
    const unsigned ix = cell(q[3 * static_cast<size_t>(i)], x0, sx);
    const unsigned iy = cell(q[3 * static_cast<size_t>(i) + 1], y0, sy);
    const unsigned iz = cell(q[3 * static_cast<size_t>(i) + 2], z0, sz);
    const int key = static_cast<int>(spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2));
    keys[i] = key;
    atomicAdd(&hist[key], 1);
}

__global__ void morton_scatter_kernel(const int *__restrict__ keys, int M,
                                      const long long *__restrict__ start, int *cursor,
                                      int *__restrict__ order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int key = keys[i];
    const int k = atomicAdd(&cursor[key], 1);
    order[start[key] + k] = i;
}

hipError_t launch_morton_sort(const double *d_q, int M, const double bbox[6], int bits,
                              int *d_order, int *d_hist, int *d_keys, long long *d_scan_tmp,
                              hipStream_t st) {
    const int nb = 1 << (3 * bits);
    hipError_t e = hipMemsetAsync(d_hist, 0, sizeof(int) * nb, st);
    if (e != hipSuccess) return e;
    auto scale = [&](int d) {
        const double r = bbox[3 + d] - bbox[d];
        return r > 0 ? static_cast<double>(1 << bits) / r : 1.0;
    };
    hipLaunchKernelGGL(morton_key_kernel, dim3((M + 255) / 256), dim3(256), 0, st, d_q, M,
                       bbox[0], bbox[1], bbox[2], scale(0), scale(1), scale(2), bits, d_keys,
                       d_hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    long long *starts = d_scan_tmp;  // [nb + 1]
    long long *tmp = d_scan_tmp + nb + 1;
    e = launch_scan_counts(d_hist, starts, nb, tmp, st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(d_hist, 0, sizeof(int) * nb, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(morton_scatter_kernel, dim3((M + 255) / 256), dim3(256), 0, st, d_keys,
                       M, starts, d_hist, d_order);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scanning

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ long long block_excl_scan(long long v, long long &total) {
    __shared__ long long wsum[kScanBlock / 64];
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    long long x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < kScanBlock / 64; j++) {
        const long long s = wsum[j];
        if (j < w) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__global__ void scan_tiles_kernel(const int *__restrict__ in, int n, long long *__restrict__ tsum) {
    const int base = blockIdx.x * kScanTile;
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    long long total;
    block_excl_scan(s, total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ void scan_sums_kernel(long long *tsum, int ntiles, long long *grand) {
    long long carry = 0;
    for (int b = 0; b < ntiles; b += kScanBlock) {
        const int i = b + threadIdx.x;
        const long long v = i < ntiles ? tsum[i] : 0;
        long long total;
        const long long e = block_excl_scan(v, total);
        if (i < ntiles) tsum[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void scan_apply_kernel(const int *__restrict__ in, int n,
                                  const long long *__restrict__ tsum, long long *__restrict__ out) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    long long total;
    long long off = block_excl_scan(s, total) + tsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) out[i] = off;
        off += v[k];
    }
}

size_t scan_tmp_elems(int n) { return static_cast<size_t>((n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_scan_counts(const int *d_in, long long *d_out, int n, long long *d_tmp,
                              hipStream_t st) {
    const int ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles > 0) {
        hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(kScanBlock), 0, st, d_in, n,
                           d_tmp);
    }
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, st, d_tmp, ntiles,
                       d_out + n);
    if (ntiles > 0) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(ntiles), dim3(kScanBlock), 0, st, d_in, n,
                           d_tmp, d_out);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ CSR gather

__global__ void gather_csr_kernel(const long long *__restrict__ rowptr64,
                                  const long long *__restrict__ row_off,
                                  const int *__restrict__ row_cnt,
                                  const double *__restrict__ pv, const int *__restrict__ pc,
                                  int M, double *__restrict__ vals, int *__restrict__ cols,
                                  int *__restrict__ rowptr32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > M) return;
    rowptr32[i] = static_cast<int>(rowptr64[i]);
    if (i == M) return;
    const long long dst = rowptr64[i], src = row_off[i];
    const int c = row_cnt[i];
    for (int j = 0; j < c; j++) {
        vals[dst + j] = pv[src + j];
        cols[dst + j] = pc[src + j];
    }
}

hipError_t launch_gather_csr(const long long *d_rowptr64, const long long *d_row_off,
                             const int *d_row_cnt, const double *pool_vals,
                             const int *pool_cols, int M, double *d_vals, int *d_cols,
                             int *d_rowptr32, hipStream_t st) {
    hipLaunchKernelGGL(gather_csr_kernel, dim3((M + 1 + 255) / 256), dim3(256), 0, st,
                       d_rowptr64, d_row_off, d_row_cnt, pool_vals, pool_cols, M, d_vals,
                       d_cols, d_rowptr32);
    return hipGetLastError();
}

// ------------------------------------------------------ row canonicalisation

// Insertion sort of each row by column (rows are short; columns are unique).
template <typename TV>
__global__ void sort_rows_kernel(int M, const int *__restrict__ rowptr, int *__restrict__ cols,
                                 TV *__restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int b = rowptr[i], e = rowptr[i + 1];
    for (int j = b + 1; j < e; j++) {
        const int c = cols[j];
        const TV v = vals[j];
        int k = j - 1;
        while (k >= b && cols[k] > c) {
            cols[k + 1] = cols[k];
            vals[k + 1] = vals[k];
            k--;
        }
        cols[k + 1] = c;
        vals[k + 1] = v;
    }
}

hipError_t launch_sort_rows(int M, const int *rowptr, int *cols, void *vals, int vdt,
                            hipStream_t st) {
    if (M <= 0) return hipSuccess;
    if (vdt == 0)
        hipLaunchKernelGGL(sort_rows_kernel<float>, dim3((M + 255) / 256), dim3(256), 0, st, M,
                           rowptr, cols, static_cast<float *>(vals));
    else
        hipLaunchKernelGGL(sort_rows_kernel<double>, dim3((M + 255) / 256), dim3(256), 0, st, M,
                           rowptr, cols, static_cast<double *>(vals));
    return hipGetLastError();
}

// ------------------------------------------------------------------- SpMM

// y[i,d] = sum over the row's entries in stored order of w * x[col,d],
// starting from zero, in the output type (scipy csr_matvec(s)).
template <typename TW, typename TX, typename TY>
__global__ void spmm_kernel(int M, int D, const int *__restrict__ rowptr,
                            const int *__restrict__ cols, const TW *__restrict__ w,
                            const TX *__restrict__ x, TY *__restrict__ y) {
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(M) * D) return;
    const int i = static_cast<int>(gid / D), d = static_cast<int>(gid % D);
    TY acc = 0;
    const int b = rowptr[i], e = rowptr[i + 1];
    for (int j = b; j < e; j++) {
        const TY a = static_cast<TY>(w[j]);
        const TY v = static_cast<TY>(x[static_cast<size_t>(cols[j]) * D + d]);
        acc += a * v;
    }
    y[gid] = acc;
}

template <typename TW, typename TX, typename TY>
static hipError_t spmm_t(int M, int D, const int *rp, const int *c, const void *w, const void *x,
                         void *y, hipStream_t st) {
    const long long n = static_cast<long long>(M) * D;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((spmm_kernel<TW, TX, TY>), dim3(static_cast<unsigned>((n + 255) / 256)),
                       dim3(256), 0, st, M, D, rp, c, static_cast<const TW *>(w),
                       static_cast<const TX *>(x), static_cast<TY *>(y));
    return hipGetLastError();
}

hipError_t launch_spmm(int M, int D, const int *rp, const int *c, const void *w, int vdt,
                       const void *x, int xdt, void *y, int ydt, hipStream_t st) {
    const int code = vdt * 4 + xdt * 2 + ydt;
    switch (code) {
        case 0: return spmm_t<float, float, float>(M, D, rp, c, w, x, y, st);
        case 1: return spmm_t<float, float, double>(M, D, rp, c, w, x, y, st);
        case 3: return spmm_t<float, double, double>(M, D, rp, c, w, x, y, st);
        case 5: return spmm_t<double, float, double>(M, D, rp, c, w, x, y, st);
        case 7: return spmm_t<double, double, double>(M, D, rp, c, w, x, y, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace nat
