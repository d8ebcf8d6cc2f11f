// delaunay.h -- the per-thread steps of the opt-in GPU Delaunay build
// (parallel Bowyer-Watson in rounds), written once for the gfx950 kernels and
// for a sequential host emulation that checks the same code on the CPU.
//
// The reference builds the mesh on the host by inserting the points one at a
// time (delaunay.c:108-161 buildMesh, :561-614 conflict search, :619-688
// addPoint, :707-750 setNeighbours).  Here many points are inserted per
// round:
//   offer    every uninserted point offers itself to the tet holding it; the
//            tet keeps the offer of least priority (a hash of point and round);
//   claim    each tet's chosen point (a candidate) searches its cavity -- the
//            tets whose circumsphere holds it, grown from its tet -- and
//            claims the cavity and the ring of tets just outside it with an
//            atomic minimum of its priority;
//   win      a candidate that holds every tet it claimed wins and counts its
//            cavity's boundary faces;
//   faces    winners list their boundary faces at scanned offsets of a face
//            pool and pair the faces' edges;
//   commit   winners replace their cavities by the star of the new point
//            (the cavity's slots first, then fresh ones at scanned offsets);
//   relocate points whose tet was replaced walk to their new tet.
// Two winners' claims are disjoint, so neither tests or rewrites a tet of the
// other, and the circumsphere of a new tet lies inside the union of the
// spheres of the two tets that shared its base face (one in the cavity, one
// in the claimed outer ring): a winner's star never holds another winner's
// point, so the round equals a sequence of single insertions and the result
// is the Delaunay tetrahedralisation.  Every insphere / orient3d sign is the
// exact one (static filter, then exact.h), so the structure never breaks on
// rounding; exactly degenerate input (a repeated point, a point on the plane
// of a boundary face) is reported instead of perturbed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "exact.h"
#include "geom.h"

namespace nat {
namespace dl {

constexpr int kCapL = 2048;    // cavity + outer-ring tets of one insertion
constexpr int kVis = 4096;     // visited-set hash entries (power of two >= 2 kCapL)
constexpr int kFaceInts = 8;   // per boundary face: a b c outer kn e01 e02 e12
constexpr int kSlotInts = kCapL + 2 * kVis;
constexpr int kWalkSteps = 1 << 16;

// error bits
constexpr unsigned kErrCavity = 1u;      // cavity and outer ring over kCapL tets
constexpr unsigned kErrRepeated = 2u;    // empty cavity: a repeated point
constexpr unsigned kErrFlat = 4u;        // a new tet of zero volume
constexpr unsigned kErrTets = 8u;        // tet capacity exhausted
constexpr unsigned kErrWalk = 16u;       // walk too long
constexpr unsigned kErrOutside = 32u;    // a point outside the super tet
constexpr unsigned kErrWide = 64u;       // a test too wide for the exact evaluator
constexpr unsigned kErrShape = 128u;     // cavity boundary not a closed surface

struct State {
    const double4 *pts;          // [n + 4] points, then the super tet's vertices
    int n;
    int4 *tv;                    // [cap] vertices, positively oriented
    int4 *tn;                    // [cap] neighbour across face i (face_slot), -1 none
    int cap;
    int *loc;                    // [n] tet holding each uninserted point, -1 inserted
    unsigned long long *best;    // [cap] least offer per tet this round
    unsigned long long *owner;   // [cap] least claim per tet this round
    int *mark;                   // [cap] tag of the cavity a tet belongs to
    int *kround;                 // [cap] round a slot was last rewritten in
    int *flag;                   // [cap] 1 when the tet has an offer
    long long *flag_off;         // [cap + 1] scan of flag
    int *cand;                   // [slots] candidate points of the round
    int slots;
    int *slot_c;                 // [slots] cavity tets
    int *slot_o;                 // [slots] outer-ring tets
    int *slot_b;                 // [slots] boundary faces of a winner (0: lost)
    long long *face_off;         // [slots + 1] scan of slot_b
    int *faces;                  // [face_cap * kFaceInts] boundary faces of the winners
    long long face_cap;
    int *need;                   // [slots] fresh slots a winner needs
    long long *need_off;         // [slots + 1] scan of need
    int *scr;                    // [slots * kSlotInts] claim list, visited set
    unsigned *err;               // [1]
    unsigned *inserted;          // [1]
};

// Shared-word updates: the kernels pass device atomics (DevOps in
// delaunay_kernels.hip); the sequential host run passes plain updates.
struct HostOps {
    static void amin(unsigned long long *p, unsigned long long v) {
        if (v < *p) *p = v;
    }
    static void add(unsigned *p, unsigned v) { *p += v; }
    static void fail(unsigned *p, unsigned v) { *p |= v; }
    static void add64(unsigned long long *p, unsigned long long v) { *p += v; }
};

__host__ __device__ __forceinline__ d3 pt(const State &s, int i) {
    const double4 v = s.pts[i];
    return {v.x, v.y, v.z};
}

__host__ __device__ __forceinline__ unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Least wins: a hash of point and round, drawn afresh each round (fixed
// priorities per point, or priorities that favour points which lost many
// rounds, take 1.3x to 2.3x the rounds on uniform points).
__host__ __device__ __forceinline__ unsigned long long priority(int u, int round) {
    const unsigned h = mix32(static_cast<unsigned>(u) * 0x9e3779b1u ^
                             static_cast<unsigned>(round) * 0x85ebca77u);
    return (static_cast<unsigned long long>(h) << 32) | static_cast<unsigned>(u);
}

__host__ __device__ __forceinline__ int get4(const int4 &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
__host__ __device__ __forceinline__ void set4(int4 &v, int i, int x) {
    if (i == 0) v.x = x;
    else if (i == 1) v.y = x;
    else if (i == 2) v.z = x;
    else v.w = x;
}

// exact sign of insphere(a, b, c, d, e) for a positively oriented (a, b, c, d)
template <class O>
__host__ __device__ int sphere_sign(const d3 &a, const d3 &b, const d3 &c, const d3 &d,
                                    const d3 &e, unsigned *err) {
    const double v = insphere(a, b, c, d, e);
    if (!insphere_uncertain(a, b, c, d, e)) return v > 0 ? 1 : -1;
    const int s = insphere_exact_sign(a, b, c, d, e);
    if (s == 2) {
        O::fail(err, kErrWide);
        return v > 0 ? 1 : (v < 0 ? -1 : 0);
    }
    return s;
}

// exact sign of orient3d(a, b, c, d)
template <class O>
__host__ __device__ int orient_sign(const d3 &a, const d3 &b, const d3 &c, const d3 &d,
                                    unsigned *err) {
    const double v = orient3d(a, b, c, d);
    if (!orient3d_uncertain(a, b, c, d)) return v > 0 ? 1 : -1;
    const int s = orient3d_exact_sign(a, b, c, d);
    if (s == 2) {
        O::fail(err, kErrWide);
        return v > 0 ? 1 : (v < 0 ? -1 : 0);
    }
    return s;
}

__host__ __device__ __forceinline__ int *slot_scr(const State &s, int sl) {
    return s.scr + static_cast<size_t>(sl) * kSlotInts;
}

// ---------------------------------------------------------------- per round

// resets of the per-tet words, t < live tets
__host__ __device__ __forceinline__ void reset_tet(const State &s, int t) {
    s.best[t] = ~0ull;
    s.owner[t] = ~0ull;
}

// point u offers itself to its tet
template <class O>
__host__ __device__ void offer(const State &s, int u, int round, unsigned live) {
    const int t = s.loc[u];
    if (t < 0) return;
    if (static_cast<unsigned>(t) >= live) {
        O::fail(s.err, kErrShape);
        return;
    }
    O::amin(&s.best[t], priority(u, round));
}

__host__ __device__ __forceinline__ void flag_tet(const State &s, int t) {
    s.flag[t] = s.best[t] != ~0ull ? 1 : 0;
}

// the first `slots` tets with an offer, in tet order, give the candidates
__host__ __device__ __forceinline__ void pick(const State &s, int t) {
    if (!s.flag[t]) return;
    const long long k = s.flag_off[t];
    if (k < s.slots) s.cand[k] = static_cast<int>(s.best[t] & 0xffffffffull);
}

__host__ __device__ __forceinline__ int n_cand(const State &s, unsigned live) {
    const long long k = s.flag_off[live];
    return static_cast<int>(k < s.slots ? k : s.slots);
}

// visited set of one search: (tet, tag) pairs, other tags count as empty
__host__ __device__ __forceinline__ bool visit(int *vis, int t, int tag) {
    unsigned h = mix32(static_cast<unsigned>(t)) & (kVis - 1);
    for (int probe = 0; probe < kVis; probe++) {
        int *e = vis + 2 * h;
        if (e[1] != tag) {
            e[0] = t;
            e[1] = tag;
            return true;
        }
        if (e[0] == t) return false;
        h = (h + 1) & (kVis - 1);
    }
    return false;  // unreachable: at most kCapL + 1 entries per search
}

__host__ __device__ __forceinline__ int tag_of(const State &s, int sl, int round) {
    return round * s.slots + sl + 1;
}

// Cavity of candidate sl and its claims: the tets whose circumsphere holds
// the point, grown breadth-first from its tet (the claim list is the queue),
// at cl[0, C); the tets tested outside -- the ring around the cavity -- at
// cl[kCapL - O, kCapL).
template <class O>
__host__ __device__ void claim(const State &s, int sl, int round, unsigned live) {
    const int p = s.cand[sl];
    const d3 q = pt(s, p);
    int *cl = slot_scr(s, sl);
    int *vis = cl + kCapL;
    const int tag = tag_of(s, sl, round);
    int C = 0, Oc = 0;
    s.slot_c[sl] = 0;
    s.slot_o[sl] = 0;
    const int start = s.loc[p];
    if (start < 0 || static_cast<unsigned>(start) >= live) {
        O::fail(s.err, kErrShape);
        return;
    }
    auto test = [&](int t) {  // false on overflow
        if (!visit(vis, t, tag)) return true;
        if (C + Oc >= kCapL) return false;
        const int4 v = s.tv[t];
        if (sphere_sign<O>(pt(s, v.x), pt(s, v.y), pt(s, v.z), pt(s, v.w), q, s.err) > 0)
            cl[C++] = t;
        else
            cl[kCapL - 1 - Oc++] = t;
        return true;
    };
    if (!test(start)) {
        O::fail(s.err, kErrCavity);
        return;
    }
    for (int head = 0; head < C; head++) {
        const int4 nb = s.tn[cl[head]];
        for (int k = 0; k < 4; k++) {
            const int x = get4(nb, k);
            if (x < 0) continue;
            if (static_cast<unsigned>(x) >= live) {
                O::fail(s.err, kErrShape);
                return;
            }
            if (!test(x)) {
                O::fail(s.err, kErrCavity);
                return;
            }
        }
    }
    if (C == 0) {
        O::fail(s.err, kErrRepeated);
        return;
    }
    s.slot_c[sl] = C;
    s.slot_o[sl] = Oc;
    const unsigned long long pr = priority(p, round);
    for (int i = 0; i < C; i++) O::amin(&s.owner[cl[i]], pr);
    for (int i = 0; i < Oc; i++) O::amin(&s.owner[cl[kCapL - 1 - i]], pr);
}

// Win check.  A winner marks its cavity and counts its boundary faces
// (slot_b = B); a loser gets slot_b = 0.
template <class O>
__host__ __device__ void win(const State &s, int sl, int round) {
    s.slot_b[sl] = 0;
    s.need[sl] = 0;
    const int C = s.slot_c[sl], Oc = s.slot_o[sl];
    if (C == 0) return;
    const int p = s.cand[sl];
    const unsigned long long pr = priority(p, round);
    const int *cl = slot_scr(s, sl);
    bool won = true;
    for (int i = 0; i < C && won; i++) won = s.owner[cl[i]] == pr;
    for (int i = 0; i < Oc && won; i++) won = s.owner[cl[kCapL - 1 - i]] == pr;
    if (!won) return;
    const int tag = tag_of(s, sl, round);
    for (int i = 0; i < C; i++) s.mark[cl[i]] = tag;
    int B = 0;
    for (int ci = 0; ci < C; ci++) {
        const int4 nb = s.tn[cl[ci]];
        for (int f = 0; f < 4; f++) {
            const int x = get4(nb, f);
            if (x < 0 || s.mark[x] != tag) B++;
        }
    }
    s.slot_b[sl] = B;
}

// A winner's boundary faces at faces[face_off[sl] ...] (each with the outer
// tet and that tet's face index), checked for positive volume with the new
// point, and for each face the faces that share its edges (a b), (a c),
// (b c); need = max(B - C, 0) fresh slots.  A winner whose faces do not fit
// the pool this round waits (slot_b = 0).
template <class O>
__host__ __device__ void faces(const State &s, int sl, int round) {
    const int B = s.slot_b[sl];
    if (B == 0) return;
    const int C = s.slot_c[sl];
    const int p = s.cand[sl];
    const long long off = s.face_off[sl];
    if (off + B > s.face_cap) {
        s.slot_b[sl] = 0;
        return;
    }
    const int *cl = slot_scr(s, sl);
    int *bd = s.faces + off * kFaceInts;
    const int tag = tag_of(s, sl, round);
    int j = 0;
    for (int ci = 0; ci < C; ci++) {
        const int t = cl[ci];
        const int4 v = s.tv[t];
        const int4 nb = s.tn[t];
        for (int f = 0; f < 4; f++) {
            const int x = get4(nb, f);
            if (x >= 0 && s.mark[x] == tag) continue;
            const int a = get4(v, face_slot(f, 0)), b = get4(v, face_slot(f, 1)),
                      c = get4(v, face_slot(f, 2));
            int kn = -1;
            if (x >= 0) {
                const int4 w = s.tv[x];
                int miss = 0;
                for (int k = 0; k < 4; k++) {
                    const int y = get4(w, k);
                    if (y != a && y != b && y != c) {
                        kn = 3 - k;
                        miss++;
                    }
                }
                if (miss != 1) {
                    O::fail(s.err, kErrShape);
                    s.slot_b[sl] = 0;
                    return;
                }
            }
            if (j >= B) {
                O::fail(s.err, kErrShape);
                s.slot_b[sl] = 0;
                return;
            }
            int *e = bd + kFaceInts * j++;
            e[0] = a;
            e[1] = b;
            e[2] = c;
            e[3] = x;
            e[4] = kn;
        }
    }
    // the new tets (a, b, c, p) must have positive volume
    const d3 q = pt(s, p);
    for (int k = 0; k < B; k++) {
        const int *e = bd + kFaceInts * k;
        if (orient_sign<O>(pt(s, e[0]), pt(s, e[1]), pt(s, e[2]), q, s.err) <= 0) {
            O::fail(s.err, kErrFlat);
            s.slot_b[sl] = 0;
            return;
        }
    }
    // edge partners: each boundary edge is shared by exactly two faces
    for (int k = 0; k < B; k++) {
        int *e = bd + kFaceInts * k;
        const int a = e[0], b = e[1], c = e[2];
        int x01 = -1, x02 = -1, x12 = -1;
        for (int k2 = 0; k2 < B; k2++) {
            if (k2 == k) continue;
            const int *g = bd + kFaceInts * k2;
            const bool ia = g[0] == a || g[1] == a || g[2] == a;
            const bool ib = g[0] == b || g[1] == b || g[2] == b;
            const bool ic = g[0] == c || g[1] == c || g[2] == c;
            if (ia && ib) x01 = x01 == -1 ? k2 : -2;
            if (ia && ic) x02 = x02 == -1 ? k2 : -2;
            if (ib && ic) x12 = x12 == -1 ? k2 : -2;
        }
        if (x01 < 0 || x02 < 0 || x12 < 0) {
            O::fail(s.err, kErrShape);
            s.slot_b[sl] = 0;
            return;
        }
        e[5] = x01;
        e[6] = x02;
        e[7] = x12;
    }
    s.need[sl] = B > C ? B - C : 0;
}

// The star of a winner: new tet j = (a, b, c, p) in slot cl[j] (j < C) or in
// fresh slot live + need_off[sl] + (j - C); face 0 (a b c) faces the outer
// tet, faces 1, 2, 3 (p with edges (a b), (a c), (b c)) the partners.  A
// cavity of more tets than boundary faces leaves slots cl[B, C) free: they
// are marked dead (vertex -1, then the slot of the star's first tet, where
// the points they held start their walk) and dropped when the mesh is packed.
template <class O>
__host__ __device__ void commit(const State &s, int sl, int round, unsigned live) {
    const int B = s.slot_b[sl];
    if (B == 0) return;
    const int C = s.slot_c[sl];
    const int p = s.cand[sl];
    const int *cl = slot_scr(s, sl);
    const int *bd = s.faces + s.face_off[sl] * kFaceInts;
    const long long fresh = static_cast<long long>(live) + s.need_off[sl];
    if (fresh + (B - C) > s.cap) {
        O::fail(s.err, kErrTets);
        return;
    }
    auto slot = [&](int j) { return j < C ? cl[j] : static_cast<int>(fresh + (j - C)); };
    for (int j = 0; j < B; j++) {
        const int *e = bd + kFaceInts * j;
        const int sj = slot(j);
        s.tv[sj] = make_int4(e[0], e[1], e[2], p);
        s.tn[sj] = make_int4(e[3], slot(e[5]), slot(e[6]), slot(e[7]));
        s.kround[sj] = round;
        if (e[3] >= 0) {
            int4 w = s.tn[e[3]];
            set4(w, e[4], sj);
            s.tn[e[3]] = w;
        }
    }
    for (int j = B; j < C; j++) {
        s.tv[cl[j]] = make_int4(-1, cl[0], -1, -1);
        s.tn[cl[j]] = make_int4(-1, -1, -1, -1);
        s.kround[cl[j]] = round;
    }
    s.loc[p] = -1;
    O::add(s.inserted, 1u);
}

// a point whose tet was rewritten this round walks to its tet (visibility
// walk, first separating face, as the query path's start-tet walk)
template <class O>
__host__ __device__ void relocate(const State &s, int u, int round) {
    int t = s.loc[u];
    if (t < 0 || s.kround[t] != round) return;
    const d3 q = pt(s, u);
    int4 v = s.tv[t];
    if (v.x < 0) {  // a slot freed this round: start in the star that replaced it
        t = v.y;
        v = s.tv[t];
    }
    int steps = 0;
    int i = 0;
    while (i < 4) {
        const d3 a = pt(s, get4(v, face_slot(i, 0)));
        const d3 b = pt(s, get4(v, face_slot(i, 1)));
        const d3 c = pt(s, get4(v, face_slot(i, 2)));
        if (orient_sign<O>(a, b, c, q, s.err) < 0) {
            const int nb = get4(s.tn[t], i);
            if (nb < 0) {
                O::fail(s.err, kErrOutside);
                return;
            }
            t = nb;
            v = s.tv[t];
            i = 0;
            if (++steps > kWalkSteps) {
                O::fail(s.err, kErrWalk);
                return;
            }
            continue;
        }
        i++;
    }
    s.loc[u] = t;
}

// ------------------------------------------------------- Delaunay audit

// Per tet t: its orientation, the symmetry of its neighbour links and, per
// face with a neighbour, whether the neighbour's far vertex is strictly
// inside t's circumsphere (local Delaunay on every face <=> Delaunay).
// out[0] += faces tested, [1] += non-positive tets, [2] += broken links,
// [3] += faces violating the empty-sphere property.
template <class O>
__host__ __device__ void audit_tet(const double4 *pts, const int4 *tv, const int4 *tn, int T,
                                   int t, unsigned *err, unsigned long long *out) {
    State s{};
    s.pts = pts;
    const int4 v = tv[t];
    if (v.x < 0) return;  // a freed slot
    const d3 P0 = pt(s, v.x), P1 = pt(s, v.y), P2 = pt(s, v.z), P3 = pt(s, v.w);
    unsigned long long faces = 0, neg = 0, links = 0, bad = 0;
    if (orient_sign<O>(P0, P1, P2, P3, err) <= 0) neg++;
    const int4 nb = tn[t];
    for (int f = 0; f < 4; f++) {
        const int x = get4(nb, f);
        if (x < 0) continue;
        if (x >= T) {
            links++;
            continue;
        }
        const int a = get4(v, face_slot(f, 0)), b = get4(v, face_slot(f, 1)),
                  c = get4(v, face_slot(f, 2));
        const int4 w = tv[x];
        int far = -1, miss = 0, back = -1;
        for (int j = 0; j < 4; j++) {
            const int y = get4(w, j);
            if (y != a && y != b && y != c) {
                far = y;
                back = 3 - j;
                miss++;
            }
        }
        if (miss != 1 || get4(tn[x], back) != t) {
            links++;
            continue;
        }
        faces++;
        if (sphere_sign<O>(P0, P1, P2, P3, pt(s, far), err) > 0) bad++;
    }
    if (faces) O::add64(out + 0, faces);
    if (neg) O::add64(out + 1, neg);
    if (links) O::add64(out + 2, links);
    if (bad) O::add64(out + 3, bad);
}

}  // namespace dl
}  // namespace nat
