// mesh_builder.cpp -- host Bowyer-Watson tetrahedralisation, written to
// reproduce the reference mesh exactly so that the device query path sees
// the same tets in the same order (delaunay.c:83-161, 383-407, 513-750,
// 832-863; utils.c:161-185, 268-291; kdtree.c:183-233).
//
// The behaviours that fix the result, and that are therefore kept:
//   * insertion in input order, each point located by a visibility walk
//     from the oldest live tet (the head of the creation-ordered list);
//   * conflict region grown by a LIFO search that re-tests non-conflicting
//     tets each time they are reached;
//   * new tets created in (conflict order) x (face 0..3) order with vertices
//     (face vertices, new point) and slot 0 facing outward;
//   * internal slots 1..3 linked by matching the edges (p0,p1), (p2,p0),
//     (p1,p2) of the new tets;
//   * a point whose insphere test is exactly 0, or that makes a new tet flat,
//     is moved by (u - 1/2) * 1e-4 * (attempt + 1) per coordinate, u drawn
//     from the libc additive-feedback generator in the state of a fresh
//     process, and its insertion restarts;
//   * k-d tree over tet centroids inserted in list order, node i = tet i.
// The perturbation generator is private to each build, so a build does not
// depend on what else the process did before.
#include "mesh_builder.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <thread>

#include "geom.h"

namespace nat {
namespace {

// Additive feedback generator r[i] = r[i-3] + r[i-31] (mod 2^32), output
// r[i] >> 1, initial state from the Lehmer sequence 16807^i * seed mod
// (2^31 - 1) with the first 310 outputs discarded: the generator behind
// libc rand() after srand(seed).
class AdditiveFeedbackRandom {
  public:
    explicit AdditiveFeedbackRandom(uint32_t seed) {
        uint64_t w = seed ? seed : 1;
        st_[0] = static_cast<uint32_t>(w);
        for (int i = 1; i < 31; i++) {
            w = (16807ull * w) % 2147483647ull;
            st_[i] = static_cast<uint32_t>(w);
        }
        front_ = 3;
        rear_ = 0;
        for (int i = 0; i < 310; i++) next();
    }
    int next() {
        st_[front_] += st_[rear_];
        const int out = static_cast<int>(st_[front_] >> 1);
        front_ = front_ == 30 ? 0 : front_ + 1;
        rear_ = rear_ == 30 ? 0 : rear_ + 1;
        return out;
    }

  private:
    uint32_t st_[31];
    int front_, rear_;
};

struct EdgeKey {
    int lo, hi, tet, slot, order;
};

class Builder {
  public:
    explicit Builder(int n) : n_(n), rng_(1) {}

    int run(const double *pts, PackedMesh *out);

  private:
    using I4 = std::array<int, 4>;

    // one 32-byte record per tet slot: the walk touches a single cache line
    struct alignas(32) Rec {
        I4 p, s;
    };
    struct PView {
        std::vector<Rec> *r;
        I4 &operator[](size_t t) const { return (*r)[t].p; }
    };
    struct SView {
        std::vector<Rec> *r;
        I4 &operator[](size_t t) const { return (*r)[t].s; }
    };

    int n_;
    std::vector<d3> X_;
    std::vector<Rec> R_;
    PView T_{&R_};  // vertex slots of tet t
    SView S_{&R_};  // neighbour across face i of tet t
    std::vector<uint64_t> mark_;
    std::vector<int> prev_, next_;
    int head_ = -1, tail_ = -1, live_ = 0;
    std::vector<int> dead_;
    uint64_t gen_ = 0;
    std::vector<int> conflicts_, updates_, stack_;
    struct NbUpdate {
        int tet, slot, old;
    };
    std::vector<NbUpdate> nb_updates_;
    std::vector<EdgeKey> edges_;
    AdditiveFeedbackRandom rng_;
    int coplanar_ = 0, cospherical_ = 0, anomalies_ = 0;

    int new_tet() {
        int t;
        if (!dead_.empty()) {
            t = dead_.back();
            dead_.pop_back();
        } else {
            t = static_cast<int>(R_.size());
            R_.push_back({{0, 0, 0, 0}, {-1, -1, -1, -1}});
            mark_.push_back(0);
            prev_.push_back(-1);
            next_.push_back(-1);
        }
        S_[t] = {-1, -1, -1, -1};
        mark_[t] = 0;
        return t;
    }
    void append(int t) {
        prev_[t] = tail_;
        next_[t] = -1;
        if (tail_ >= 0)
            next_[tail_] = t;
        else
            head_ = t;
        tail_ = t;
        live_++;
    }
    void unlink(int t) {
        if (prev_[t] >= 0)
            next_[prev_[t]] = next_[t];
        else
            head_ = next_[t];
        if (next_[t] >= 0)
            prev_[next_[t]] = prev_[t];
        else
            tail_ = prev_[t];
        live_--;
    }
    // delaunay.c:513-522
    void perturb(int v, int attempt) {
        double *c = &X_[v].x;
        for (int i = 0; i < 3; i++) {
            const double u = static_cast<double>(rng_.next()) / (2147483647.0 + 1.0);
            const double p = (u - 0.5) * 0.0001 * static_cast<double>(attempt + 1);
            c[i] += p;
        }
    }
    // delaunay.c:383-407 with the list head as the start (no k-d tree yet)
    int walk(int v) {
        int s = head_;
        const d3 p = X_[v];
        long steps = 0;
        const long limit = 16L * (static_cast<long>(R_.size()) + 16);
        for (int i = 0; i < 4; i++) {
            const I4 &tv = T_[s];
            if (orient3d(X_[tv[face_slot(i, 0)]], X_[tv[face_slot(i, 1)]],
                         X_[tv[face_slot(i, 2)]], p) < 0) {
                const int nb = S_[s][i];
                if (nb < 0) return -1;
                s = nb;
                i = -1;
                if (++steps > limit) return -1;
            }
        }
        return s;
    }
    // delaunay.c:531-555: 1 = outside the circumsphere, 0 = inside,
    // -1 = exactly on it, -2 = tet not positively oriented
    int delaunay_state(int t, int v) {
        const I4 &tv = T_[t];
        const double o = orient3d(X_[tv[0]], X_[tv[1]], X_[tv[2]], X_[tv[3]]);
        if (o <= 0) return -2;
        const double ins = insphere(X_[tv[0]], X_[tv[1]], X_[tv[2]], X_[tv[3]], X_[v]);
        if (ins == 0) return -1;
        return ins < 0 ? 1 : 0;
    }
    // delaunay.c:561-614
    int find_conflicts(int v) {
        for (;;) {
            const int s0 = walk(v);
            if (s0 < 0) return -3;
            const uint64_t g = ++gen_;
            conflicts_.clear();
            stack_.clear();
            stack_.push_back(s0);
            bool restart = false;
            while (!stack_.empty()) {
                const int cur = stack_.back();
                stack_.pop_back();
                if (mark_[cur] == g) continue;
                int st = delaunay_state(cur, v);
                if (st == -2) return -2;
                if (st == -1) {
                    cospherical_++;
                    int i = 0;
                    while (st == -1) {
                        perturb(v, i);
                        st = delaunay_state(cur, v);
                        if (st == -2) return -2;
                        i++;
                    }
                    restart = true;
                    break;
                }
                if (st == 0) {
                    mark_[cur] = g;
                    conflicts_.push_back(cur);
                    for (int i = 0; i < 4; i++)
                        if (S_[cur][i] >= 0) stack_.push_back(S_[cur][i]);
                }
            }
            if (!restart) return 0;
        }
    }
    // delaunay.c:707-750
    void link_new_tets() {
        edges_.clear();
        for (size_t j = 0; j < updates_.size(); j++) {
            const int t = updates_[j];
            const I4 &p = T_[t];
            const int pairs[3][2] = {{p[0], p[1]}, {p[2], p[0]}, {p[1], p[2]}};
            for (int e = 0; e < 3; e++) {
                const int a = pairs[e][0], b = pairs[e][1];
                edges_.push_back({std::min(a, b), std::max(a, b), t, e + 1,
                                  static_cast<int>(edges_.size())});
            }
        }
        std::sort(edges_.begin(), edges_.end(), [](const EdgeKey &x, const EdgeKey &y) {
            if (x.lo != y.lo) return x.lo < y.lo;
            if (x.hi != y.hi) return x.hi < y.hi;
            return x.order < y.order;
        });
        size_t run = 1;
        for (size_t i = 0; i + 1 < edges_.size(); i++) {
            if (edges_[i].lo == edges_[i + 1].lo && edges_[i].hi == edges_[i + 1].hi) {
                S_[edges_[i].tet][edges_[i].slot] = edges_[i + 1].tet;
                S_[edges_[i + 1].tet][edges_[i + 1].slot] = edges_[i].tet;
                run++;
            } else {
                if (run != 2) anomalies_++;
                run = 1;
            }
        }
        if (!edges_.empty() && run != 2) anomalies_++;
    }
    // delaunay.c:619-688, plus the pool push of delaunay.c:124-127
    int add_point(int v) {
        for (;;) {
            const int rc = find_conflicts(v);
            if (rc) return rc;
            const uint64_t g = gen_;
            updates_.clear();
            nb_updates_.clear();
            bool restart = false;
            for (size_t j = 0; j < conflicts_.size() && !restart; j++) {
                const int s = conflicts_[j];
                for (int i = 0; i < 4; i++) {
                    const int v1 = T_[s][face_slot(i, 0)];
                    const int v2 = T_[s][face_slot(i, 1)];
                    const int v3 = T_[s][face_slot(i, 2)];
                    const int nb = S_[s][i];
                    if (nb >= 0 && mark_[nb] == g) continue;
                    const int t = new_tet();
                    T_[t] = {v1, v2, v3, v};
                    double o = orient3d(X_[v1], X_[v2], X_[v3], X_[v]);
                    if (o <= 0) {
                        coplanar_++;
                        int attempt = 0;
                        while (o <= 0) {
                            perturb(v, attempt);
                            o = orient3d(X_[v1], X_[v2], X_[v3], X_[v]);
                            attempt++;
                        }
                        dead_.push_back(t);
                        for (size_t k = nb_updates_.size(); k-- > 0;)
                            S_[nb_updates_[k].tet][nb_updates_[k].slot] = nb_updates_[k].old;
                        nb_updates_.clear();
                        for (int u : updates_) {
                            unlink(u);
                            dead_.push_back(u);
                        }
                        updates_.clear();
                        conflicts_.clear();
                        restart = true;
                        break;
                    }
                    S_[t][0] = nb;
                    if (nb >= 0) {
                        int k = 0;
                        while (k < 4 && S_[nb][k] != s) k++;
                        if (k == 4) return -2;
                        S_[nb][k] = t;
                        nb_updates_.push_back({nb, k, s});
                    }
                    updates_.push_back(t);
                    append(t);
                }
            }
            if (restart) continue;
            link_new_tets();
            for (int s : conflicts_) unlink(s);
            for (int s : conflicts_) dead_.push_back(s);
            return 0;
        }
    }
};

int Builder::run(const double *pts, PackedMesh *out) {
    X_.resize(static_cast<size_t>(n_) + 4);
    for (int i = 0; i < n_; i++) X_[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};

    double sv[12];
    super_vertices(pts, n_, sv);
    for (int k = 0; k < 4; k++) X_[n_ + k] = {sv[3 * k], sv[3 * k + 1], sv[3 * k + 2]};

    const size_t reserve = static_cast<size_t>(n_) * 8 + 64;
    R_.reserve(reserve);
    mark_.reserve(reserve);
    prev_.reserve(reserve);
    next_.reserve(reserve);
    const int sup = new_tet();
    T_[sup] = {n_, n_ + 1, n_ + 2, n_ + 3};
    append(sup);

    for (int i = 0; i < n_; i++) {
        const int rc = add_point(i);
        if (rc) return rc;
        conflicts_.clear();
        updates_.clear();
        nb_updates_.clear();
    }

    // pack in list order (natural_insertionfree.c:413-435)
    const int T = live_;
    std::vector<int> index(R_.size(), -1);
    std::vector<int> order;
    order.reserve(T);
    for (int t = head_; t >= 0; t = next_[t]) {
        index[t] = static_cast<int>(order.size());
        order.push_back(t);
    }
    out->n_data = n_;
    out->n_tets = T;
    out->verts.resize((static_cast<size_t>(n_) + 4) * 3);
    for (int i = 0; i < n_ + 4; i++) {
        out->verts[3 * i] = X_[i].x;
        out->verts[3 * i + 1] = X_[i].y;
        out->verts[3 * i + 2] = X_[i].z;
    }
    out->tet_v.resize(static_cast<size_t>(T) * 4);
    out->tet_n.resize(static_cast<size_t>(T) * 4);
    for (int i = 0; i < T; i++) {
        const int t = order[i];
        for (int j = 0; j < 4; j++) {
            out->tet_v[4 * i + j] = T_[t][j];
            out->tet_n[4 * i + j] = S_[t][j] >= 0 ? index[S_[t][j]] : -1;
        }
    }

    // k-d tree over centroids, delaunay.c:142-159 and kdtree.c:183-233
    out->kd_left.assign(T, -1);
    out->kd_right.assign(T, -1);
    out->kd_pos.resize(static_cast<size_t>(T) * 3);
    int depth_max = T > 0 ? 1 : 0;
    for (int i = 0; i < T; i++) {
        const I4 &p = T_[order[i]];
        double *c = &out->kd_pos[3 * i];
        c[0] = (X_[p[0]].x + X_[p[1]].x + X_[p[2]].x + X_[p[3]].x) / 4;
        c[1] = (X_[p[0]].y + X_[p[1]].y + X_[p[2]].y + X_[p[3]].y) / 4;
        c[2] = (X_[p[0]].z + X_[p[1]].z + X_[p[2]].z + X_[p[3]].z) / 4;
        if (i == 0) {
            for (int d = 0; d < 3; d++) out->kd_rect[d] = out->kd_rect[3 + d] = c[d];
            continue;
        }
        int cur = 0, depth = 0;
        for (;;) {
            const int dir = depth % 3;
            int &child = c[dir] < out->kd_pos[3 * cur + dir] ? out->kd_left[cur]
                                                                 : out->kd_right[cur];
            depth++;
            if (child < 0) {
                child = i;
                break;
            }
            cur = child;
        }
        depth_max = std::max(depth_max, depth + 1);
        for (int d = 0; d < 3; d++) {
            if (c[d] < out->kd_rect[d]) out->kd_rect[d] = c[d];
            if (c[d] > out->kd_rect[3 + d]) out->kd_rect[3 + d] = c[d];
        }
    }
    out->kd_depth = depth_max;
    out->coplanar = coplanar_;
    out->cospherical = cospherical_;
    out->edge_anomalies = anomalies_;
    return 0;
}

// k-d tree of median splits over cent[idx[lo, hi)]: the subtree's root moves
// to idx[lo], its left subtree to [lo + 1, lo + 1 + L), its right after.
void kd_split(std::vector<int> &idx, const std::vector<double> &cent, int lo, int hi, int depth,
              std::vector<int32_t> &left, std::vector<int32_t> &right, int &depth_max,
              int spawn_depth) {
    while (hi > lo) {
        depth_max = std::max(depth_max, depth + 1);
        const int dir = depth % 3;
        const int L = (hi - lo) / 2;
        auto key = [&](int t) { return cent[3 * static_cast<size_t>(t) + dir]; };
        std::nth_element(idx.begin() + lo, idx.begin() + lo + L, idx.begin() + hi,
                         [&](int a, int b) { return key(a) < key(b) || (key(a) == key(b) && a < b); });
        std::swap(idx[lo], idx[lo + L]);
        const int l0 = lo + 1, r0 = lo + 1 + L;
        left[lo] = L > 0 ? l0 : -1;
        right[lo] = hi > r0 ? r0 : -1;
        if (depth < spawn_depth && L > 4096) {
            int dm = 0;
            std::thread th([&, l0, r0, depth] {
                kd_split(idx, cent, l0, r0, depth + 1, left, right, dm, spawn_depth);
            });
            int dr = 0;
            kd_split(idx, cent, r0, hi, depth + 1, left, right, dr, spawn_depth);
            th.join();
            depth_max = std::max({depth_max, dm, dr});
            return;
        }
        kd_split(idx, cent, l0, r0, depth + 1, left, right, depth_max, spawn_depth);
        lo = r0;
        depth++;
    }
}

}  // namespace

void super_vertices(const double *pts, int n, double *out) {
    // super simplex, delaunay.c:832-863 and getRange delaunay.c:1086-1111
    d3 mn = {pts[0], pts[1], pts[2]}, mx = mn;
    for (int i = 0; i < n; i++) {
        const d3 p = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        mx.x = mx.x < p.x ? p.x : mx.x;
        mx.y = mx.y < p.y ? p.y : mx.y;
        mx.z = mx.z < p.z ? p.z : mx.z;
        mn.x = mn.x > p.x ? p.x : mn.x;
        mn.y = mn.y > p.y ? p.y : mn.y;
        mn.z = mn.z > p.z ? p.z : mn.z;
    }
    const d3 r = {mx.x - mn.x, mx.y - mn.y, mx.z - mn.z};
    const d3 s[4] = {{mn.x + r.x / 2, mx.y + 3 * r.y, mn.z - r.z},
                     {mx.x + 2 * r.x, mn.y - 2 * r.y, mn.z - r.z},
                     {mn.x - 2 * r.x, mn.y - 2 * r.y, mn.z - r.z},
                     {mn.x + r.x / 2, mn.y + r.y / 2, mx.z + 2 * r.z}};
    for (int k = 0; k < 4; k++) {
        out[3 * k] = s[k].x;
        out[3 * k + 1] = s[k].y;
        out[3 * k + 2] = s[k].z;
    }
}

void pack_balanced(int n, std::vector<double> verts, const std::vector<int32_t> &tet_v_in,
                   const std::vector<int32_t> &tet_n_in, int threads, PackedMesh *out) {
    // drop freed slots (first vertex -1), renumbering the neighbour links
    const int T0 = static_cast<int>(tet_v_in.size() / 4);
    std::vector<int> keep(T0, -1);
    int T = 0;
    for (int t = 0; t < T0; t++)
        if (tet_v_in[4 * static_cast<size_t>(t)] >= 0) keep[t] = T++;
    std::vector<int32_t> tet_v(static_cast<size_t>(T) * 4), tet_n(static_cast<size_t>(T) * 4);
    for (int t = 0; t < T0; t++) {
        if (keep[t] < 0) continue;
        for (int j = 0; j < 4; j++) {
            tet_v[4 * static_cast<size_t>(keep[t]) + j] = tet_v_in[4 * static_cast<size_t>(t) + j];
            const int x = tet_n_in[4 * static_cast<size_t>(t) + j];
            tet_n[4 * static_cast<size_t>(keep[t]) + j] = x >= 0 ? keep[x] : -1;
        }
    }
    // centroids as the reference's k-d tree takes them ((a + b + c + d) / 4)
    std::vector<double> cent(static_cast<size_t>(T) * 3);
    for (int t = 0; t < T; t++)
        for (int d = 0; d < 3; d++) {
            const int32_t *p = &tet_v[4 * static_cast<size_t>(t)];
            cent[3 * static_cast<size_t>(t) + d] =
                (verts[3 * p[0] + d] + verts[3 * p[1] + d] + verts[3 * p[2] + d] +
                 verts[3 * p[3] + d]) / 4;
        }
    std::vector<int> idx(T);
    for (int t = 0; t < T; t++) idx[t] = t;
    out->kd_left.assign(T, -1);
    out->kd_right.assign(T, -1);
    int depth_max = 0;
    int spawn = 0;
    while ((1 << spawn) < threads && spawn < 6) spawn++;
    kd_split(idx, cent, 0, T, 0, out->kd_left, out->kd_right, depth_max, spawn);
    std::vector<int> inv(T);
    for (int i = 0; i < T; i++) inv[idx[i]] = i;
    out->n_data = n;
    out->n_tets = T;
    out->verts = std::move(verts);
    out->tet_v.resize(static_cast<size_t>(T) * 4);
    out->tet_n.resize(static_cast<size_t>(T) * 4);
    out->kd_pos.resize(static_cast<size_t>(T) * 3);
    for (int d = 0; d < 3; d++) {
        out->kd_rect[d] = T ? cent[d] : 0.0;
        out->kd_rect[3 + d] = T ? cent[d] : 0.0;
    }
    for (int i = 0; i < T; i++) {
        const int t = idx[i];
        for (int j = 0; j < 4; j++) {
            out->tet_v[4 * static_cast<size_t>(i) + j] = tet_v[4 * static_cast<size_t>(t) + j];
            const int x = tet_n[4 * static_cast<size_t>(t) + j];
            out->tet_n[4 * static_cast<size_t>(i) + j] = x >= 0 ? inv[x] : -1;
        }
        for (int d = 0; d < 3; d++) {
            const double c = cent[3 * static_cast<size_t>(t) + d];
            out->kd_pos[3 * static_cast<size_t>(i) + d] = c;
            out->kd_rect[d] = std::min(out->kd_rect[d], c);
            out->kd_rect[3 + d] = std::max(out->kd_rect[3 + d], c);
        }
    }
    out->kd_depth = depth_max;
    out->coplanar = out->cospherical = out->edge_anomalies = 0;
}

int build_packed_mesh(const double *points, int n, PackedMesh *out) {
    if (!points || n <= 0 || !out) return -1;
    Builder b(n);
    return b.run(points, out);
}

}  // namespace nat
