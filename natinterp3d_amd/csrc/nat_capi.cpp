// nat_capi.cpp -- the C ABI (include/natinterp3d_hip.h): mesh build and
// upload, device workspace, the capacity-level driver of the row kernel and
// CSR assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <sys/mman.h>
#include <vector>

#include "delaunay_build.h"
#include "exact.h"
#include "geom.h"
#include "mesh_builder.h"
#include "nat_internal.h"
#include "natinterp3d_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_OR_FAIL(expr)                                                               \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(-10, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),      \
                        __FILE__, __LINE__);                                            \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need, bool zero = false) {
        if (need <= bytes && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        size_t alloc = std::max<size_t>(need, 256);
        hipError_t e = hipMalloc(&p, alloc);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = alloc;
        if (zero) return hipMemset(p, 0, alloc);
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

constexpr int kLevels = 4;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

struct nat_mesh {
    int device = 0;
    int num_cus = 0;
    // one problem per part (a plain mesh has one); global numbering in dev
    std::vector<nat::PackedMesh> parts;
    std::vector<long long> vbase, tbase;  // [P + 1] data-vertex and tet offsets
    int kd_depth_max = 0;
    std::vector<nat::LocParams> lph;      // [P]
    nat::DevMesh dev{};
    DevBuf verts, tet_v, tet_n, fat, sph, kd, cell_start, cell_pts, lp;
    DevBuf vorig;                         // [N+4P] device vertex -> key index
    DevBuf vals_dev;                      // backward: values in device vertex order
    DevBuf vals_fwd;                      // nat_weights_spmm: values in device vertex order
    DevBuf old_vol;                       // step-3 table (built on the first compute)
    DevBuf itab;                          // the search's item entries (Topo::itab)
    bool old_vol_tried = false;
    DevBuf prob, qoff;                    // batch calls: problem of each query
    std::mutex mu;
    int builder = 0;                      // 0 host build, 1 GPU build
    int build_rounds = 0;
    double build_seconds = 0.0;

    // row-path policy (nat_policy).  route[p] = visited-tet budget of the
    // group kernel | first thread-per-query level << 8, read by the kernels at
    // launch time; grids are resident blocks times the multipliers.
    std::vector<unsigned char> budget;  // [P]
    DevBuf route;
    int first_level = 1;
    int work_sort = 1;
    int topo_grid = 1, geom_grid = 1, geom_lanes = 2;
    int wave_rows = 1;                  // wave-per-query level before the thread-per-query ones
    DevBuf wscr;                        // its per-wave scratch
    int wave_cap = 0;                   // waves the scratch holds

    // workspace
    nat::Caps caps[kLevels];
    int nthr[kLevels] = {0, 0, 0, 0};  // [0]: locate kernel threads; the most per level
    int scap[kLevels] = {0, 0, 0, 0};  // threads the level's scratch is laid out for
    int nthr_geom = 0;
    int geom_rows_blocks = 0, geom_rows_blocks_mixed = 0;
    int grad_blocks = 0;               // resident blocks of the gradient geometry kernel
    int topo_blocks = 0;
    DevBuf scratch[kLevels];
    DevBuf start;                      // start tet per query
    DevBuf rec, rec_off;               // topology records
    long long rec_cap = 0;
    double units_per_row = 40.0;
    DevBuf q;                    // host-path query upload
    DevBuf order, sortw;         // query order and its sort's workspace
    DevBuf geom_order;           // work order of the forward geometry kernel
    DevBuf scan;                 // int64 scan scratch
    DevBuf row_off, row_cnt, rowptr64, row_total;
    DevBuf ovf[2];
    DevBuf counters;             // [0..7] unsigned: work counters, ovf counts; [8] u64 pool_used
    DevBuf pool_vals, pool_cols;
    long long pool_cap = 0;
    DevBuf csr_vals, csr_cols, csr_rowptr;  // host-path CSR staging
    double nnz_per_row = 24.0;

    int last_M = -1;
    // one-shot caller buffer for the next compute's topology records
    // (nat_topology_target) and the buffer the last compute wrote them to
    char *ext_topo = nullptr;
    long long ext_capacity = 0;
    int ext_M = -1;
    char *ext_written = nullptr;
    long long last_nnz = 0;
    // test hook: entries added to the row total the overflow checks see
    // (NATINTERP3D_TEST_FORGE_NNZ, read by each compute); the rows are untouched
    long long nnz_forge = 0;
    long long reported_nnz() const { return last_nnz + nnz_forge; }
    long long last_units = 0;  // record units of the last compute

    // row arithmetic: 0 = FP64, bit-identical to the reference; 1 = mixed
    int precision = 0;
    float mixed_tau = 0.02f;
    float mixed_kappa = 20.0f;
    // error statement of mixed rows: every mixed_check-th query again in
    // FP64 (0 = off); chk holds the sampled rows' items and saved rows
    int mixed_check = 1024;
    DevBuf chk;

    // optional kernel timing (HIP events on the launch stream)
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    nat_timings times{};

    int64_t device_bytes() const {
        return static_cast<int64_t>(verts.bytes + tet_v.bytes + tet_n.bytes + fat.bytes +
                                    sph.bytes + kd.bytes +
                                    cell_start.bytes + cell_pts.bytes + lp.bytes +
                                    old_vol.bytes + vorig.bytes);
    }
    int problems() const { return static_cast<int>(parts.size()); }
    int64_t workspace_bytes() const {
        size_t s = q.bytes + order.bytes + geom_order.bytes + sortw.bytes + scan.bytes +
                   prob.bytes + qoff.bytes +
                   row_off.bytes + row_cnt.bytes + rowptr64.bytes + row_total.bytes +
                   ovf[0].bytes +
                   ovf[1].bytes + counters.bytes + pool_vals.bytes + pool_cols.bytes +
                   csr_vals.bytes + csr_cols.bytes + csr_rowptr.bytes +
                   rec.bytes + rec_off.bytes + start.bytes + wscr.bytes + itab.bytes +
                   vals_dev.bytes +
                   vals_fwd.bytes + chk.bytes;
        for (int l = 0; l < kLevels; l++) s += scratch[l].bytes;
        return static_cast<int64_t>(s);
    }
    void release_all() {
        DevBuf *all[] = {&verts,    &tet_v,     &tet_n,     &fat,      &sph,      &kd,
                         &cell_start, &cell_pts,
                         &q,        &order,     &geom_order, &sortw,   &scan,
                         &row_off,  &row_cnt,   &rowptr64,  &row_total, &ovf[0],    &ovf[1],
                         &counters, &pool_vals, &pool_cols, &csr_vals, &csr_cols,   &csr_rowptr,
                         &rec,      &rec_off,   &start,     &lp,       &prob,       &qoff,
                         &route,    &wscr,      &itab,      &old_vol,   &vorig,    &vals_dev,
                         &vals_fwd, &chk};
        for (DevBuf *b : all) b->release();
        for (int l = 0; l < kLevels; l++) scratch[l].release();
        for (hipEvent_t &e : ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
};

namespace {

void mark(nat_mesh *m, int i, hipStream_t st) {
    if (m->timing) (void)hipEventRecord(m->ev[i], st);
}
double elapsed(nat_mesh *m, int a, int b) {
    if (!m->timing) return 0.0;
    float ms = 0.f;
    if (hipEventSynchronize(m->ev[b]) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, m->ev[a], m->ev[b]) != hipSuccess) return -1.0;
    return ms;
}

void set_levels(nat_mesh *m) {
    const int kd = m->kd_depth_max + 2;
    const int cus = std::max(1, m->num_cus);
    // level 0 (locate, topology and geometry kernels) sees every query; the
    // topology kernel's LDS capacities cover all but a handful of rows of
    // uniform data.  Level-0 scratch only holds the k-d fallback stacks of
    // the locate kernel.  Persistent grids: exactly the resident blocks.
    m->caps[0] = {kd, 1, 1, 1, 1, 1, 1};
    m->nthr[0] = cus * nat::locate_blocks_per_cu() * 256;
    m->topo_blocks = cus * nat::topo_blocks_per_cu();
    m->nthr_geom = cus * nat::geom_blocks_per_cu() * nat::kGeomThreads;
    m->geom_rows_blocks = cus * nat::geom_rows_blocks_per_cu(false);
    m->geom_rows_blocks_mixed = cus * nat::geom_rows_blocks_per_cu(true);
    m->grad_blocks = cus * nat::grad_geom_blocks_per_cu();
    // larger levels (fused kernel) only see the queries that overflowed; a
    // level's scratch holds as many threads as it has queries, up to these
    // counts (level 1 at most ~31 GB, level 2 ~7 GB, level 3 ~2 GB).  Level 1
    // takes rows of up to 4096 cavity tets and 2048 neighbours (hollow shells
    // of 20k keys: up to 1951 and 821) on 2 waves per SIMD: these rows wait on
    // memory, so its speed follows its thread count
    m->caps[1] = {kd, 4096, 13, 4096, 8192, 2048, 12};
    m->nthr[1] = 131072;
    m->caps[2] = {kd, 65536, 17, 32768, 81920, 32768, 16};
    m->nthr[2] = 2048;
    m->caps[3] = {kd, 524288, 20, 262144, 655360, 393216, 20};
    m->nthr[3] = 64;
}

// Uniform grid over the tet centroids (the k-d tree's points) for the exact
// nearest-centroid search: the data bounding box padded by 0.1%, about two
// centroids per cell; centroids outside the box are filed into the nearest
// edge cell.
void build_grid(const nat::PackedMesh &h, double g0[3], double &gh, int gn[3],
                std::vector<int> &start, std::vector<double4> &pts) {
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) lo[k] = hi[k] = h.verts[k];
    for (int i = 0; i < h.n_data; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], h.verts[3 * i + k]);
            hi[k] = std::max(hi[k], h.verts[3 * i + k]);
        }
    double ext = 0;
    for (int k = 0; k < 3; k++) ext = std::max(ext, hi[k] - lo[k]);
    if (!(ext > 0)) ext = 1.0;
    double vol = 1.0;
    for (int k = 0; k < 3; k++) {
        const double pad = 1e-3 * ext;
        lo[k] -= pad;
        hi[k] += pad;
        vol *= std::max(hi[k] - lo[k], 1e-3 * ext);
    }
    const int T = h.n_tets;
    double target = std::max(1.0, std::min(T / 2.0, 1.0 * (1 << 24)));
    gh = std::cbrt(vol / target);
    for (;;) {
        long long total = 1;
        for (int k = 0; k < 3; k++) {
            gn[k] = std::max(1, static_cast<int>(std::ceil((hi[k] - lo[k]) / gh)));
            total *= gn[k];
        }
        if (total <= (1 << 25) && gn[0] <= 4096 && gn[1] <= 4096 && gn[2] <= 4096) break;
        gh *= 1.25;
    }
    for (int k = 0; k < 3; k++) g0[k] = lo[k];
    const double inv = 1.0 / gh;
    const long long ncell = static_cast<long long>(gn[0]) * gn[1] * gn[2];
    std::vector<int> cell(T);
    start.assign(ncell + 1, 0);
    for (int t = 0; t < T; t++) {
        int c[3];
        for (int k = 0; k < 3; k++) {
            const double f = (h.kd_pos[3 * t + k] - g0[k]) * inv;
            int i = f >= 0.0 ? (f < gn[k] ? static_cast<int>(f) : gn[k] - 1) : 0;
            if (!(f == f)) i = 0;
            c[k] = i;
        }
        cell[t] = (c[2] * gn[1] + c[1]) * gn[0] + c[0];
        start[cell[t] + 1]++;
    }
    for (long long c = 0; c < ncell; c++) start[c + 1] += start[c];
    std::vector<int> fill(start.begin(), start.end() - 1);
    pts.resize(std::max(T, 1));
    for (int t = 0; t < T; t++) {
        const long long id = t;
        double w;
        std::memcpy(&w, &id, sizeof(w));
        pts[fill[cell[t]]++] =
            make_double4(h.kd_pos[3 * t], h.kd_pos[3 * t + 1], h.kd_pos[3 * t + 2], w);
    }
}

// 10-bit coordinate spread to every third bit (x -> bits 0, 3, 6, ...).
unsigned spread3(unsigned x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// perm[new] = old: the n points pts[3i..3i+2] in ascending 30-bit Morton
// code over the box lo..hi (10 bits per axis; points outside are clamped
// onto it), equal codes in index order -- a stable three-pass radix sort.
void morton_perm(const double *pts, int n, const double lo[3], const double hi[3],
                 std::vector<int> &perm) {
    std::vector<unsigned> code(n);
    double scale[3];
    for (int k = 0; k < 3; k++) scale[k] = hi[k] > lo[k] ? 1024.0 / (hi[k] - lo[k]) : 0.0;
    for (int i = 0; i < n; i++) {
        unsigned c = 0;
        for (int k = 0; k < 3; k++) {
            const double f = (pts[3 * i + k] - lo[k]) * scale[k];
            const unsigned u = f >= 1023.0 ? 1023u : (f > 0.0 ? static_cast<unsigned>(f) : 0u);
            c |= spread3(u) << k;
        }
        code[i] = c;
    }
    perm.resize(n);
    std::vector<int> tmp(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int pass = 0; pass < 3; pass++) {
        const int sh = 10 * pass;
        std::vector<int> cnt(1025, 0);
        for (int i = 0; i < n; i++) cnt[((code[perm[i]] >> sh) & 1023u) + 1]++;
        for (int b = 0; b < 1024; b++) cnt[b + 1] += cnt[b];
        for (int i = 0; i < n; i++) tmp[cnt[(code[perm[i]] >> sh) & 1023u]++] = perm[i];
        perm.swap(tmp);
    }
}

// Offsets of the parts in the global numbering.
void index_parts(nat_mesh *m) {
    const int P = m->problems();
    m->vbase.assign(P + 1, 0);
    m->tbase.assign(P + 1, 0);
    m->kd_depth_max = 0;
    for (int p = 0; p < P; p++) {
        m->vbase[p + 1] = m->vbase[p] + m->parts[p].n_data;
        m->tbase[p + 1] = m->tbase[p] + m->parts[p].n_tets;
        m->kd_depth_max = std::max(m->kd_depth_max, m->parts[p].kd_depth);
    }
}

int upload_mesh(nat_mesh *m);

// Most tets the group kernel's LDS tables hold (its default budget).
constexpr int kGroupBudget = 96;

int upload_route(nat_mesh *m) {
    std::vector<unsigned short> r(m->budget.size());
    for (size_t p = 0; p < r.size(); p++)
        r[p] = static_cast<unsigned short>(m->budget[p] | (m->first_level << 8));
    HIP_OR_FAIL(hipMemcpy(m->route.p, r.data(), sizeof(unsigned short) * r.size(),
                          hipMemcpyHostToDevice));
    return 0;
}

// Device set-up of a built mesh (any number of parts).
int finish_build(nat_mesh *m) {
    index_parts(m);
    if (m->device == -1) return 0;  // host-only mesh: export and inspection, no queries
    DeviceGuard g(m->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, m->device) != hipSuccess)
        return fail(-10, "hipGetDeviceProperties failed");
    m->num_cus = prop.multiProcessorCount;
    set_levels(m);
    const int urc = upload_mesh(m);
    if (urc) m->release_all();
    return urc;
}

int check_device(int device) {
    if (device >= 0) {
        int ndev = 0;
        HIP_OR_FAIL(hipGetDeviceCount(&ndev));
        if (device >= ndev) return fail(-1, "device %d out of range (%d)", device, ndev);
    } else if (device != -1) {
        return fail(-1, "device %d invalid", device);
    }
    return 0;
}

int builder_error(int rc) {
    if (rc == -2) return fail(-2, "mesh builder met a tet of non-positive orientation");
    if (rc == -3) return fail(-3, "mesh builder could not locate a point");
    return fail(rc, "mesh builder failed (%d)", rc);
}

// Uploads the parts as one mesh with global numbering: data vertices of all
// problems first (problem p's from vbase[p]), then the four super vertices of
// each problem; tets and k-d nodes from tbase[p]; per-problem grid blocks.
// Within a problem, data vertices and tets are renumbered in Morton order of
// position and centroid over the data bounding box; every record keeps its
// slot order (vertex slots, adjacency slots, k-d children), so the searches
// and the arithmetic see the same mesh.  vorig maps each device vertex back
// to its global key index (super vertices map to themselves).
int upload_mesh(nat_mesh *m) {
    const int P = m->problems();
    const long long NT = m->vbase[P], TT = m->tbase[P];
    if (NT + 4LL * P > 2147483647LL || TT > 2147483647LL)
        return fail(-1, "batch too large for 32-bit vertex and tet indices");
    const size_t nv = static_cast<size_t>(NT + 4LL * P), T = static_cast<size_t>(TT);
    std::vector<double4> verts(nv);
    std::vector<int4> tv(std::max<size_t>(T, 1)), tn(std::max<size_t>(T, 1));
    std::vector<double4> kd(std::max<size_t>(T, 1));
    std::vector<int> start_all;
    std::vector<double4> pts_all(std::max<size_t>(T, 1));
    std::vector<int> vorig(nv);
    std::vector<int> vperm, vnew, tperm, tnew;
    m->lph.assign(P, nat::LocParams{});
    for (int p = 0; p < P; p++) {
        const nat::PackedMesh &h = m->parts[p];
        const int N = h.n_data;
        const int vb = static_cast<int>(m->vbase[p]), tb = static_cast<int>(m->tbase[p]);
        const int sb = static_cast<int>(NT) + 4 * p;
        nat::LocParams &L = m->lph[p];
        // bounding box of the data points: the Morton grid of the renumbering
        // and of the queries
        for (int k = 0; k < 3; k++) L.bbox[k] = L.bbox[3 + k] = h.verts[k];
        for (int i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) {
                L.bbox[k] = std::min(L.bbox[k], h.verts[3 * i + k]);
                L.bbox[3 + k] = std::max(L.bbox[3 + k], h.verts[3 * i + k]);
            }
        morton_perm(h.verts.data(), N, L.bbox, L.bbox + 3, vperm);
        morton_perm(h.kd_pos.data(), h.n_tets, L.bbox, L.bbox + 3, tperm);
        vnew.assign(N, 0);
        for (int i = 0; i < N; i++) vnew[vperm[i]] = i;
        tnew.assign(h.n_tets, 0);
        for (int t = 0; t < h.n_tets; t++) tnew[tperm[t]] = t;
        auto gv = [&](int v) { return v < N ? vb + vnew[v] : sb + (v - N); };
        auto gt = [&](int x) { return x < 0 ? -1 : tb + tnew[x]; };
        for (int i = 0; i < N + 4; i++) {
            verts[gv(i)] =
                make_double4(h.verts[3 * i], h.verts[3 * i + 1], h.verts[3 * i + 2], 0.0);
            vorig[gv(i)] = i < N ? vb + i : sb + (i - N);
        }
        for (int nt = 0; nt < h.n_tets; nt++) {
            const int t = tperm[nt];
            const int *pv = &h.tet_v[4 * t];
            auto pt = [&](int v) {
                return nat::d3{h.verts[3 * v], h.verts[3 * v + 1], h.verts[3 * v + 2]};
            };
            // orientation bit (natural_insertionfree.c:457-476), stored inverted
            // in bit 31 of p3 so that the common positive case reads unchanged
            const bool pos = nat::orient3d(pt(pv[0]), pt(pv[1]), pt(pv[2]), pt(pv[3])) > 0;
            const int w = gv(pv[3]);
            tv[tb + nt] = make_int4(gv(pv[0]), gv(pv[1]), gv(pv[2]),
                                    pos ? w : (w | static_cast<int>(0x80000000u)));
            tn[tb + nt] = make_int4(gt(h.tet_n[4 * t]), gt(h.tet_n[4 * t + 1]),
                                    gt(h.tet_n[4 * t + 2]), gt(h.tet_n[4 * t + 3]));
            int2 lr = make_int2(gt(h.kd_left[t]), gt(h.kd_right[t]));
            double wd;
            std::memcpy(&wd, &lr, sizeof(wd));
            kd[tb + nt] =
                make_double4(h.kd_pos[3 * t], h.kd_pos[3 * t + 1], h.kd_pos[3 * t + 2], wd);
        }
        std::vector<int> start;
        std::vector<double4> pts;
        build_grid(h, L.g0, L.gh, L.gn, start, pts);
        L.ginv = 1.0 / L.gh;
        L.cell_base = static_cast<int>(start_all.size());
        for (int v : start) start_all.push_back(v + tb);
        for (int t = 0; t < h.n_tets; t++) {
            long long id;
            std::memcpy(&id, &pts[t].w, sizeof(id));
            id = gt(static_cast<int>(id));
            double wd;
            std::memcpy(&wd, &id, sizeof(wd));
            pts_all[tb + t] = make_double4(pts[t].x, pts[t].y, pts[t].z, wd);
        }
        L.kd_root = h.n_tets > 0 ? gt(0) : -1;
        for (int k = 0; k < 3; k++) {
            L.rmin[k] = h.kd_rect[k];
            L.rmax[k] = h.kd_rect[3 + k];
        }
        if (start_all.size() > 2147483647ULL) return fail(-1, "batch grid too large");
    }
    HIP_OR_FAIL(m->verts.ensure(sizeof(double4) * nv));
    HIP_OR_FAIL(m->tet_v.ensure(sizeof(int4) * tv.size()));
    HIP_OR_FAIL(m->tet_n.ensure(sizeof(int4) * tn.size()));
    HIP_OR_FAIL(m->kd.ensure(sizeof(double4) * kd.size()));
    HIP_OR_FAIL(m->fat.ensure(sizeof(nat::FatTet) * tv.size()));
    HIP_OR_FAIL(m->cell_start.ensure(sizeof(int) * start_all.size()));
    HIP_OR_FAIL(m->cell_pts.ensure(sizeof(double4) * pts_all.size()));
    HIP_OR_FAIL(m->lp.ensure(sizeof(nat::LocParams) * P));
    HIP_OR_FAIL(m->vorig.ensure(sizeof(int) * nv));
    HIP_OR_FAIL(hipMemcpy(m->vorig.p, vorig.data(), sizeof(int) * nv, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->verts.p, verts.data(), sizeof(double4) * nv, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->tet_v.p, tv.data(), sizeof(int4) * tv.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->tet_n.p, tn.data(), sizeof(int4) * tn.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->kd.p, kd.data(), sizeof(double4) * kd.size(), hipMemcpyHostToDevice));
    {
        // the tets with their vertex coordinates inline, in device numbering
        std::vector<nat::FatTet> ft(tv.size());
        for (size_t t = 0; t < tv.size(); t++) {
            const int4 v = tv[t];
            const int s[4] = {v.x, v.y, v.z, v.w & 0x7fffffff};
            double c[12];
            for (int k = 0; k < 4; k++) {
                const double4 p = T ? verts[s[k]] : make_double4(0, 0, 0, 0);
                c[3 * k] = p.x;
                c[3 * k + 1] = p.y;
                c[3 * k + 2] = p.z;
            }
            ft[t].c0 = make_double4(c[0], c[1], c[2], c[3]);
            ft[t].c1 = make_double4(c[4], c[5], c[6], c[7]);
            ft[t].c2 = make_double4(c[8], c[9], c[10], c[11]);
            ft[t].v = v;
            ft[t].n = tn[t];
        }
        HIP_OR_FAIL(hipMemcpy(m->fat.p, ft.data(), sizeof(nat::FatTet) * ft.size(),
                              hipMemcpyHostToDevice));
    }
    HIP_OR_FAIL(hipMemcpy(m->cell_start.p, start_all.data(), sizeof(int) * start_all.size(),
                          hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->cell_pts.p, pts_all.data(), sizeof(double4) * pts_all.size(),
                          hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(m->lp.p, m->lph.data(), sizeof(nat::LocParams) * P,
                          hipMemcpyHostToDevice));
    nat::DevMesh &d = m->dev;
    d.verts = m->verts.as<double4>();
    d.tet_v = m->tet_v.as<int4>();
    d.tet_n = m->tet_n.as<int4>();
    d.fat = m->fat.as<nat::FatTet>();
    d.sph = nullptr;
    d.kd = m->kd.as<double4>();
    d.n_data = static_cast<int>(NT);
    d.n_tets = static_cast<int>(TT);
    d.cell_start = m->cell_start.as<int>();
    d.cell_pts = m->cell_pts.as<double4>();
    d.lp = m->lp.as<nat::LocParams>();
    d.prob = nullptr;
    // the cavity search's circumspheres, from the fat tets (64 B per tet)
    HIP_OR_FAIL(m->sph.ensure(64 * std::max<size_t>(1, tv.size())));
    HIP_OR_FAIL(nat::launch_sphere(d, m->sph.as<nat::SphereTet>(), nullptr));
    HIP_OR_FAIL(hipStreamSynchronize(nullptr));
    d.sph = m->sph.as<nat::SphereTet>();
    HIP_OR_FAIL(m->route.ensure(sizeof(unsigned short) * P));
    d.route = m->route.as<unsigned short>();
    m->budget.assign(P, static_cast<unsigned char>(kGroupBudget));
    return upload_route(m);
}

// Per-thread scratch of a level for a launch of `want` threads: the level's
// thread count is capped by the work present, so small calls hold little
// memory.  The layout (stride = capacity in threads) only changes when the
// capacity grows, and then the scratch is cleared: the per-thread stamps of
// the hash tables stay consistent.  Returns the threads to launch; the carve
// stride is m->scap[level].
int ensure_level_scratch(nat_mesh *m, int level, long long want, int *launch) {
    const int cap_max = m->nthr[level];
    const int n = static_cast<int>(std::min<long long>(cap_max, std::max<long long>(want, 1)));
    if (n > m->scap[level]) {
        int cap = std::max(n, 2 * m->scap[level]);
        cap = std::min(cap_max, (cap + 255) / 256 * 256);
        // each of the (at most 16) arrays is rounded up to 256 bytes when carved
        const size_t per = nat::scratch_bytes_per_thread(m->caps[level]);
        const size_t need = per * static_cast<size_t>(cap) + 16 * 256;
        m->scratch[level].release();
        HIP_OR_FAIL(m->scratch[level].ensure(need, /*zero=*/true));
        m->scap[level] = cap;
    }
    *launch = n;
    return 0;
}

// Scratch of the wave-per-query level for `want` queries: one wave per
// query up to the resident waves of the device.  Returns the waves to launch.
int ensure_wave_scratch(nat_mesh *m, long long want, int *launch) {
    const int cap_max = std::max(1, m->num_cus) * nat::wave_blocks_per_cu();
    const int n = static_cast<int>(std::min<long long>(cap_max, std::max<long long>(want, 1)));
    if (n > m->wave_cap) {
        const int cap = std::min(cap_max, std::max(n, 2 * m->wave_cap));
        m->wscr.release();
        HIP_OR_FAIL(m->wscr.ensure(nat::wave_scratch_size() * static_cast<size_t>(cap)));
        m->wave_cap = cap;
    }
    *launch = n;
    return 0;
}

// d_prob: problem of each query (batch meshes), NULL for a single problem
// The step-3 table (64 doubles per tet), built once per mesh on its first
// compute; without the memory for it the geometry is computed inline.
// Grid of the two-lanes-per-row geometry kernel of the mesh's precision.
int geom_rows_grid(const nat_mesh *m) {
    return (m->precision == 1 ? m->geom_rows_blocks_mixed : m->geom_rows_blocks) * m->geom_grid;
}

int ensure_old_vol(nat_mesh *m, hipStream_t st) {
    if (m->old_vol_tried) return 0;
    m->old_vol_tried = true;
    const size_t bytes = sizeof(double) * 64 * static_cast<size_t>(m->dev.n_tets);
    if (m->old_vol.ensure(bytes) != hipSuccess) {
        (void)hipGetLastError();
        m->old_vol.release();
        return 0;
    }
    nat::DevMesh d = m->dev;
    d.old_vol = nullptr;
    // the kernel also puts each tet's table face mask into its circumsphere
    // record (sphere_dm), where the search hands it on to the order kernel
    HIP_OR_FAIL(nat::launch_old_vol(d, m->old_vol.as<double>(), m->sph.as<nat::SphereTet>(), st));
    m->dev.old_vol = m->old_vol.as<double>();
    return 0;
}

size_t topo_rec_off(long long M);

// Entries the overflow checks add to a compute's true total: the value of
// NATINTERP3D_TEST_FORGE_NNZ (a non-negative integer), else 0.  It lets a
// test take the nnz > INT32_MAX branches of the boundary without 2^31
// entries; the rows themselves are never changed.
long long forged_entries() {
    const char *s = std::getenv("NATINTERP3D_TEST_FORGE_NNZ");
    if (!s || !*s) return 0;
    char *end = nullptr;
    const long long v = std::strtoll(s, &end, 10);
    return (end && *end == '\0' && v > 0) ? v : 0;
}

// Error statement of a mixed compute whose rows are all in the pool: the
// sampled rows (every m->mixed_check-th query of the processing order) are
// made again by the FP64 geometry kernel from the same records, behind the
// compute's entries in the pool, compared with the mixed rows, and the mixed
// rows' offsets are put back.
int mixed_check(nat_mesh *m, const double *d_q, int M, int n_check, const nat::Topo &topo,
                nat::RowOut out, unsigned *ctr, hipStream_t st) {
    const size_t n = static_cast<size_t>(n_check);
    const size_t a = (n + 63) & ~static_cast<size_t>(63);
    HIP_OR_FAIL(m->chk.ensure(a * (4 + 4 + 8 + 8)));
    int *items = m->chk.as<int>();
    int *s_cnt = items + a;
    long long *s_off = reinterpret_cast<long long *>(s_cnt + a);
    double *s_total = reinterpret_cast<double *>(s_off + a);
    unsigned long long *res = reinterpret_cast<unsigned long long *>(ctr + 54);
    HIP_OR_FAIL(nat::launch_zero_u32(ctr + 40, 8, st));
    HIP_OR_FAIL(nat::launch_zero_u32(res, 6, st));
    HIP_OR_FAIL(nat::launch_mixed_check_save(m->order.as<int>(), n_check, m->mixed_check,
                                             out.row_off, out.row_cnt, out.row_total, items,
                                             s_off, s_cnt, s_total, st));
    nat::Topo t = topo;
    t.rec_used = nullptr;
    HIP_OR_FAIL(nat::launch_geom(m->dev, m->nthr_geom / nat::kGeomThreads * m->geom_grid,
                                 m->geom_lanes == 2 ? m->geom_rows_blocks * m->geom_grid : 0,
                                 d_q, items, n_check, ctr + 40, t, out, 0.0f, 0.0f, nullptr,
                                 st));
    HIP_OR_FAIL(nat::launch_mixed_check_compare(n_check, items, topo.rec_off, out.row_off,
                                                out.row_cnt, out.row_total, s_off, s_cnt,
                                                s_total, out.pool_vals, out.pool_cols,
                                                out.pool_cap, res, st));
    unsigned long long h[3] = {0, 0, 0};
    HIP_OR_FAIL(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    double err;
    std::memcpy(&err, &h[0], sizeof(err));
    m->times.mixed_err_max = err;
    m->times.mixed_err_rows = static_cast<int64_t>(h[1]);
    m->times.mixed_err_entries = static_cast<int64_t>(h[2]);
    (void)M;
    return 0;
}

int compute_rows(nat_mesh *m, const double *d_q, int M, const int *d_prob, hipStream_t st,
                 long long *nnz_out) {
    m->last_M = -1;
    m->nnz_forge = forged_entries();
    // the records go straight into a caller's saved-topology buffer when one
    // was set for this many queries and they fit; else into the workspace
    char *ext_base = m->ext_topo && m->ext_M == M ? m->ext_topo : nullptr;
    // a topology will be saved: the wave-per-query rows keep big records too
    const bool save_big = ext_base != nullptr;
    long long ext_units = 0;
    if (ext_base) {
        ext_units = (m->ext_capacity - static_cast<long long>(topo_rec_off(M))) /
                    static_cast<long long>(sizeof(int4));
        if (ext_units < 1) ext_base = nullptr;
    }
    m->ext_topo = nullptr;
    m->ext_written = nullptr;
    if (M > 0)
        if (int rc = ensure_old_vol(m, st)) return rc;
    m->dev.prob = d_prob;
    if (M == 0) {
        m->last_M = 0;
        m->last_nnz = 0;
        *nnz_out = m->reported_nnz();
        return 0;
    }
    const size_t Mz = static_cast<size_t>(M);
    // Morton bucketing: per problem, 3*bits-bit cells over its data bounding box
    const int P = m->problems();
    int bits = 1;
    while (bits < 7 && (static_cast<unsigned long long>(P) << (3 * (bits + 1))) <= Mz * 2 &&
           (static_cast<unsigned long long>(P) << (3 * (bits + 1))) <= (1ull << 24))
        bits++;
    if ((static_cast<unsigned long long>(P) << (3 * bits)) > (1ull << 26))
        return fail(-1, "too many problems in one batch (%d)", P);
    const size_t scan_elems = nat::scan_tmp_elems(M) + Mz + 1;
    HIP_OR_FAIL(m->order.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->geom_order.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->sortw.ensure(nat::morton_sort_bytes(M)));
    HIP_OR_FAIL(m->scan.ensure(sizeof(long long) * scan_elems));
    HIP_OR_FAIL(m->row_off.ensure(sizeof(long long) * Mz));
    HIP_OR_FAIL(m->row_cnt.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->row_total.ensure(sizeof(double) * Mz));
    HIP_OR_FAIL(m->rowptr64.ensure(sizeof(long long) * (Mz + 1)));
    HIP_OR_FAIL(m->ovf[0].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->ovf[1].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->counters.ensure(256));
    m->dev.robust_ctr = reinterpret_cast<unsigned long long *>(m->counters.as<unsigned>() + 48);
    int nthr0 = 0;
    if (int rc = ensure_level_scratch(m, 0, M, &nthr0)) return rc;
    // mixed rows' error statement: the sampled rows are made again in FP64
    // behind the compute's own entries in the pool
    const int n_check = m->precision == 1 && m->mixed_check > 0
                            ? static_cast<int>((Mz + m->mixed_check - 1) / m->mixed_check)
                            : 0;

    m->times = nat_timings{};
    mark(m, 0, st);
    HIP_OR_FAIL(nat::launch_morton_sort(d_q, M, m->lp.as<nat::LocParams>(), d_prob, P, bits,
                                        m->order.as<int>(), m->sortw.p, st));
    mark(m, 1, st);
    m->times.order_ms = elapsed(m, 0, 1);

    HIP_OR_FAIL(m->rec_off.ensure(sizeof(long long) * Mz));
    HIP_OR_FAIL(m->start.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->itab.ensure(sizeof(int4) * Mz));
    const long long check_room =
        n_check ? static_cast<long long>(std::ceil(2.0 * m->nnz_per_row * n_check)) + 1024 : 0;
    long long want = static_cast<long long>(std::ceil(m->nnz_per_row * M)) + 1024 + check_room;
    long long want_rec = static_cast<long long>(std::ceil(m->units_per_row * M)) + 1024;
    for (int attempt = 0; attempt < 4; attempt++) {
        m->times.levels = 0;
        if (m->pool_cap < want) {
            HIP_OR_FAIL(m->pool_vals.ensure(sizeof(double) * want));
            HIP_OR_FAIL(m->pool_cols.ensure(sizeof(int) * want));
            m->pool_cap = want;
        }
        if (!ext_base && m->rec_cap < want_rec) {
            HIP_OR_FAIL(m->rec.ensure(sizeof(int4) * want_rec));
            m->rec_cap = want_rec;
        }
        // counters: [0..3] fused-level work, [6] wave-level work, [8..11]
        // overflow counts, [12] wave-level overflow, u64 pool use at [16],
        // records at [18], mixed-path FP32 / FP64 counts at [20], [22], rows
        // redone at [24]; per-XCD work parts of the topology kernel at
        // [32..39], of the geometry kernel at [40..47]; u64 robust-predicate
        // counts at [48], [50], [52]; u64 mixed error statement at [54..59]
        unsigned *ctr = m->counters.as<unsigned>();
        unsigned long long *pool_used = reinterpret_cast<unsigned long long *>(ctr + 16);
        unsigned long long *rec_used = reinterpret_cast<unsigned long long *>(ctr + 18);
        HIP_OR_FAIL(nat::launch_zero_u32(m->counters.p, 64, st));
        nat::RowOut out;
        out.pool_vals = m->pool_vals.as<double>();
        out.pool_cols = m->pool_cols.as<int>();
        out.pool_cap = m->pool_cap;
        out.pool_used = pool_used;
        out.row_off = m->row_off.as<long long>();
        out.row_cnt = m->row_cnt.as<int>();
        out.row_total = m->row_total.as<double>();

        // level 0: start tets, topology records, then rows from them
        nat::Scratch sc0;
        nat::scratch_carve(m->scratch[0].p, m->caps[0], m->scap[0], &sc0);
        nat::Topo topo;
        topo.rec = ext_base ? reinterpret_cast<int4 *>(ext_base + topo_rec_off(M))
                            : m->rec.as<int4>();
        topo.rec_cap = ext_base ? ext_units : m->rec_cap;
        topo.rec_used = rec_used;
        topo.rec_off = m->rec_off.as<long long>();
        topo.itab = m->itab.as<int4>();
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_locate(m->dev, m->caps[0], sc0, (nthr0 + 255) / 256, d_q,
                                       m->order.as<int>(), M, m->start.as<int>(), st));
        mark(m, 3, st);
        m->times.locate_ms = elapsed(m, 2, 3);
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_topo_search(m->dev, m->topo_blocks * m->topo_grid, d_q,
                                            m->order.as<int>(), M, m->start.as<int>(), ctr + 32,
                                            topo, ctr + 8, m->ovf[0].as<int>(), st));
        mark(m, 1, st);
        HIP_OR_FAIL(nat::launch_topo_order(m->dev, m->order.as<int>(), M, topo, ctr + 8,
                                           m->ovf[0].as<int>(), st));
        mark(m, 3, st);
        unsigned long long meta[2] = {0, 0};
        HIP_OR_FAIL(hipMemcpyAsync(meta, ctr + 16, sizeof(meta), hipMemcpyDeviceToHost, st));
        unsigned novf = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&novf, ctr + 8, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        m->times.topo_ms = elapsed(m, 2, 3);
        m->times.topo_search_ms = elapsed(m, 2, 1);
        m->times.topo_order_ms = elapsed(m, 1, 3);
        const long long units = static_cast<long long>(meta[1]);
        if (units > topo.rec_cap) {
            want_rec = units + units / 8 + 1024;
            ext_base = nullptr;  // the caller's buffer is too small: the workspace
            continue;
        }
        m->units_per_row = std::max(8.0, 1.125 * static_cast<double>(units) / M);
        m->last_units = units;
        mark(m, 2, st);
        if (m->work_sort)
            HIP_OR_FAIL(nat::launch_work_sort(m->order.as<int>(), M, topo,
                                              m->geom_order.as<int>(), st));
        HIP_OR_FAIL(nat::launch_geom(
            m->dev, m->nthr_geom / nat::kGeomThreads * m->geom_grid,
            m->geom_lanes == 2 ? geom_rows_grid(m) : 0, d_q,
            (m->work_sort ? m->geom_order : m->order).as<int>(), M, ctr + 40,
            topo, out, m->mixed_tau, m->mixed_kappa,
            m->precision == 1 ? reinterpret_cast<unsigned long long *>(ctr + 20) : nullptr, st));
        mark(m, 3, st);
        m->times.geom_ms = elapsed(m, 2, 3);
        m->times.rows_ms[0] = m->times.topo_ms + m->times.geom_ms;
        m->times.rows_items[0] = M;
        m->times.levels = 1;

        // larger rows: the queries that overflowed the previous level, first
        // one wave per query, then one thread per query
        int cur = 0;  // ovf buffer holding the items
        const int *items = m->ovf[0].as<int>();
        int n_items = static_cast<int>(novf);
        m->times.wave_ms = 0.0;
        m->times.wave_items = 0;
        if (n_items > 0 && m->wave_rows) {
            int nw = 0;
            if (int rc = ensure_wave_scratch(m, n_items, &nw)) return rc;
            nat::RowOut wout = out;
            if (save_big) wout.save = topo;
            mark(m, 2, st);
            HIP_OR_FAIL(nat::launch_wave_rows(m->dev, m->wscr.as<char>(), nw, d_q, items, n_items,
                                              nullptr, m->start.as<int>(), ctr + 6, wout,
                                              m->ovf[1].as<int>(), ctr + 12, st));
            mark(m, 3, st);
            unsigned nnext = 0;
            unsigned long long used_rec = 0;
            HIP_OR_FAIL(
                hipMemcpyAsync(&nnext, ctr + 12, sizeof(unsigned), hipMemcpyDeviceToHost, st));
            HIP_OR_FAIL(hipMemcpyAsync(&used_rec, rec_used, sizeof(used_rec),
                                       hipMemcpyDeviceToHost, st));
            HIP_OR_FAIL(hipStreamSynchronize(st));
            m->times.wave_ms = elapsed(m, 2, 3);
            m->times.wave_items = n_items;
            if (save_big) {
                const long long u2 = static_cast<long long>(used_rec);
                if (u2 > topo.rec_cap) {  // the big records did not fit: again, larger
                    want_rec = u2 + u2 / 8 + 1024;
                    ext_base = nullptr;
                    continue;
                }
                m->units_per_row = std::max(8.0, 1.125 * static_cast<double>(u2) / M);
                m->last_units = u2;
            }
            n_items = static_cast<int>(nnext);
            cur = 1;
            items = m->ovf[1].as<int>();
        }
        for (int level = 1; level < kLevels && n_items > 0; level++) {
            int nthr = 0;
            if (int rc = ensure_level_scratch(m, level, n_items, &nthr)) return rc;
            nat::Scratch sc;
            nat::scratch_carve(m->scratch[level].p, m->caps[level], m->scap[level], &sc);
            out.ovf_list = m->ovf[cur ^ 1].as<int>();
            out.ovf_count = ctr + 8 + level;
            const int nblocks = (nthr + 255) / 256;
            mark(m, 2, st);
            HIP_OR_FAIL(nat::launch_rows(m->dev, m->caps[level], sc, level, nblocks, d_q, items,
                                         n_items, nullptr, ctr + level, out, st));
            mark(m, 3, st);
            unsigned nnext = 0;
            HIP_OR_FAIL(hipMemcpyAsync(&nnext, ctr + 8 + level, sizeof(unsigned),
                                       hipMemcpyDeviceToHost, st));
            HIP_OR_FAIL(hipStreamSynchronize(st));
            m->times.rows_ms[level] = elapsed(m, 2, 3);
            m->times.rows_items[level] = n_items;
            m->times.levels = level + 1;
            n_items = static_cast<int>(nnext);
            cur ^= 1;
            items = m->ovf[cur].as<int>();
        }
        if (n_items > 0)
            return fail(-4, "%d queries need more than %d cavity tets", n_items,
                        m->caps[kLevels - 1].cav);
        // pool use, record use, FP32 / FP64 tets and faces and FP64 rows of
        // the mixed path
        unsigned long long tail[5] = {0, 0, 0, 0, 0};
        HIP_OR_FAIL(hipMemcpyAsync(tail, pool_used, sizeof(tail), hipMemcpyDeviceToHost, st));
        unsigned long long rob[3] = {0, 0, 0};
        HIP_OR_FAIL(hipMemcpyAsync(rob, ctr + 48, sizeof(rob), hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        const unsigned long long used = tail[0];
        m->times.mixed_fp32 = static_cast<int64_t>(tail[2]);
        m->times.mixed_fp64 = static_cast<int64_t>(tail[3]);
        m->times.mixed_rows_fp64 = static_cast<int64_t>(tail[4]);
        m->times.robust_exact = static_cast<int64_t>(rob[0]);
        m->times.robust_changed = static_cast<int64_t>(rob[1]);
        m->times.robust_unresolved = static_cast<int64_t>(rob[2]);
        if (static_cast<long long>(used) <= m->pool_cap) {
            if (n_check > 0 && static_cast<long long>(used) + check_room <= m->pool_cap)
                if (int rc = mixed_check(m, d_q, M, n_check, topo, out, ctr, st)) return rc;
            m->last_M = M;
            m->ext_written = ext_base;
            m->last_nnz = static_cast<long long>(used);
            m->nnz_per_row = std::max(4.0, 1.125 * static_cast<double>(used) / M);
            *nnz_out = m->reported_nnz();
            return 0;
        }
        want = static_cast<long long>(used) + 1024 + check_room;
    }
    return fail(-5, "row pool did not converge");
}

// Saved topology: int64 header {magic, M, record units, N, P, 0, 0, 0},
// rec_off int64[M], the query order int32[M], the problem of each query
// int32[M] (filled for batches, P > 1), the start tet of each query int32[M]
// (the forward's locate, so that the backward does not locate the rows
// without a record again), the row totals float64[M], then the records
// (16-byte aligned).
constexpr long long kTopoMagic = 0x4e41545450304c34LL;
constexpr size_t kTopoHeader = 64;
size_t topo_order_off(long long M) { return kTopoHeader + 8 * static_cast<size_t>(M); }
size_t topo_prob_off(long long M) { return topo_order_off(M) + 4 * static_cast<size_t>(M); }
size_t topo_start_off(long long M) { return topo_prob_off(M) + 4 * static_cast<size_t>(M); }
size_t topo_total_off(long long M) {
    return (topo_start_off(M) + 4 * static_cast<size_t>(M) + 7) & ~static_cast<size_t>(7);
}
size_t topo_rec_off(long long M) {
    return (topo_total_off(M) + 8 * static_cast<size_t>(M) + 15) & ~static_cast<size_t>(15);
}

int grad_rows(nat_mesh *m, const char *blob, const double *d_q, int M, const int *d_rowptr,
              const double *d_w, const double *d_gy, const double *d_values, int D,
              double *d_dq, hipStream_t st) {
    long long hdr[8];
    HIP_OR_FAIL(hipMemcpyAsync(hdr, blob, sizeof(hdr), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    if (hdr[0] != kTopoMagic) return fail(-1, "d_topo is not a saved topology");
    if (hdr[1] != M)
        return fail(-1, "saved topology has %lld queries, not %d", hdr[1], M);
    if (hdr[3] != m->dev.n_data || hdr[4] != m->problems())
        return fail(-1, "saved topology is of another mesh");
    // the larger levels locate again: they need each query's problem
    m->dev.prob = hdr[4] > 1 ? reinterpret_cast<const int *>(blob + topo_prob_off(M)) : nullptr;
    const size_t Mz = static_cast<size_t>(M);
    HIP_OR_FAIL(m->ovf[0].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->ovf[1].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->counters.ensure(256));
    // counters: [1..3] fused-level work, [6] wave-level work, [8..11]
    // overflow counts, [12] wave-level overflow, [32..39] per-XCD work parts
    // of the record kernel
    unsigned *ctr = m->counters.as<unsigned>();
    HIP_OR_FAIL(hipMemsetAsync(m->counters.p, 0, 256, st));
    m->dev.robust_ctr = reinterpret_cast<unsigned long long *>(ctr + 48);
    nat::Topo topo;
    topo.rec = reinterpret_cast<int4 *>(const_cast<char *>(blob) + topo_rec_off(M));
    topo.rec_cap = hdr[2];
    topo.rec_used = nullptr;
    topo.rec_off = reinterpret_cast<long long *>(const_cast<char *>(blob) + kTopoHeader);
    const int *order = reinterpret_cast<const int *>(blob + topo_order_off(M));
    // the kernels read values by device vertex index: the caller's rows in
    // that order (data vertices only; super vertices carry no value)
    const int nd = m->dev.n_data;
    HIP_OR_FAIL(m->vals_dev.ensure(sizeof(double) * std::max<size_t>(1, static_cast<size_t>(nd) * D)));
    HIP_OR_FAIL(nat::launch_gather_rows(d_values, m->vorig.as<int>(), nd, D,
                                        m->vals_dev.as<double>(), st));
    nat::GradIn g{d_gy, m->vals_dev.as<double>(), D, d_dq, d_rowptr, d_w,
                  reinterpret_cast<const double *>(blob + topo_total_off(M)), topo.rec,
                  topo.rec_off};
    for (int l = 0; l < kLevels; l++) {
        m->times.grad_ms[l] = 0.0;
        m->times.grad_items[l] = 0;
    }
    mark(m, 2, st);
    HIP_OR_FAIL(nat::launch_grad_geom(m->dev, m->grad_blocks, d_q, order, M,
                                      ctr + 32, topo, g, m->ovf[0].as<int>(), ctr + 8, st));
    mark(m, 3, st);
    unsigned novf = 0;
    HIP_OR_FAIL(hipMemcpyAsync(&novf, ctr + 8, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    m->times.grad_ms[0] = elapsed(m, 2, 3);
    m->times.grad_items[0] = M;
    int cur = 0;
    const int *items = m->ovf[0].as<int>();
    int n_items = static_cast<int>(novf);
    m->times.grad_wave_ms = 0.0;
    m->times.grad_wave_items = 0;
    if (n_items > 0 && m->wave_rows) {
        // one wave per row without a record, from the forward's start tet
        // (the same locate with the same capacities: a k-d overflow, -2, goes
        // on to the thread-per-query levels, which locate with larger ones)
        int nw = 0;
        if (int rc = ensure_wave_scratch(m, n_items, &nw)) return rc;
        const int *start = reinterpret_cast<const int *>(blob + topo_start_off(M));
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_wave_grad(m->dev, m->wscr.as<char>(), nw, d_q, items, n_items,
                                          start, ctr + 6, g, m->ovf[1].as<int>(), ctr + 12, st));
        mark(m, 3, st);
        unsigned nnext = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&nnext, ctr + 12, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        m->times.grad_wave_ms = elapsed(m, 2, 3);
        m->times.grad_wave_items = n_items;
        n_items = static_cast<int>(nnext);
        cur = 1;
        items = m->ovf[1].as<int>();
    }
    for (int level = 1; level < kLevels && n_items > 0; level++) {
        int nthr = 0;
        if (int rc = ensure_level_scratch(m, level, n_items, &nthr)) return rc;
        nat::Scratch sc;
        nat::scratch_carve(m->scratch[level].p, m->caps[level], m->scap[level], &sc);
        const int nblocks = (nthr + 255) / 256;
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_grad_rows(m->dev, m->caps[level], sc, level, nblocks, d_q, items,
                                          n_items, ctr + level, g, m->ovf[cur ^ 1].as<int>(),
                                          ctr + 8 + level, st));
        mark(m, 3, st);
        unsigned nnext = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&nnext, ctr + 8 + level, sizeof(unsigned),
                                   hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        m->times.grad_ms[level] = elapsed(m, 2, 3);
        m->times.grad_items[level] = n_items;
        n_items = static_cast<int>(nnext);
        cur ^= 1;
        items = m->ovf[cur].as<int>();
    }
    if (n_items > 0)
        return fail(-4, "%d queries need more than %d cavity tets", n_items,
                    m->caps[kLevels - 1].cav);
    return 0;
}

int emit_rows(nat_mesh *m, double *d_vals, int *d_cols, int *d_rowptr, hipStream_t st) {
    if (m->last_M < 0) return fail(-6, "nat_weights_emit without a preceding compute");
    if (m->reported_nnz() > 2147483647LL) return -1;
    const int M = m->last_M;
    if (M == 0) {
        HIP_OR_FAIL(hipMemsetAsync(d_rowptr, 0, sizeof(int), st));
        return 0;
    }
    long long *rp64 = m->rowptr64.as<long long>();
    mark(m, 0, st);
    HIP_OR_FAIL(nat::launch_scan_counts(m->row_cnt.as<int>(), rp64, M, m->scan.as<long long>(),
                                        st));
    HIP_OR_FAIL(nat::launch_gather_csr(rp64, m->row_off.as<long long>(), m->row_cnt.as<int>(),
                                       m->pool_vals.as<double>(), m->pool_cols.as<int>(),
                                       m->vorig.as<int>(), M, m->last_nnz, d_vals, d_cols,
                                       d_rowptr, st));
    mark(m, 1, st);
    m->times.emit_ms = elapsed(m, 0, 1);
    return 0;
}

// ---------------------------------------------------------------- graphs
//
// The launch sequence of one weights computation without host round trips:
// capacities fixed by a warm-up call, level kernels reading their item counts
// on the device, counters cleared by kernels.  Overflows land in counters
// that nat_graph_launch checks after the replay.

// Device buffers of the mesh workspace a graph's nodes point to.
std::vector<std::pair<const void *, size_t>> workspace_fingerprint(const nat_mesh *m) {
    std::vector<std::pair<const void *, size_t>> f;
    const DevBuf *bufs[] = {&m->order, &m->geom_order, &m->sortw, &m->scan,
                            &m->row_off, &m->row_cnt, &m->row_total, &m->rowptr64, &m->ovf[0],
                            &m->ovf[1], &m->counters, &m->pool_vals, &m->pool_cols, &m->rec,
                            &m->rec_off, &m->start};
    for (const DevBuf *b : bufs) f.emplace_back(b->p, b->bytes);
    for (int l = 0; l < kLevels; l++) f.emplace_back(m->scratch[l].p, m->scratch[l].bytes);
    f.emplace_back(m->wscr.p, m->wscr.bytes);
    return f;
}

// ---- host transfers of the host-buffer entry point (nat_get_weights)
//
// Pageable hipMemcpy moved the CSR at about 4 GB/s (one staging thread, and
// page faults on the freshly allocated output).  Here kXferLanes host threads
// each move one contiguous slice through two pinned chunks on a stream of
// their own: the DMA of one chunk overlaps the host copy of the previous one,
// and the output pages are first touched by all threads at once.  The pinned
// chunks are allocated on the first large transfer to a device and kept for
// the life of the process.
constexpr int kXferLanes = 8;
constexpr size_t kXferChunk = static_cast<size_t>(16) << 20;
constexpr size_t kXferDirect = static_cast<size_t>(8) << 20;  // smaller: one hipMemcpy
constexpr int kXferMaxDevices = 64;

struct XferLane {
    hipStream_t st = nullptr;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};
struct XferPool {
    std::mutex mu;
    bool ready = false;
    XferLane lane[kXferLanes];
};

XferPool *xfer_pool(int device) {
    static std::mutex mu;
    static XferPool *pools[kXferMaxDevices] = {};
    if (device < 0 || device >= kXferMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!pools[device]) pools[device] = new XferPool;  // lives to the process end
    return pools[device];
}

hipError_t xfer_ready(XferPool &p) {  // the caller holds p.mu, device current
    if (p.ready) return hipSuccess;
    for (XferLane &L : p.lane) {
        hipError_t e = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking);
        for (int k = 0; k < 2 && e == hipSuccess; k++) {
            e = hipHostMalloc(&L.buf[k], kXferChunk, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&L.ev[k], hipEventDisableTiming);
        }
        if (e != hipSuccess) return e;
    }
    p.ready = true;
    return hipSuccess;
}

// One lane's slice: device -> host (to_host) or host -> device, chunk by chunk.
hipError_t xfer_slice(XferLane &L, char *dst, const char *src, size_t n, bool to_host) {
    const size_t nch = (n + kXferChunk - 1) / kXferChunk;
    auto len_of = [&](size_t k) { return std::min(kXferChunk, n - k * kXferChunk); };
    if (to_host) {
        for (size_t k = 0; k <= nch; k++) {
            if (k < nch) {  // buffer k & 1 was emptied by the host copy of chunk k - 2
                hipError_t e = hipMemcpyAsync(L.buf[k & 1], src + k * kXferChunk, len_of(k),
                                              hipMemcpyDeviceToHost, L.st);
                if (e == hipSuccess) e = hipEventRecord(L.ev[k & 1], L.st);
                if (e != hipSuccess) return e;
            }
            if (k > 0) {
                const size_t j = k - 1;
                hipError_t e = hipEventSynchronize(L.ev[j & 1]);
                if (e != hipSuccess) return e;
                std::memcpy(dst + j * kXferChunk, L.buf[j & 1], len_of(j));
            }
        }
        return hipSuccess;
    }
    for (size_t k = 0; k < nch; k++) {
        if (k >= 2) {  // the DMA of chunk k - 2 has read buffer k & 1
            hipError_t e = hipEventSynchronize(L.ev[k & 1]);
            if (e != hipSuccess) return e;
        }
        std::memcpy(L.buf[k & 1], src + k * kXferChunk, len_of(k));
        hipError_t e = hipMemcpyAsync(dst + k * kXferChunk, L.buf[k & 1], len_of(k),
                                      hipMemcpyHostToDevice, L.st);
        if (e == hipSuccess) e = hipEventRecord(L.ev[k & 1], L.st);
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(L.st);
}

// Copies n bytes between device memory and pageable host memory; returns
// when the copy is complete.  The caller has the device current and has
// finished (synchronised) the work that produces or consumes the data.
hipError_t xfer(int device, void *dst, const void *src, size_t n, bool to_host) {
    const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    XferPool *p = n >= kXferDirect ? xfer_pool(device) : nullptr;
    if (!p) return n ? hipMemcpy(dst, src, n, kind) : hipSuccess;
    std::lock_guard<std::mutex> lk(p->mu);
    hipError_t e = xfer_ready(*p);
    if (e != hipSuccess) return e;
    // slices of whole chunks, one per lane (fewer lanes for smaller copies)
    const size_t chunks = (n + kXferChunk - 1) / kXferChunk;
    const int lanes = static_cast<int>(std::min<size_t>(kXferLanes, chunks));
    std::vector<hipError_t> err(lanes, hipSuccess);
    std::vector<std::thread> th;
    auto run = [&](int l) {
        const size_t c0 = chunks * l / lanes, c1 = chunks * (l + 1) / lanes;
        const size_t b0 = c0 * kXferChunk, b1 = std::min(n, c1 * kXferChunk);
        hipError_t le = hipSetDevice(device);
        if (le == hipSuccess)
            le = xfer_slice(p->lane[l], static_cast<char *>(dst) + b0,
                            static_cast<const char *>(src) + b0, b1 - b0, to_host);
        err[l] = le;
    };
    bool spawned = true;
    try {
        for (int l = 1; l < lanes; l++) th.emplace_back(run, l);
    } catch (...) {  // no thread: the lanes started so far finish, the rest in one copy
        spawned = false;
    }
    if (spawned) run(0);
    for (std::thread &t : th) t.join();
    if (!spawned) {
        const size_t b0 = chunks * (th.size() + 1) / lanes * kXferChunk;
        hipError_t le = hipSetDevice(device);
        if (le == hipSuccess) {
            // slices 1..th.size() are done; slice 0 and those past them remain
            const size_t s0 = chunks * 1 / lanes * kXferChunk;
            le = hipMemcpy(dst, src, std::min(n, s0), kind);
            if (le == hipSuccess && b0 < n)
                le = hipMemcpy(static_cast<char *>(dst) + b0,
                               static_cast<const char *>(src) + b0, n - b0, kind);
        }
        if (le != hipSuccess) return le;
        for (size_t l = 1; l <= th.size(); l++)
            if (err[l] != hipSuccess) return err[l];
        return hipSuccess;
    }
    for (hipError_t le : err)
        if (le != hipSuccess) return le;
    return hipSuccess;
}

// Host output of n bytes that the caller releases with free(): 2 MB aligned
// with transparent huge pages advised, so that first touch faults ~500 times
// less often.
void *alloc_host_out(size_t n) {
    void *p = nullptr;
    const size_t huge = static_cast<size_t>(2) << 20;
    if (posix_memalign(&p, n >= huge ? huge : 64, std::max<size_t>(n, 1)) != 0) return nullptr;
    if (n >= huge) (void)madvise(p, n, MADV_HUGEPAGE);
    return p;
}

}  // namespace

struct nat_graph {
    nat_mesh *m = nullptr;
    int M = 0;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf prob, qoff;                  // problem of each query (batch)
    DevBuf vals, cols, rowptr;          // the CSR of the last replay
    long long cap = 0;                  // CSR capacity (entries)
    std::vector<std::pair<const void *, size_t>> fingerprint;
    long long last_nnz = -1;
    // row arithmetic baked into the captured kernels (kernel choice and its
    // thresholds): a replay after nat_mesh_set_precision changed them is refused
    int precision = 0;
    float mixed_tau = 0.f, mixed_kappa = 0.f;
    int robust = 0;
};

namespace {

int issue_graph_rows(nat_mesh *m, nat_graph *g, const double *d_q, hipStream_t st) {
    const int M = g->M;
    const int P = m->problems();
    const size_t Mz = static_cast<size_t>(M);
    int bits = 1;
    while (bits < 7 && (static_cast<unsigned long long>(P) << (3 * (bits + 1))) <= Mz * 2 &&
           (static_cast<unsigned long long>(P) << (3 * (bits + 1))) <= (1ull << 24))
        bits++;
    unsigned *ctr = m->counters.as<unsigned>();
    HIP_OR_FAIL(nat::launch_zero_u32(m->counters.p, 64, st));
    const int *d_prob = P > 1 ? g->prob.as<int>() : nullptr;
    if (d_prob)
        HIP_OR_FAIL(nat::launch_problem_ids(g->qoff.as<long long>(), P, M, g->prob.as<int>(), st));
    m->dev.prob = d_prob;
    HIP_OR_FAIL(nat::launch_morton_sort(d_q, M, m->lp.as<nat::LocParams>(), d_prob, P, bits,
                                        m->order.as<int>(), m->sortw.p, st));
    nat::RowOut out;
    out.pool_vals = m->pool_vals.as<double>();
    out.pool_cols = m->pool_cols.as<int>();
    out.pool_cap = m->pool_cap;
    out.pool_used = reinterpret_cast<unsigned long long *>(ctr + 16);
    out.row_off = m->row_off.as<long long>();
    out.row_cnt = m->row_cnt.as<int>();
    out.row_total = m->row_total.as<double>();
    nat::Scratch sc0;
    nat::scratch_carve(m->scratch[0].p, m->caps[0], m->scap[0], &sc0);
    nat::Topo topo;
    topo.rec = m->rec.as<int4>();
    topo.rec_cap = m->rec_cap;
    topo.rec_used = reinterpret_cast<unsigned long long *>(ctr + 18);
    topo.rec_off = m->rec_off.as<long long>();
    HIP_OR_FAIL(nat::launch_locate(m->dev, m->caps[0], sc0, m->scap[0] / 256, d_q,
                                   m->order.as<int>(), M, m->start.as<int>(), st));
    HIP_OR_FAIL(nat::launch_topo(m->dev, m->topo_blocks * m->topo_grid, d_q, m->order.as<int>(),
                                 M, m->start.as<int>(), ctr + 32, topo, ctr + 8,
                                 m->ovf[0].as<int>(), st));
    if (m->work_sort)
        HIP_OR_FAIL(
            nat::launch_work_sort(m->order.as<int>(), M, topo, m->geom_order.as<int>(), st));
    HIP_OR_FAIL(nat::launch_geom(
        m->dev, m->nthr_geom / nat::kGeomThreads * m->geom_grid,
        m->geom_lanes == 2 ? geom_rows_grid(m) : 0, d_q,
        (m->work_sort ? m->geom_order : m->order).as<int>(), M, ctr + 40, topo,
        out, m->mixed_tau, m->mixed_kappa,
        m->precision == 1 ? reinterpret_cast<unsigned long long *>(ctr + 20) : nullptr, st));
    // the larger levels read their item counts from the device
    int cur = 0;
    const unsigned *n_dev = ctr + 8;
    if (m->wave_rows) {
        HIP_OR_FAIL(nat::launch_wave_rows(m->dev, m->wscr.as<char>(), m->wave_cap, d_q,
                                          m->ovf[0].as<int>(), 0, n_dev, m->start.as<int>(),
                                          ctr + 6, out, m->ovf[1].as<int>(), ctr + 12, st));
        cur = 1;
        n_dev = ctr + 12;
    }
    for (int level = 1; level < kLevels; level++) {
        nat::Scratch sc;
        nat::scratch_carve(m->scratch[level].p, m->caps[level], m->scap[level], &sc);
        out.ovf_list = m->ovf[cur ^ 1].as<int>();
        out.ovf_count = ctr + 8 + level;
        HIP_OR_FAIL(nat::launch_rows(m->dev, m->caps[level], sc, level,
                                     (m->scap[level] + 255) / 256, d_q, m->ovf[cur].as<int>(), 0,
                                     n_dev, ctr + level, out, st));
        cur ^= 1;
        n_dev = ctr + 8 + level;
    }
    long long *rp64 = m->rowptr64.as<long long>();
    HIP_OR_FAIL(nat::launch_scan_counts(m->row_cnt.as<int>(), rp64, M, m->scan.as<long long>(),
                                        st));
    HIP_OR_FAIL(nat::launch_gather_csr(rp64, m->row_off.as<long long>(), m->row_cnt.as<int>(),
                                       m->pool_vals.as<double>(), m->pool_cols.as<int>(),
                                       m->vorig.as<int>(), M, g->cap, g->vals.as<double>(),
                                       g->cols.as<int>(), g->rowptr.as<int>(), st));
    return 0;
}

int check_offsets(const nat_mesh *m, const int64_t *query_offsets) {
    const int P = m->problems();
    if (query_offsets[0] != 0) return fail(-1, "query_offsets[0] must be 0");
    for (int p = 0; p < P; p++)
        if (query_offsets[p + 1] < query_offsets[p])
            return fail(-1, "query_offsets must not decrease (problem %d)", p);
    if (query_offsets[P] > 2147483647LL) return fail(-1, "too many queries (%lld)",
                                                     static_cast<long long>(query_offsets[P]));
    return 0;
}

// Problem of each query of a batch call, in m->prob (*d_prob stays NULL for
// a single problem or no queries).
int problem_ids(nat_mesh *m, const int64_t *query_offsets, long long M, hipStream_t st,
                const int **d_prob) {
    const int P = m->problems();
    *d_prob = nullptr;
    if (P > 1 && M > 0) {
        HIP_OR_FAIL(m->qoff.ensure(sizeof(long long) * (P + 1)));
        HIP_OR_FAIL(m->prob.ensure(sizeof(int) * static_cast<size_t>(M)));
        HIP_OR_FAIL(hipMemcpyAsync(m->qoff.p, query_offsets, sizeof(long long) * (P + 1),
                                   hipMemcpyHostToDevice, st));
        HIP_OR_FAIL(nat::launch_problem_ids(m->qoff.as<long long>(), P, static_cast<int>(M),
                                            m->prob.as<int>(), st));
        *d_prob = m->prob.as<int>();
    }
    return 0;
}

// ---- row-path policy and autotuner

struct Policy {
    std::vector<unsigned char> budget;
    int first_level = 1, work_sort = 1, topo_grid = 1, geom_grid = 1, wave_rows = 1;
    int geom_lanes = 2;
};

Policy policy_of(const nat_mesh *m) {
    Policy p;
    p.budget = m->budget;
    p.first_level = m->first_level;
    p.work_sort = m->work_sort;
    p.topo_grid = m->topo_grid;
    p.geom_grid = m->geom_grid;
    p.wave_rows = m->wave_rows;
    p.geom_lanes = m->geom_lanes;
    return p;
}

int apply_policy(nat_mesh *m, const Policy &p) {
    m->budget = p.budget;
    m->first_level = p.first_level;
    m->work_sort = p.work_sort;
    m->topo_grid = p.topo_grid;
    m->geom_grid = p.geom_grid;
    m->wave_rows = p.wave_rows;
    m->geom_lanes = p.geom_lanes;
    return upload_route(m);
}

// Best wall time of `repeats` synchronised computes under policy p.
int time_policy(nat_mesh *m, const Policy &p, const double *d_q, int M, const int *d_prob,
                hipStream_t st, int repeats, double *best_ms) {
    if (int rc = apply_policy(m, p)) return rc;
    double best = 1e300;
    for (int r = 0; r < repeats; r++) {
        HIP_OR_FAIL(hipStreamSynchronize(st));
        const auto t0 = std::chrono::steady_clock::now();
        long long nnz = 0;
        if (int rc = compute_rows(m, d_q, M, d_prob, st, &nnz)) return rc;
        HIP_OR_FAIL(hipStreamSynchronize(st));
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
        best = std::min(best, ms);
    }
    *best_ms = best;
    return 0;
}

}  // namespace

extern "C" {

int nat_mesh_build(const double *data_points, int num_data_points, int device, nat_mesh **out) {
    g_err.clear();
    if (!out) return fail(-1, "out is NULL");
    *out = nullptr;
    if (!data_points || num_data_points <= 0)
        return fail(-1, "need at least one data point (got %d)", num_data_points);
    if (int rc = check_device(device)) return rc;
    nat_mesh *m = new nat_mesh();
    m->device = device;
    m->parts.resize(1);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = nat::build_packed_mesh(data_points, num_data_points, &m->parts[0]);
    m->build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
        delete m;
        return builder_error(rc);
    }
    if (int urc = finish_build(m)) {
        delete m;
        return urc;
    }
    *out = m;
    return 0;
}

int nat_mesh_build_ex(const double *data_points, int num_data_points, int device, int flags,
                      nat_mesh **out) {
    g_err.clear();
    if (flags == 0) return nat_mesh_build(data_points, num_data_points, device, out);
    if (!out) return fail(-1, "out is NULL");
    *out = nullptr;
    if (flags != NAT_BUILD_GPU) return fail(-1, "unknown build flags %d", flags);
    if (!data_points || num_data_points <= 0)
        return fail(-1, "need at least one data point (got %d)", num_data_points);
    if (device < 0) return fail(-1, "the GPU build needs a device (got %d)", device);
    if (int rc = check_device(device)) return rc;
    nat_mesh *m = new nat_mesh();
    m->device = device;
    m->parts.resize(1);
    {
        DeviceGuard g(device);
        nat::GpuBuild gb;
        std::string msg;
        const int rc = nat::gpu_build_tets(data_points, num_data_points, &gb, &msg);
        if (rc) {
            delete m;
            return fail(rc, "%s", msg.c_str());
        }
        const unsigned hw = std::thread::hardware_concurrency();
        nat::pack_balanced(num_data_points, std::move(gb.verts), gb.tet_v, gb.tet_n,
                           static_cast<int>(std::min(16u, hw ? hw : 1u)), &m->parts[0]);
        m->builder = 1;
        m->build_rounds = gb.rounds;
        m->build_seconds = gb.seconds;
    }
    if (int urc = finish_build(m)) {
        delete m;
        return urc;
    }
    *out = m;
    return 0;
}

int nat_mesh_from_packed(const double *verts, int num_data_points, const int32_t *tet_v,
                         const int32_t *tet_n, int num_tets, const int32_t *kd_left,
                         const int32_t *kd_right, const double *kd_pos, const double *kd_rect,
                         const int32_t *degeneracies, int device, nat_mesh **out) {
    g_err.clear();
    if (!out) return fail(-1, "out is NULL");
    *out = nullptr;
    if (!verts || !tet_v || !tet_n || !kd_left || !kd_right || !kd_pos || !kd_rect)
        return fail(-1, "NULL array");
    if (num_data_points <= 0 || num_tets <= 0)
        return fail(-1, "need data points and tets (got %d, %d)", num_data_points, num_tets);
    if (static_cast<long long>(num_tets) * 4 > 2147483647LL)
        return fail(-1, "too many tets (%d)", num_tets);
    if (int rc = check_device(device)) return rc;
    const int N = num_data_points, T = num_tets;
    // every index in range
    for (long long i = 0; i < 4LL * T; i++) {
        if (tet_v[i] < 0 || tet_v[i] >= N + 4)
            return fail(-1, "tet %lld: vertex slot %d outside [0, %d)", i / 4, tet_v[i], N + 4);
        if (tet_n[i] < -1 || tet_n[i] >= T)
            return fail(-1, "tet %lld: neighbour %d outside [-1, %d)", i / 4, tet_n[i], T);
    }
    for (int t = 0; t < T; t++)
        if (kd_left[t] < -1 || kd_left[t] >= T || kd_right[t] < -1 || kd_right[t] >= T)
            return fail(-1, "k-d node %d: child outside [-1, %d)", t, T);
    // the k-d tree: every node reached from the root (node 0) once, depth bounded
    constexpr int kMaxDepth = 4096;
    std::vector<unsigned char> seen(T, 0);
    std::vector<std::pair<int, int>> stack{{0, 1}};
    int depth = 0, reached = 0;
    while (!stack.empty()) {
        const auto [node, d] = stack.back();
        stack.pop_back();
        if (seen[node]) return fail(-1, "k-d node %d reached twice", node);
        seen[node] = 1;
        reached++;
        depth = std::max(depth, d);
        if (depth > kMaxDepth) return fail(-1, "k-d tree deeper than %d", kMaxDepth);
        if (kd_left[node] >= 0) stack.emplace_back(kd_left[node], d + 1);
        if (kd_right[node] >= 0) stack.emplace_back(kd_right[node], d + 1);
    }
    if (reached != T) return fail(-1, "k-d tree reaches %d of %d nodes", reached, T);
    nat_mesh *m = new nat_mesh();
    m->device = device;
    m->parts.resize(1);
    nat::PackedMesh &h = m->parts[0];
    h.n_data = N;
    h.n_tets = T;
    h.verts.assign(verts, verts + 3 * static_cast<size_t>(N + 4));
    h.tet_v.assign(tet_v, tet_v + 4 * static_cast<size_t>(T));
    h.tet_n.assign(tet_n, tet_n + 4 * static_cast<size_t>(T));
    h.kd_left.assign(kd_left, kd_left + T);
    h.kd_right.assign(kd_right, kd_right + T);
    h.kd_pos.assign(kd_pos, kd_pos + 3 * static_cast<size_t>(T));
    std::copy(kd_rect, kd_rect + 6, h.kd_rect);
    h.kd_depth = depth;
    if (degeneracies) {
        h.coplanar = degeneracies[0];
        h.cospherical = degeneracies[1];
        h.edge_anomalies = degeneracies[2];
    }
    if (int urc = finish_build(m)) {
        delete m;
        return urc;
    }
    *out = m;
    return 0;
}

int nat_batch_build(const double *const *data_points, const int32_t *num_data_points,
                    int num_problems, int device, int num_threads, nat_mesh **out) {
    g_err.clear();
    if (!out) return fail(-1, "out is NULL");
    *out = nullptr;
    if (num_problems <= 0 || !data_points || !num_data_points)
        return fail(-1, "need at least one problem (got %d)", num_problems);
    for (int p = 0; p < num_problems; p++)
        if (!data_points[p] || num_data_points[p] <= 0)
            return fail(-1, "problem %d needs at least one data point (got %d)", p,
                        num_data_points[p]);
    if (int rc = check_device(device)) return rc;
    nat_mesh *m = new nat_mesh();
    m->device = device;
    m->parts.resize(num_problems);
    // independent host builds, largest first over a small thread pool
    std::vector<int> order(num_problems), rcs(num_problems, 0);
    for (int p = 0; p < num_problems; p++) order[p] = p;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return num_data_points[a] > num_data_points[b]; });
    int nt = num_threads > 0 ? num_threads : static_cast<int>(std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, num_problems));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= num_problems) return;
            const int p = order[i];
            rcs[p] = nat::build_packed_mesh(data_points[p], num_data_points[p], &m->parts[p]);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    for (int p = 0; p < num_problems; p++)
        if (rcs[p]) {
            delete m;
            const int rc = builder_error(rcs[p]);
            g_err = "problem " + std::to_string(p) + ": " + g_err;
            return rc;
        }
    if (int urc = finish_build(m)) {
        delete m;
        return urc;
    }
    *out = m;
    return 0;
}

int nat_batch_offsets(const nat_mesh *m, int64_t *vertex_offsets, int64_t *tet_offsets) {
    if (!m) return fail(-1, "NULL mesh");
    const int P = m->problems();
    for (int p = 0; p <= P; p++) {
        if (vertex_offsets) vertex_offsets[p] = m->vbase[p];
        if (tet_offsets) tet_offsets[p] = m->tbase[p];
    }
    return 0;
}

void nat_mesh_free(nat_mesh *m) {
    if (!m) return;
    if (m->device >= 0) {
        DeviceGuard g(m->device);
        m->release_all();
    }
    delete m;
}

int nat_mesh_get_info(const nat_mesh *m, nat_mesh_info *info) {
    if (!m || !info) return fail(-1, "NULL argument");
    const int P = m->problems();
    info->n_data = static_cast<int32_t>(m->vbase[P]);
    info->n_tets = static_cast<int32_t>(m->tbase[P]);
    info->kd_depth = m->kd_depth_max;
    info->coplanar = info->cospherical = info->edge_anomalies = 0;
    for (const nat::PackedMesh &h : m->parts) {
        info->coplanar += h.coplanar;
        info->cospherical += h.cospherical;
        info->edge_anomalies += h.edge_anomalies;
    }
    info->n_problems = P;
    info->device = m->device;
    info->num_cus = m->num_cus;
    info->device_bytes = m->device_bytes();
    info->workspace_bytes = m->workspace_bytes();
    info->builder = m->builder;
    info->build_rounds = m->build_rounds;
    info->build_seconds = m->build_seconds;
    return 0;
}

int nat_mesh_delaunay_audit(const nat_mesh *m, int64_t *counts) {
    g_err.clear();
    if (!m || !counts) return fail(-1, "NULL argument");
    if (m->problems() != 1) return fail(-1, "nat_mesh_delaunay_audit needs a single-problem mesh");
    if (m->device < 0) return fail(-1, "host-only mesh has no device");
    DeviceGuard g(m->device);
    const nat::PackedMesh &h = m->parts[0];
    long long c[4] = {0, 0, 0, 0};
    std::string msg;
    const int rc = nat::gpu_delaunay_audit(h.verts.data(), h.n_data + 4, h.tet_v.data(),
                                           h.tet_n.data(), h.n_tets, c, &msg);
    if (rc) return fail(rc, "%s", msg.c_str());
    for (int i = 0; i < 4; i++) counts[i] = c[i];
    return 0;
}

int nat_mesh_export(const nat_mesh *m, double *verts, int32_t *tet_v, int32_t *tet_n,
                    int32_t *kd_left, int32_t *kd_right, double *kd_pos, double *kd_rect) {
    if (!m) return fail(-1, "NULL mesh");
    if (m->problems() != 1) return fail(-1, "nat_mesh_export needs a single-problem mesh");
    const nat::PackedMesh &h = m->parts[0];
    if (verts) std::memcpy(verts, h.verts.data(), sizeof(double) * h.verts.size());
    if (tet_v) std::memcpy(tet_v, h.tet_v.data(), sizeof(int32_t) * h.tet_v.size());
    if (tet_n) std::memcpy(tet_n, h.tet_n.data(), sizeof(int32_t) * h.tet_n.size());
    if (kd_left) std::memcpy(kd_left, h.kd_left.data(), sizeof(int32_t) * h.kd_left.size());
    if (kd_right) std::memcpy(kd_right, h.kd_right.data(), sizeof(int32_t) * h.kd_right.size());
    if (kd_pos) std::memcpy(kd_pos, h.kd_pos.data(), sizeof(double) * h.kd_pos.size());
    if (kd_rect) std::memcpy(kd_rect, h.kd_rect, sizeof(double) * 6);
    return 0;
}

int nat_weights_compute(nat_mesh *m, const double *d_queries, int num_query_points,
                        void *stream, int64_t *nnz_out) {
    g_err.clear();
    if (!m || !nnz_out) return fail(-1, "NULL argument");
    if (num_query_points < 0) return fail(-1, "negative query count");
    if (num_query_points > 0 && !d_queries) return fail(-1, "NULL queries");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (m->problems() != 1)
        return fail(-1, "a batch mesh needs nat_batch_weights_compute (query offsets)");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    long long nnz = 0;
    const int rc = compute_rows(m, d_queries, num_query_points, nullptr,
                                static_cast<hipStream_t>(stream), &nnz);
    if (rc) return rc;
    *nnz_out = nnz;
    return 0;
}

int nat_batch_weights_compute(nat_mesh *m, const double *d_queries,
                              const int64_t *query_offsets, void *stream, int64_t *nnz_out) {
    g_err.clear();
    if (!m || !nnz_out || !query_offsets) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (int rc = check_offsets(m, query_offsets)) return rc;
    const long long M = query_offsets[m->problems()];
    if (M > 0 && !d_queries) return fail(-1, "NULL queries");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int *d_prob = nullptr;
    if (int rc = problem_ids(m, query_offsets, M, st, &d_prob)) return rc;
    long long nnz = 0;
    const int rc = compute_rows(m, d_queries, static_cast<int>(M), d_prob, st, &nnz);
    if (rc) return rc;
    *nnz_out = nnz;
    return 0;
}

int nat_weights_emit(nat_mesh *m, double *d_vals, int32_t *d_cols, int32_t *d_rowptr,
                     void *stream) {
    g_err.clear();
    if (!m || !d_rowptr) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    return emit_rows(m, d_vals, d_cols, d_rowptr, static_cast<hipStream_t>(stream));
}

int nat_get_weights(const double *query_points, int num_query_points, nat_mesh *m,
                    int num_data_points, double **weight_values, int **weight_col_inds,
                    int *weight_row_ptrs) {
    g_err.clear();
    if (!m || !weight_values || !weight_col_inds || !weight_row_ptrs)
        return fail(-2, "NULL argument");
    if (num_query_points < 0) return fail(-2, "negative query count");
    if (m->problems() != 1)
        return fail(-2, "nat_get_weights needs a single-problem mesh (use the batch calls)");
    if (num_data_points != m->parts[0].n_data)
        return fail(-2, "num_data_points %d does not match the mesh (%d)", num_data_points,
                    m->parts[0].n_data);
    if (m->device < 0) return fail(-2, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    const int M = num_query_points;
    const size_t Mz = static_cast<size_t>(M);
    hipStream_t st = nullptr;
    if (M > 0) {
        HIP_OR_FAIL(m->q.ensure(sizeof(double) * 3 * Mz));
        HIP_OR_FAIL(hipStreamSynchronize(nullptr));  // no earlier work still reads m->q
        HIP_OR_FAIL(xfer(m->device, m->q.p, query_points, sizeof(double) * 3 * Mz, false));
    }
    long long nnz = 0;
    int rc = compute_rows(m, m->q.as<double>(), M, nullptr, st, &nnz);
    if (rc) return rc;
    if (nnz > 2147483647LL) return -1;  // nothing allocated (natural_insertionfree.c:805-812)
    const size_t nz = static_cast<size_t>(m->last_nnz);
    HIP_OR_FAIL(m->csr_vals.ensure(sizeof(double) * std::max<size_t>(nz, 1)));
    HIP_OR_FAIL(m->csr_cols.ensure(sizeof(int) * std::max<size_t>(nz, 1)));
    HIP_OR_FAIL(m->csr_rowptr.ensure(sizeof(int) * (Mz + 1)));
    rc = emit_rows(m, m->csr_vals.as<double>(), m->csr_cols.as<int>(), m->csr_rowptr.as<int>(),
                   st);
    if (rc) return rc;
    double *hv = static_cast<double *>(alloc_host_out(sizeof(double) * nz));
    int *hc = static_cast<int *>(alloc_host_out(sizeof(int) * nz));
    if (!hv || !hc) {
        std::free(hv);
        std::free(hc);
        return fail(-3, "host allocation of %zu entries failed", nz);
    }
    hipError_t e = hipStreamSynchronize(st);  // the CSR is complete
    if (e == hipSuccess)
        e = xfer(m->device, weight_row_ptrs, m->csr_rowptr.p, sizeof(int) * (Mz + 1), true);
    if (e == hipSuccess && nz) e = xfer(m->device, hv, m->csr_vals.p, sizeof(double) * nz, true);
    if (e == hipSuccess && nz) e = xfer(m->device, hc, m->csr_cols.p, sizeof(int) * nz, true);
    if (e != hipSuccess) {
        std::free(hv);
        std::free(hc);
        return fail(-10, "copy-out failed: %s", hipGetErrorString(e));
    }
    *weight_values = hv;
    *weight_col_inds = hc;
    return 0;
}

int nat_spmm_csr(int num_rows, int num_cols, int dim, const int32_t *d_rowptr,
                 const int32_t *d_cols, const void *d_vals, int vals_dtype,
                 const void *d_values, int values_dtype, const int32_t *d_row_order,
                 void *d_out, int out_dtype, void *stream) {
    g_err.clear();
    (void)num_cols;
    if (num_rows < 0 || dim < 0) return fail(-1, "negative size");
    if (num_rows == 0 || dim == 0) return 0;
    if (!d_rowptr || !d_out) return fail(-1, "NULL argument");
    if (vals_dtype < 0 || vals_dtype > 1 || values_dtype < 0 || values_dtype > 1 ||
        out_dtype < (vals_dtype | values_dtype) || out_dtype > 1)
        return fail(-1, "unsupported dtype combination %d %d %d", vals_dtype, values_dtype,
                    out_dtype);
    hipError_t e = nat::launch_spmm(num_rows, dim, d_rowptr, d_cols, d_vals, vals_dtype,
                                    d_values, values_dtype, d_row_order, d_out, out_dtype,
                                    static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(-10, "spmm launch failed: %s", hipGetErrorString(e));
    return 0;
}

int nat_weights_spmm(nat_mesh *m, int num_query_points, const void *d_values,
                     int num_data_points, int values_dtype, int dim, double *d_out,
                     void *stream) {
    g_err.clear();
    if (!m) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (values_dtype < 0 || values_dtype > 1) return fail(-1, "values dtype %d", values_dtype);
    if (dim < 1 || dim > 4)
        return fail(-1, "dim %d: nat_weights_spmm takes 1 to 4 value columns", dim);
    if (num_data_points != m->dev.n_data)
        return fail(-1, "values have %d rows, the mesh %d data points", num_data_points,
                    m->dev.n_data);
    std::lock_guard<std::mutex> lk(m->mu);
    if (m->last_M < 0 || num_query_points != m->last_M)
        return fail(-6, "last compute had %d queries, not %d", m->last_M, num_query_points);
    if (num_query_points == 0) return 0;
    if (!d_values || !d_out) return fail(-1, "NULL argument");
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nd = m->dev.n_data;
    const int elt = values_dtype ? 8 : 4;
    HIP_OR_FAIL(m->vals_fwd.ensure(static_cast<size_t>(elt) *
                                   std::max<size_t>(1, static_cast<size_t>(nd) * dim)));
    HIP_OR_FAIL(nat::launch_gather_rows_bytes(d_values, m->vorig.as<int>(), nd, dim, elt,
                                              m->vals_fwd.p, st));
    HIP_OR_FAIL(nat::launch_pool_spmm(num_query_points, dim, m->row_off.as<long long>(),
                                      m->row_cnt.as<int>(), m->pool_cols.as<int>(),
                                      m->pool_vals.as<double>(), m->vals_fwd.p, values_dtype,
                                      m->order.as<int>(), d_out, st));
    return 0;
}

int nat_weights_row_order(const nat_mesh *m, int num_query_points, const int32_t **out) {
    g_err.clear();
    if (!m || !out) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (num_query_points != m->last_M || m->last_M < 0)
        return fail(-1, "last compute had %d queries, not %d", m->last_M, num_query_points);
    *out = m->order.as<const int32_t>();
    return 0;
}

int nat_topology_save(nat_mesh *m, void *d_dst, int64_t capacity, int64_t *bytes,
                      void *stream) {
    g_err.clear();
    if (!m || !bytes) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    if (m->last_M < 0) return fail(-6, "nat_topology_save without a preceding compute");
    const long long M = m->last_M;
    const long long units = M > 0 ? m->last_units : 0;
    const size_t need = topo_rec_off(M) + sizeof(int4) * static_cast<size_t>(units);
    *bytes = static_cast<int64_t>(need);
    if (!d_dst) return 0;
    if (capacity < static_cast<int64_t>(need))
        return fail(-7, "topology needs %zu bytes, capacity %lld", need,
                    static_cast<long long>(capacity));
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int P = m->problems();
    const long long hdr[8] = {kTopoMagic, M, units, m->dev.n_data, P, 0, 0, 0};
    char *dst = static_cast<char *>(d_dst);
    HIP_OR_FAIL(hipMemcpyAsync(dst, hdr, sizeof(hdr), hipMemcpyHostToDevice, st));
    if (M > 0) {
        HIP_OR_FAIL(hipMemcpyAsync(dst + kTopoHeader, m->rec_off.p, 8 * static_cast<size_t>(M),
                                   hipMemcpyDeviceToDevice, st));
        if (P > 1)
            HIP_OR_FAIL(hipMemcpyAsync(dst + topo_prob_off(M), m->prob.p,
                                       4 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        // the backward keeps the Morton order: its gathers (values, dL/dy)
        // favour spatial locality over grouping by cavity size
        HIP_OR_FAIL(hipMemcpyAsync(dst + topo_order_off(M), m->order.p,
                                   4 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        HIP_OR_FAIL(hipMemcpyAsync(dst + topo_start_off(M), m->start.p,
                                   4 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        HIP_OR_FAIL(hipMemcpyAsync(dst + topo_total_off(M), m->row_total.p,
                                   8 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        // records already in place when the compute wrote them to d_dst
        // (nat_topology_target); else from wherever it wrote them
        const char *rsrc = m->ext_written ? m->ext_written + topo_rec_off(M)
                                          : static_cast<const char *>(m->rec.p);
        if (units > 0 && m->ext_written != dst)
            HIP_OR_FAIL(hipMemcpyAsync(dst + topo_rec_off(M), rsrc,
                                       sizeof(int4) * static_cast<size_t>(units),
                                       hipMemcpyDeviceToDevice, st));
    }
    // the header comes from the stack: finish before returning
    HIP_OR_FAIL(hipStreamSynchronize(st));
    return 0;
}

int nat_topology_target(nat_mesh *m, void *d_dst, int64_t capacity, int num_query_points) {
    g_err.clear();
    if (!m) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (d_dst && (capacity < 0 || num_query_points < 0))
        return fail(-1, "bad size (capacity %lld, M %d)", static_cast<long long>(capacity),
                    num_query_points);
    if (reinterpret_cast<uintptr_t>(d_dst) & 15u) return fail(-1, "d_dst must be 16-byte aligned");
    std::lock_guard<std::mutex> lk(m->mu);
    m->ext_topo = static_cast<char *>(d_dst);
    m->ext_capacity = capacity;
    m->ext_M = num_query_points;
    return 0;
}

int nat_grad_queries(nat_mesh *m, const void *d_topo, const double *d_queries,
                     int num_query_points, const int32_t *d_rowptr, const double *d_vals,
                     const double *d_gy, const double *d_values, int dim, double *d_dq,
                     void *stream) {
    g_err.clear();
    if (!m || !d_topo) return fail(-1, "NULL argument");
    if (num_query_points < 0 || dim <= 0) return fail(-1, "bad size (M %d, D %d)",
                                                      num_query_points, dim);
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (num_query_points > 0 &&
        (!d_queries || !d_rowptr || !d_gy || !d_values || !d_dq))
        return fail(-1, "NULL argument");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    return grad_rows(m, static_cast<const char *>(d_topo), d_queries, num_query_points, d_rowptr,
                     d_vals, d_gy, d_values, dim, d_dq, static_cast<hipStream_t>(stream));
}

int nat_graph_create(nat_mesh *m, const double *d_queries, int num_query_points,
                     const int64_t *query_offsets, void *stream, nat_graph **out) {
    g_err.clear();
    if (!m || !out || !d_queries) return fail(-1, "NULL argument");
    *out = nullptr;
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    const int P = m->problems();
    if (num_query_points <= 0) return fail(-1, "a graph needs queries (got %d)", num_query_points);
    if (P > 1) {
        if (!query_offsets) return fail(-1, "a batch mesh needs query offsets");
        if (query_offsets[0] != 0 || query_offsets[P] != num_query_points)
            return fail(-1, "query_offsets must run from 0 to num_query_points");
        for (int p = 0; p < P; p++)
            if (query_offsets[p + 1] < query_offsets[p])
                return fail(-1, "query_offsets must not decrease (problem %d)", p);
    }
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard dg(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    nat_graph *g = new nat_graph();
    g->m = m;
    g->M = num_query_points;
    const size_t Mz = static_cast<size_t>(g->M);
    auto bail = [&](int rc) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->st) (void)hipStreamDestroy(g->st);
        g->prob.release();
        g->qoff.release();
        g->vals.release();
        g->cols.release();
        g->rowptr.release();
        delete g;
        return rc;
    };
    int rc = 0;
    if (P > 1) {
        if (g->qoff.ensure(sizeof(long long) * (P + 1)) != hipSuccess ||
            g->prob.ensure(sizeof(int) * Mz) != hipSuccess)
            return bail(fail(-10, "graph buffer allocation failed"));
        if (hipMemcpy(g->qoff.p, query_offsets, sizeof(long long) * (P + 1),
                      hipMemcpyHostToDevice) != hipSuccess ||
            nat::launch_problem_ids(g->qoff.as<long long>(), P, g->M, g->prob.as<int>(), st) !=
                hipSuccess)
            return bail(fail(-10, "graph problem ids failed"));
    }
    // warm-up: sizes every capacity, then 25% headroom for the replays
    long long nnz = 0;
    const bool timing = m->timing;
    m->timing = false;
    rc = compute_rows(m, d_queries, g->M, P > 1 ? g->prob.as<int>() : nullptr, st, &nnz);
    m->timing = timing;
    if (rc) return bail(rc);
    int items[kLevels];
    for (int l = 0; l < kLevels; l++) items[l] = l < m->times.levels ? m->times.rows_items[l] : 0;
    const long long pool = nnz + nnz / 4 + 4096, rec = m->last_units + m->last_units / 4 + 4096;
    if (pool > 2147483647LL) return bail(fail(-1, "graph output above INT32_MAX entries"));
    if (m->pool_cap < pool) {
        if (m->pool_vals.ensure(sizeof(double) * pool) != hipSuccess ||
            m->pool_cols.ensure(sizeof(int) * pool) != hipSuccess)
            return bail(fail(-10, "graph pool allocation failed"));
        m->pool_cap = pool;
    }
    if (m->rec_cap < rec) {
        if (m->rec.ensure(sizeof(int4) * rec) != hipSuccess)
            return bail(fail(-10, "graph record allocation failed"));
        m->rec_cap = rec;
    }
    for (int l = 1; l < kLevels; l++) {
        int n = 0;
        if ((rc = ensure_level_scratch(m, l, std::max(256, 2 * items[l]), &n))) return bail(rc);
    }
    if (m->wave_rows) {
        int n = 0;
        if ((rc = ensure_wave_scratch(m, std::max(64, 2 * m->times.wave_items), &n)))
            return bail(rc);
    }
    g->cap = m->pool_cap;
    if (g->vals.ensure(sizeof(double) * g->cap) != hipSuccess ||
        g->cols.ensure(sizeof(int) * g->cap) != hipSuccess ||
        g->rowptr.ensure(sizeof(int) * (Mz + 1)) != hipSuccess)
        return bail(fail(-10, "graph output allocation failed"));
    if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(-10, "warm-up failed"));
    g->fingerprint = workspace_fingerprint(m);
    g->precision = m->precision;
    g->mixed_tau = m->mixed_tau;
    g->mixed_kappa = m->mixed_kappa;
    g->robust = m->dev.robust;
    if (hipStreamCreate(&g->st) != hipSuccess) return bail(fail(-10, "stream creation failed"));
    if (hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bail(fail(-10, "hipStreamBeginCapture failed"));
    const nat::DevMesh saved = m->dev;
    rc = issue_graph_rows(m, g, d_queries, g->st);
    m->dev = saved;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(g->st, &graph);
    g->graph = graph;
    if (rc) return bail(rc);
    if (ec != hipSuccess) return bail(fail(-10, "hipStreamEndCapture: %s", hipGetErrorString(ec)));
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        return bail(fail(-10, "hipGraphInstantiate failed"));
    m->last_M = -1;  // the workspace now serves the graph
    *out = g;
    return 0;
}

int nat_graph_launch(nat_graph *g, void *stream, int64_t *nnz_out) {
    g_err.clear();
    if (!g || !nnz_out) return fail(-1, "NULL argument");
    nat_mesh *m = g->m;
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard dg(m->device);
    if (workspace_fingerprint(m) != g->fingerprint)
        return fail(-9, "the mesh workspace was reallocated since the graph was made");
    if (m->precision != g->precision ||
        (g->precision == 1 && (m->mixed_tau != g->mixed_tau || m->mixed_kappa != g->mixed_kappa)))
        return fail(-11, "the mesh's precision changed since the graph was made (captured %s, "
                         "now %s)", g->precision ? "mixed" : "fp64",
                    m->precision ? "mixed" : "fp64");
    if (m->dev.robust != g->robust)
        return fail(-11, "the mesh's predicates changed since the graph was made (captured %s, "
                         "now %s)", g->robust ? "robust" : "fp64",
                    m->dev.robust ? "robust" : "fp64");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_OR_FAIL(hipGraphLaunch(g->exec, st));
    unsigned h[20];
    HIP_OR_FAIL(hipMemcpyAsync(h, m->counters.p, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    unsigned long long pool_used, rec_used;
    std::memcpy(&pool_used, h + 16, sizeof(pool_used));
    std::memcpy(&rec_used, h + 18, sizeof(rec_used));
    m->last_M = -1;
    g->last_nnz = -1;
    if (static_cast<long long>(pool_used) > g->cap || static_cast<long long>(rec_used) > m->rec_cap)
        return fail(-8, "graph capacity exceeded (entries %llu of %lld, records %llu of %lld)",
                    pool_used, g->cap, rec_used, m->rec_cap);
    if (h[11] > 0)
        return fail(-4, "%u queries need more than %d cavity tets", h[11],
                    m->caps[kLevels - 1].cav);
    g->last_nnz = static_cast<long long>(pool_used);
    *nnz_out = g->last_nnz;
    return 0;
}

int nat_graph_copy_out(nat_graph *g, double *d_vals, int32_t *d_cols, int32_t *d_rowptr,
                       void *stream) {
    g_err.clear();
    if (!g || !d_rowptr) return fail(-1, "NULL argument");
    if (g->last_nnz < 0) return fail(-6, "no successful replay to copy");
    if (g->last_nnz > 0 && (!d_vals || !d_cols)) return fail(-1, "NULL argument");
    DeviceGuard dg(g->m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t n = static_cast<size_t>(g->last_nnz);
    if (n) {
        HIP_OR_FAIL(hipMemcpyAsync(d_vals, g->vals.p, sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, st));
        HIP_OR_FAIL(hipMemcpyAsync(d_cols, g->cols.p, sizeof(int) * n, hipMemcpyDeviceToDevice,
                                   st));
    }
    HIP_OR_FAIL(hipMemcpyAsync(d_rowptr, g->rowptr.p, sizeof(int) * (static_cast<size_t>(g->M) + 1),
                               hipMemcpyDeviceToDevice, st));
    return 0;
}

void nat_graph_free(nat_graph *g) {
    if (!g) return;
    DeviceGuard dg(g->m->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->st) (void)hipStreamDestroy(g->st);
    g->prob.release();
    g->qoff.release();
    g->vals.release();
    g->cols.release();
    g->rowptr.release();
    delete g;
}

int nat_insphere_classify(const double *points, int n, double *det, int32_t *uncertain) {
    g_err.clear();
    if (n < 0 || (n > 0 && (!points || !det || !uncertain))) return fail(-1, "bad arguments");
    for (int i = 0; i < n; i++) {
        const double *p = points + 15 * i;
        const nat::d3 a = {p[0], p[1], p[2]}, b = {p[3], p[4], p[5]}, c = {p[6], p[7], p[8]};
        const nat::d3 d = {p[9], p[10], p[11]}, e = {p[12], p[13], p[14]};
        det[i] = nat::insphere(a, b, c, d, e);
        uncertain[i] = nat::insphere_uncertain(a, b, c, d, e) ? 1 : 0;
    }
    return 0;
}

int nat_predicate_audit(nat_mesh *m, const void *d_topo, const double *d_queries,
                        int num_query_points, int64_t *decisions, int64_t *uncertain,
                        void *stream) {
    g_err.clear();
    if (!m || !d_topo || !decisions || !uncertain) return fail(-1, "NULL argument");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (num_query_points > 0 && !d_queries) return fail(-1, "NULL queries");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const char *blob = static_cast<const char *>(d_topo);
    long long hdr[8];
    HIP_OR_FAIL(hipMemcpyAsync(hdr, blob, sizeof(hdr), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    const int M = num_query_points;
    if (hdr[0] != kTopoMagic || hdr[1] != M || hdr[3] != m->dev.n_data)
        return fail(-1, "d_topo is not a saved topology of these queries on this mesh");
    HIP_OR_FAIL(m->counters.ensure(256));
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(m->counters.p) + 16;
    HIP_OR_FAIL(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), st));
    nat::Topo topo;
    topo.rec = reinterpret_cast<int4 *>(const_cast<char *>(blob) + topo_rec_off(M));
    topo.rec_cap = hdr[2];
    topo.rec_used = nullptr;
    topo.rec_off = reinterpret_cast<long long *>(const_cast<char *>(blob) + kTopoHeader);
    if (M > 0) HIP_OR_FAIL(nat::launch_predicate_audit(m->dev, d_queries, M, topo, cnt, st));
    unsigned long long h[2] = {0, 0};
    HIP_OR_FAIL(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    *decisions = static_cast<int64_t>(h[0]);
    *uncertain = static_cast<int64_t>(h[1]);
    return 0;
}

int nat_topology_row_order(const void *d_topo, int num_query_points, const int32_t **out) {
    g_err.clear();
    if (!d_topo || !out) return fail(-1, "NULL argument");
    long long hdr[8];
    HIP_OR_FAIL(hipMemcpy(hdr, d_topo, sizeof(hdr), hipMemcpyDeviceToHost));
    if (hdr[0] != kTopoMagic) return fail(-1, "d_topo is not a saved topology");
    if (hdr[1] != num_query_points)
        return fail(-1, "saved topology has %lld queries, not %d", hdr[1], num_query_points);
    *out = reinterpret_cast<const int32_t *>(static_cast<const char *>(d_topo) +
                                             topo_order_off(num_query_points));
    return 0;
}

int nat_spmm_t_csr(int num_rows, int num_cols, int dim, const int32_t *d_rowptr,
                   const int32_t *d_cols, const double *d_vals, const double *d_gy,
                   const int32_t *d_row_order, double *d_out, void *stream) {
    g_err.clear();
    if (num_rows < 0 || num_cols < 0 || dim < 0) return fail(-1, "negative size");
    if (num_cols == 0 || dim == 0) return 0;
    if (!d_out || (num_rows > 0 && (!d_rowptr || !d_gy))) return fail(-1, "NULL argument");
    // few columns (at most 512 KB of output): 8 to 64 private copies of the
    // output, 32 MB at most, stream-ordered, so that the blocks' atomics do
    // not all queue on the same addresses
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t out_bytes = sizeof(double) * static_cast<size_t>(num_cols) * dim;
    int P = 1;
    double *priv = nullptr;
    if (num_rows > 0 && out_bytes <= (512u << 10)) {
        P = static_cast<int>(std::min<size_t>(64, (32u << 20) / out_bytes));
        while (P & (P - 1)) P &= P - 1;  // a power of two
        if (P < 8 || hipMallocAsync(reinterpret_cast<void **>(&priv), out_bytes * P, st) !=
                         hipSuccess) {
            (void)hipGetLastError();
            P = 1;
            priv = nullptr;
        }
    }
    hipError_t e = nat::launch_spmm_t(num_rows, num_cols, dim, d_rowptr, d_cols, d_vals, d_gy,
                                      d_row_order, d_out, priv, P, st);
    if (priv) {
        const hipError_t ef = hipFreeAsync(priv, st);
        if (e == hipSuccess) e = ef;
    }
    if (e != hipSuccess) return fail(-10, "spmm_t launch failed: %s", hipGetErrorString(e));
    return 0;
}

int nat_csr_sort_rows(int num_rows, const int32_t *d_rowptr, int32_t *d_cols, void *d_vals,
                      int vals_dtype, void *stream) {
    g_err.clear();
    if (num_rows < 0) return fail(-1, "negative size");
    if (num_rows == 0) return 0;
    if (!d_rowptr || vals_dtype < 0 || vals_dtype > 1) return fail(-1, "bad argument");
    hipError_t e = nat::launch_sort_rows(num_rows, d_rowptr, d_cols, d_vals, vals_dtype,
                                         static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(-10, "sort launch failed: %s", hipGetErrorString(e));
    return 0;
}

int nat_mesh_set_precision(nat_mesh *m, int mode, double cond_threshold,
                           double cancel_threshold) {
    g_err.clear();
    if (!m) return fail(-1, "NULL mesh");
    if (mode != 0 && mode != 1) return fail(-1, "precision mode %d (0 = fp64, 1 = mixed)", mode);
    if (mode == 1 && !(cond_threshold > 0.0 && cond_threshold <= 1.0))
        return fail(-1, "condition threshold %g outside (0, 1]", cond_threshold);
    if (mode == 1 && !(cancel_threshold >= 1.0))
        return fail(-1, "cancellation threshold %g below 1", cancel_threshold);
    std::lock_guard<std::mutex> lk(m->mu);
    m->precision = mode;
    if (mode == 1) {
        m->mixed_tau = static_cast<float>(cond_threshold);
        m->mixed_kappa = static_cast<float>(cancel_threshold);
    }
    return 0;
}

int nat_mesh_set_predicates(nat_mesh *m, int mode) {
    g_err.clear();
    if (!m) return fail(-1, "NULL mesh");
    if (mode != 0 && mode != 1) return fail(-1, "predicates mode %d (0 = fp64, 1 = robust)", mode);
    std::lock_guard<std::mutex> lk(m->mu);
    m->dev.robust = mode;
    return 0;
}

int nat_mesh_set_mixed_check(nat_mesh *m, int stride) {
    g_err.clear();
    if (!m) return fail(-1, "NULL mesh");
    if (stride < 0) return fail(-1, "mixed check stride %d below 0", stride);
    std::lock_guard<std::mutex> lk(m->mu);
    m->mixed_check = stride;
    return 0;
}

int nat_predicate_exact(const double *points, int n, int kind, int32_t *sign,
                        int32_t *uncertain) {
    g_err.clear();
    if (n < 0 || (kind != 0 && kind != 1) || (n > 0 && (!points || !sign || !uncertain)))
        return fail(-1, "bad arguments");
    const int w = kind == 0 ? 12 : 15;
    for (int i = 0; i < n; i++) {
        const double *p = points + static_cast<size_t>(w) * i;
        const nat::d3 a = {p[0], p[1], p[2]}, b = {p[3], p[4], p[5]}, c = {p[6], p[7], p[8]};
        const nat::d3 d = {p[9], p[10], p[11]};
        if (kind == 0) {
            sign[i] = nat::orient3d_exact_sign(a, b, c, d);
            uncertain[i] = nat::orient3d_uncertain(a, b, c, d) ? 1 : 0;
        } else {
            const nat::d3 e = {p[12], p[13], p[14]};
            sign[i] = nat::insphere_exact_sign(a, b, c, d, e);
            uncertain[i] = nat::insphere_uncertain(a, b, c, d, e) ? 1 : 0;
        }
    }
    return 0;
}

int nat_mesh_set_policy(nat_mesh *m, const nat_policy *policy, const uint8_t *budgets) {
    g_err.clear();
    if (!m || !policy) return fail(-1, "NULL argument");
    if (policy->first_level != 1 && policy->first_level != 2)
        return fail(-1, "first_level %d (1 or 2)", policy->first_level);
    if (policy->work_sort != 0 && policy->work_sort != 1)
        return fail(-1, "work_sort %d (0 or 1)", policy->work_sort);
    if (policy->wave_rows != 0 && policy->wave_rows != 1)
        return fail(-1, "wave_rows %d (0 or 1)", policy->wave_rows);
    if (policy->geom_lanes < 0 || policy->geom_lanes > 2)
        return fail(-1, "geom_lanes %d (0, 1 or 2)", policy->geom_lanes);
    if (policy->topo_grid < 1 || policy->topo_grid > 4 || policy->geom_grid < 1 ||
        policy->geom_grid > 4)
        return fail(-1, "grid multipliers %d, %d outside 1..4", policy->topo_grid,
                    policy->geom_grid);
    if (budgets)
        for (int i = 0; i < m->problems(); i++)
            if (budgets[i] > kGroupBudget)
                return fail(-1, "problem %d: budget %d above %d", i, budgets[i], kGroupBudget);
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    Policy p = policy_of(m);
    if (budgets) p.budget.assign(budgets, budgets + m->problems());
    p.first_level = policy->first_level;
    p.work_sort = policy->work_sort;
    p.topo_grid = policy->topo_grid;
    p.geom_grid = policy->geom_grid;
    p.wave_rows = policy->wave_rows;
    p.geom_lanes = policy->geom_lanes == 1 ? 1 : 2;
    return apply_policy(m, p);
}

int nat_mesh_get_policy(const nat_mesh *m, nat_policy *policy, uint8_t *budgets) {
    g_err.clear();
    if (!m || !policy) return fail(-1, "NULL argument");
    *policy = nat_policy{};
    policy->first_level = m->first_level;
    policy->work_sort = m->work_sort;
    policy->topo_grid = m->topo_grid;
    policy->geom_grid = m->geom_grid;
    policy->wave_rows = m->wave_rows;
    policy->geom_lanes = m->geom_lanes;
    if (budgets)
        for (int p = 0; p < m->problems(); p++)
            budgets[p] = p < static_cast<int>(m->budget.size()) ? m->budget[p] : kGroupBudget;
    return 0;
}

int nat_mesh_autotune(nat_mesh *m, const double *d_queries, int64_t num_query_points,
                      const int64_t *query_offsets, int repeats, void *stream,
                      nat_tune_report *report) {
    g_err.clear();
    if (!m || !report) return fail(-1, "NULL argument");
    if (repeats < 1 || repeats > 100) return fail(-1, "repeats %d outside 1..100", repeats);
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    const int P = m->problems();
    if (P > 1 && !query_offsets)
        return fail(-1, "a batch mesh needs query offsets");
    long long M = num_query_points;
    if (query_offsets) {
        if (int rc = check_offsets(m, query_offsets)) return rc;
        if (query_offsets[P] != M)
            return fail(-1, "query_offsets end at %lld, not at the %lld queries",
                        static_cast<long long>(query_offsets[P]), M);
    }
    if (M < 0 || M > 2147483647LL) return fail(-1, "query count %lld", M);
    if (M > 0 && !d_queries) return fail(-1, "NULL queries");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    *report = nat_tune_report{};
    if (M == 0) return 0;
    const int *d_prob = nullptr;
    if (int rc = problem_ids(m, query_offsets, M, st, &d_prob)) return rc;
    const bool timing = m->timing;
    m->timing = false;
    struct Restore {  // timing flag back on every exit
        nat_mesh *m;
        bool t;
        ~Restore() { m->timing = t; }
    } restore{m, timing};
    const int Mi = static_cast<int>(M);
    // on an error the caller's policy comes back
    const Policy orig = policy_of(m);
    bool tuned = false;
    struct Keep {
        nat_mesh *m;
        const Policy &p;
        const bool &tuned;
        ~Keep() {
            if (!tuned) (void)apply_policy(m, p);
        }
    } keep{m, orig, tuned};

    Policy def;
    def.budget.assign(P, static_cast<unsigned char>(kGroupBudget));
    // pilot under the default policy: also sizes the workspace
    double ms = 0;
    if (int rc = time_policy(m, def, d_queries, Mi, d_prob, st, 1, &ms)) return rc;
    // queries that left the group kernel, per problem (no record)
    std::vector<long long> roff(static_cast<size_t>(M));
    HIP_OR_FAIL(hipMemcpy(roff.data(), m->rec_off.p, sizeof(long long) * roff.size(),
                          hipMemcpyDeviceToHost));
    std::vector<long long> cnt(P, 0), left(P, 0);
    long long left_all = 0;
    for (int p = 0; p < P; p++) {
        const long long a = query_offsets ? query_offsets[p] : 0;
        const long long b = query_offsets ? query_offsets[p + 1] : M;
        cnt[p] = b - a;
        for (long long i = a; i < b; i++) left[p] += roff[i] < 0;
        left_all += left[p];
    }
    report->left_group = static_cast<double>(left_all) / static_cast<double>(M);

    double best_ms = 0;
    if (int rc = time_policy(m, def, d_queries, Mi, d_prob, st, repeats, &best_ms)) return rc;
    report->default_ms = best_ms;
    Policy best = def;
    int tried = 1;
    auto consider = [&](const Policy &c) -> int {
        double t = 0;
        if (int rc = time_policy(m, c, d_queries, Mi, d_prob, st, repeats, &t)) return rc;
        tried++;
        if (t < 0.99 * best_ms) {  // 1% margin against timing noise
            best_ms = t;
            best = c;
        }
        return 0;
    };
    // 1. route problems whose queries mostly leave the group kernel
    std::vector<std::vector<unsigned char>> seen;
    for (double thr : {0.5, 0.25, 0.1}) {
        Policy c = best;
        bool any = false;
        for (int p = 0; p < P; p++)
            if (cnt[p] > 0 && static_cast<double>(left[p]) > thr * static_cast<double>(cnt[p])) {
                c.budget[p] = 0;
                any = true;
            }
        if (!any || std::find(seen.begin(), seen.end(), c.budget) != seen.end()) continue;
        seen.push_back(c.budget);
        if (int rc = consider(c)) return rc;
    }
    // 2. the wave-per-query level, where thread-per-query rows start
    if (left_all > 0 || best.budget != def.budget) {
        Policy c = best;
        c.wave_rows = 0;
        if (int rc = consider(c)) return rc;
        c = best;
        c.first_level = 2;
        if (int rc = consider(c)) return rc;
    }
    // 3. smaller group budgets for the problems still in the group kernel:
    //    big cavities give up sooner
    if (left_all > 0) {
        const Policy base = best;
        for (int b : {64, 48}) {
            Policy c = base;
            for (int p = 0; p < P; p++)
                if (c.budget[p] > 0) c.budget[p] = static_cast<unsigned char>(b);
            if (int rc = consider(c)) return rc;
        }
    }
    // 4. geometry work order, 5. grids
    {
        Policy c = best;
        c.work_sort = 0;
        if (int rc = consider(c)) return rc;
    }
    {
        Policy c = best;
        c.topo_grid = 2;
        if (int rc = consider(c)) return rc;
    }
    {
        Policy c = best;
        c.geom_grid = 2;
        if (int rc = consider(c)) return rc;
    }
    // final compute under the chosen policy (nat_weights_emit may follow)
    if (int rc = time_policy(m, best, d_queries, Mi, d_prob, st, 1, &ms)) return rc;
    tuned = true;
    report->tuned_ms = best_ms;
    report->candidates = tried;
    for (int p = 0; p < P; p++)
        if (best.budget[p] == 0) {
            report->routed_problems++;
            report->routed_queries += cnt[p];
        }
    return 0;
}

int nat_mesh_set_timing(nat_mesh *m, int enable) {
    if (!m) return fail(-1, "NULL mesh");
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    if (enable && !m->ev[0]) {
        for (hipEvent_t &e : m->ev) HIP_OR_FAIL(hipEventCreate(&e));
    }
    m->timing = enable != 0;
    return 0;
}

int nat_mesh_get_timings(const nat_mesh *m, nat_timings *out) {
    if (!m || !out) return fail(-1, "NULL argument");
    *out = m->times;
    return 0;
}

void nat_free_host(void *p) { std::free(p); }

const char *nat_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// The entry point that allocates nothing (nat_weights_noalloc).
//
// Its workspace is `scratch`, a fixed device array that every GPU the library
// runs on holds from module load to process end.  The kernels get its device
// address, looked up once per device with hipGetSymbolAddress (the array's
// name is only a host-side stand-in).  Every call on a GPU shares that GPU's
// copy and leaves its contents behind; what must start at zero (counters, the
// thread-per-query level's tables and stamps) is cleared by fill kernels at
// each call, so calls may be captured in a graph.  Calls that overlap on one
// GPU -- from several threads or on several streams, on any mesh -- race on it.

namespace {

constexpr size_t kNoAllocBytes = static_cast<size_t>(64) << 20;
constexpr int kNoAllocMaxQueries = 32768;
constexpr int kNoAllocLocateThreads = 4096;  // k-d fallback stacks of the start-tet search
constexpr int kNoAllocWaves = 8;             // wave-per-query rows in flight
constexpr int kNoAllocLevel1Threads = 64;    // thread-per-query rows (level 1) in flight

static __device__ __align__(256) char scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;
char *g_noalloc_ptr[kXferMaxDevices] = {};

// Device address of this device's `scratch` (the device is current).
int noalloc_scratch(int device, char **out) {
    if (device < 0 || device >= kXferMaxDevices) return fail(-1, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-10, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_noalloc_ptr[device] = static_cast<char *>(p);
    }
    *out = g_noalloc_ptr[device];
    return 0;
}

// Bump allocation of the scratch for one call (256-byte aligned pieces).
struct Carve {
    char *base;
    size_t used = 0, cap;
    char *take(size_t bytes) {
        char *r = base + used;
        used += (bytes + 255) & ~static_cast<size_t>(255);
        return used <= cap ? r : nullptr;
    }
};

int noalloc_rows(nat_mesh *m, char *base, const double *d_q, int M, double *d_vals,
                 int32_t *d_cols, int32_t *d_rowptr, int64_t capacity, hipStream_t st,
                 int64_t *nnz_out) {
    const size_t Mz = static_cast<size_t>(M);
    int bits = 1;
    while (bits < 7 && (1ull << (3 * (bits + 1))) <= Mz * 2) bits++;
    const size_t scan_elems = nat::scan_tmp_elems(M) + Mz + 1;
    Carve c{base, 0, kNoAllocBytes};
    unsigned *ctr = reinterpret_cast<unsigned *>(c.take(256));
    int *order = reinterpret_cast<int *>(c.take(4 * Mz));
    int *geom_order = reinterpret_cast<int *>(c.take(4 * Mz));
    char *sortw = c.take(nat::morton_sort_bytes(M));
    long long *scan = reinterpret_cast<long long *>(c.take(8 * scan_elems));
    long long *row_off = reinterpret_cast<long long *>(c.take(8 * Mz));
    int *row_cnt = reinterpret_cast<int *>(c.take(4 * Mz));
    double *row_total = reinterpret_cast<double *>(c.take(8 * Mz));
    long long *rowptr64 = reinterpret_cast<long long *>(c.take(8 * (Mz + 1)));
    int *ovf0 = reinterpret_cast<int *>(c.take(4 * Mz));
    int *ovf1 = reinterpret_cast<int *>(c.take(4 * Mz));
    int *start = reinterpret_cast<int *>(c.take(4 * Mz));
    long long *rec_off = reinterpret_cast<long long *>(c.take(8 * Mz));
    // start-tet search: k-d fallback stacks for up to kNoAllocLocateThreads threads
    const int nthr0 = std::min(kNoAllocLocateThreads, (M + 255) / 256 * 256);
    char *sc0_base = c.take(nat::scratch_bytes_per_thread(m->caps[0]) * nthr0 + 16 * 256);
    // wave-per-query rows and thread-per-query rows of level 1
    char *wscr = c.take(nat::wave_scratch_size() * kNoAllocWaves);
    const size_t l1_bytes =
        nat::scratch_bytes_per_thread(m->caps[1]) * kNoAllocLevel1Threads + 16 * 256;
    char *sc1_base = c.take(l1_bytes);
    if (!sc1_base)
        return fail(-1, "nat_weights_noalloc: %d queries leave no room in the %zu-byte scratch",
                    M, kNoAllocBytes);
    // the rest: topology records (60 %) and the row pool (40 %, 12 bytes per entry)
    const size_t rest = kNoAllocBytes - c.used;
    const long long rec_cap = static_cast<long long>(rest * 6 / 10 / sizeof(int4));
    int4 *rec = reinterpret_cast<int4 *>(c.take(sizeof(int4) * static_cast<size_t>(rec_cap)));
    const long long pool_cap =
        static_cast<long long>((kNoAllocBytes - c.used - 2 * 256) / (sizeof(double) + sizeof(int)));
    double *pool_vals =
        reinterpret_cast<double *>(c.take(sizeof(double) * static_cast<size_t>(pool_cap)));
    int *pool_cols = reinterpret_cast<int *>(c.take(sizeof(int) * static_cast<size_t>(pool_cap)));
    if (!rec || !pool_vals || !pool_cols) return fail(-1, "nat_weights_noalloc: scratch carve");

    // counters as compute_rows lays them out
    unsigned long long *pool_used = reinterpret_cast<unsigned long long *>(ctr + 16);
    unsigned long long *rec_used = reinterpret_cast<unsigned long long *>(ctr + 18);
    HIP_OR_FAIL(nat::launch_zero_u32(ctr, 64, st));
    nat::DevMesh d = m->dev;  // the step-3 table is used when an earlier compute built it
    d.prob = nullptr;
    d.robust_ctr = reinterpret_cast<unsigned long long *>(ctr + 48);
    HIP_OR_FAIL(nat::launch_morton_sort(d_q, M, m->lp.as<nat::LocParams>(), nullptr, 1, bits,
                                        order, sortw, st));
    nat::RowOut out{};
    out.pool_vals = pool_vals;
    out.pool_cols = pool_cols;
    out.pool_cap = pool_cap;
    out.pool_used = pool_used;
    out.row_off = row_off;
    out.row_cnt = row_cnt;
    out.row_total = row_total;
    nat::Scratch sc0;
    nat::scratch_carve(sc0_base, m->caps[0], nthr0, &sc0);
    nat::Topo topo;
    topo.rec = rec;
    topo.rec_cap = rec_cap;
    topo.rec_used = rec_used;
    topo.rec_off = rec_off;
    HIP_OR_FAIL(nat::launch_locate(d, m->caps[0], sc0, nthr0 / 256, d_q, order, M, start, st));
    HIP_OR_FAIL(nat::launch_topo(d, m->topo_blocks * m->topo_grid, d_q, order, M, start,
                                 ctr + 32, topo, ctr + 8, ovf0, st));
    if (m->work_sort) HIP_OR_FAIL(nat::launch_work_sort(order, M, topo, geom_order, st));
    HIP_OR_FAIL(nat::launch_geom(
        d, m->nthr_geom / nat::kGeomThreads * m->geom_grid,
        m->geom_lanes == 2 ? geom_rows_grid(m) : 0, d_q, m->work_sort ? geom_order : order, M,
        ctr + 40, topo, out, m->mixed_tau, m->mixed_kappa,
        m->precision == 1 ? reinterpret_cast<unsigned long long *>(ctr + 20) : nullptr, st));
    // rows beyond the group kernel: a wave each, then one thread each (level 1);
    // both read their item counts on the device
    const unsigned *n_dev = ctr + 8;
    int *items = ovf0, *next = ovf1;
    if (m->wave_rows) {
        HIP_OR_FAIL(nat::launch_wave_rows(d, wscr, kNoAllocWaves, d_q, ovf0, 0, n_dev, start,
                                          ctr + 6, out, ovf1, ctr + 12, st));
        n_dev = ctr + 12;
        std::swap(items, next);
    }
    // the level's hash tables and per-thread stamps start from zero
    HIP_OR_FAIL(nat::launch_zero_u32(sc1_base, l1_bytes / 4, st));
    nat::Scratch sc1;
    nat::scratch_carve(sc1_base, m->caps[1], kNoAllocLevel1Threads, &sc1);
    out.ovf_list = next;
    out.ovf_count = ctr + 9;
    HIP_OR_FAIL(nat::launch_rows(d, m->caps[1], sc1, 1, 1, d_q, items, 0, n_dev, ctr + 1, out,
                                 st));
    // counters: [9] rows beyond level 1, u64 pool use at [16], records at [18]
    unsigned h[20];
    HIP_OR_FAIL(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    unsigned long long used = 0, rec_units = 0;
    std::memcpy(&used, h + 16, sizeof(used));
    std::memcpy(&rec_units, h + 18, sizeof(rec_units));
    if (static_cast<long long>(rec_units) > rec_cap)
        return fail(-8, "nat_weights_noalloc: topology records need %llu of %lld units of the "
                        "scratch (fewer queries per call)", rec_units, rec_cap);
    if (h[9] > 0)
        return fail(-4, "nat_weights_noalloc: %u queries need more than %d cavity tets "
                        "(use nat_weights_compute)", h[9], m->caps[1].cav);
    if (static_cast<long long>(used) > pool_cap)
        return fail(-8, "nat_weights_noalloc: rows need %llu of %lld pool entries of the "
                        "scratch (fewer queries per call)", used, pool_cap);
    *nnz_out = static_cast<int64_t>(used);
    if (static_cast<long long>(used) > capacity)
        return fail(-7, "nat_weights_noalloc: %llu entries, output capacity %lld", used,
                    static_cast<long long>(capacity));
    if (used > 0 && (!d_vals || !d_cols)) return fail(-1, "NULL output");
    HIP_OR_FAIL(nat::launch_scan_counts(row_cnt, rowptr64, M, scan, st));
    HIP_OR_FAIL(nat::launch_gather_csr(rowptr64, row_off, row_cnt, pool_vals, pool_cols,
                                       m->vorig.as<int>(), M, static_cast<long long>(used),
                                       d_vals, d_cols, d_rowptr, st));
    return 0;
}

}  // namespace

extern "C" int nat_weights_noalloc(nat_mesh *m, const double *d_queries, int num_query_points,
                                   double *d_vals, int32_t *d_cols, int32_t *d_rowptr,
                                   int64_t capacity, void *stream, int64_t *nnz_out) {
    g_err.clear();
    if (!m || !nnz_out || !d_rowptr) return fail(-1, "NULL argument");
    *nnz_out = 0;
    if (m->device < 0) return fail(-1, "host-only mesh has no device copy");
    if (m->problems() != 1) return fail(-1, "nat_weights_noalloc needs a single-problem mesh");
    if (num_query_points < 0 || capacity < 0) return fail(-1, "negative size");
    if (num_query_points > kNoAllocMaxQueries)
        return fail(-1, "nat_weights_noalloc takes at most %d queries per call (got %d)",
                    kNoAllocMaxQueries, num_query_points);
    if (num_query_points > 0 && !d_queries) return fail(-1, "NULL queries");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    char *base = nullptr;
    if (int rc = noalloc_scratch(m->device, &base)) return rc;
    if (num_query_points == 0) {
        // rowptr[0] = 0 by a fill kernel (no memset node if the caller captures)
        HIP_OR_FAIL(nat::launch_zero_u32(d_rowptr, 1, st));
        return 0;
    }
    return noalloc_rows(m, base, d_queries, num_query_points, d_vals, d_cols, d_rowptr, capacity,
                        st, nnz_out);
}
