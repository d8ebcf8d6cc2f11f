// nat_capi.cpp -- the C ABI (include/natinterp3d_hip.h): mesh build and
// upload, device workspace, the capacity-level driver of the row kernel and
// CSR assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

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
    nat::PackedMesh host;
    nat::DevMesh dev{};
    DevBuf verts, tet_v, tet_n, kd, cell_start, cell_pts;
    double bbox[6] = {0, 0, 0, 1, 1, 1};
    std::mutex mu;

    // workspace
    nat::Caps caps[kLevels];
    int nthr[kLevels] = {0, 0, 0, 0};  // [0]: locate kernel threads
    int nthr_geom = 0;
    int topo_blocks = 0;
    DevBuf scratch[kLevels];
    DevBuf start;                      // start tet per query
    DevBuf rec, rec_off;               // topology records
    long long rec_cap = 0;
    double units_per_row = 40.0;
    DevBuf q;                    // host-path query upload
    DevBuf order, keys, hist;    // Morton bucketing
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
    long long last_nnz = 0;
    long long last_units = 0;  // record units of the last compute

    // optional kernel timing (HIP events on the launch stream)
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    nat_timings times{};

    int64_t device_bytes() const {
        return static_cast<int64_t>(verts.bytes + tet_v.bytes + tet_n.bytes + kd.bytes +
                                    cell_start.bytes + cell_pts.bytes);
    }
    int64_t workspace_bytes() const {
        size_t s = q.bytes + order.bytes + geom_order.bytes + keys.bytes + hist.bytes + scan.bytes +
                   row_off.bytes + row_cnt.bytes + rowptr64.bytes + row_total.bytes +
                   ovf[0].bytes +
                   ovf[1].bytes + counters.bytes + pool_vals.bytes + pool_cols.bytes +
                   csr_vals.bytes + csr_cols.bytes + csr_rowptr.bytes +
                   rec.bytes + rec_off.bytes + start.bytes;
        for (int l = 0; l < kLevels; l++) s += scratch[l].bytes;
        return static_cast<int64_t>(s);
    }
    void release_all() {
        DevBuf *all[] = {&verts,    &tet_v,     &tet_n,     &kd,       &cell_start, &cell_pts,
                         &q,        &order,     &geom_order, &keys,    &hist,       &scan,
                         &row_off,  &row_cnt,   &rowptr64,  &row_total, &ovf[0],    &ovf[1],
                         &counters, &pool_vals, &pool_cols, &csr_vals, &csr_cols,   &csr_rowptr,
                         &rec,      &rec_off,   &start};
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
    const int kd = m->host.kd_depth + 2;
    const int cus = std::max(1, m->num_cus);
    // level 0 (locate, topology and geometry kernels) sees every query; the
    // topology kernel's LDS capacities cover all but a handful of rows of
    // uniform data.  Level-0 scratch only holds the k-d fallback stacks of
    // the locate kernel.  Persistent grids: exactly the resident blocks.
    m->caps[0] = {kd, 1, 1, 1, 1, 1, 1};
    m->nthr[0] = cus * nat::locate_blocks_per_cu() * 256;
    m->topo_blocks = cus * nat::topo_blocks_per_cu();
    m->nthr_geom = cus * nat::geom_blocks_per_cu() * nat::kGeomThreads;
    // larger levels (fused kernel) only see the queries that overflowed
    m->caps[1] = {kd, 2048, 12, 1024, 2560, 1536, 12};
    m->nthr[1] = 8192;
    m->caps[2] = {kd, 32768, 16, 16384, 40960, 24576, 16};
    m->nthr[2] = 512;
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

int upload_mesh(nat_mesh *m) {
    const nat::PackedMesh &h = m->host;
    const int nv = h.n_data + 4, T = h.n_tets;
    std::vector<double4> verts(nv);
    for (int i = 0; i < nv; i++)
        verts[i] = make_double4(h.verts[3 * i], h.verts[3 * i + 1], h.verts[3 * i + 2], 0.0);
    std::vector<int4> tv(T), tn(T);
    std::vector<double4> kd(T);
    for (int t = 0; t < T; t++) {
        const int *p = &h.tet_v[4 * t];
        const nat::d3 a = {h.verts[3 * p[0]], h.verts[3 * p[0] + 1], h.verts[3 * p[0] + 2]};
        const nat::d3 b = {h.verts[3 * p[1]], h.verts[3 * p[1] + 1], h.verts[3 * p[1] + 2]};
        const nat::d3 c = {h.verts[3 * p[2]], h.verts[3 * p[2] + 1], h.verts[3 * p[2] + 2]};
        const nat::d3 d = {h.verts[3 * p[3]], h.verts[3 * p[3] + 1], h.verts[3 * p[3] + 2]};
        // orientation bit (natural_insertionfree.c:457-476), stored inverted in
        // bit 31 of p3 so that the common positive case reads unchanged
        const bool pos = nat::orient3d(a, b, c, d) > 0;
        tv[t] = make_int4(p[0], p[1], p[2], pos ? p[3] : (p[3] | static_cast<int>(0x80000000u)));
        tn[t] = make_int4(h.tet_n[4 * t], h.tet_n[4 * t + 1], h.tet_n[4 * t + 2],
                          h.tet_n[4 * t + 3]);
        int2 lr = make_int2(h.kd_left[t], h.kd_right[t]);
        double w;
        std::memcpy(&w, &lr, sizeof(w));
        kd[t] = make_double4(h.kd_pos[3 * t], h.kd_pos[3 * t + 1], h.kd_pos[3 * t + 2], w);
    }
    HIP_OR_FAIL(m->verts.ensure(sizeof(double4) * nv));
    HIP_OR_FAIL(m->tet_v.ensure(sizeof(int4) * std::max(T, 1)));
    HIP_OR_FAIL(m->tet_n.ensure(sizeof(int4) * std::max(T, 1)));
    HIP_OR_FAIL(m->kd.ensure(sizeof(double4) * std::max(T, 1)));
    HIP_OR_FAIL(hipMemcpy(m->verts.p, verts.data(), sizeof(double4) * nv, hipMemcpyHostToDevice));
    if (T > 0) {
        HIP_OR_FAIL(hipMemcpy(m->tet_v.p, tv.data(), sizeof(int4) * T, hipMemcpyHostToDevice));
        HIP_OR_FAIL(hipMemcpy(m->tet_n.p, tn.data(), sizeof(int4) * T, hipMemcpyHostToDevice));
        HIP_OR_FAIL(hipMemcpy(m->kd.p, kd.data(), sizeof(double4) * T, hipMemcpyHostToDevice));
    }
    nat::DevMesh &d = m->dev;
    d.verts = m->verts.as<double4>();
    d.tet_v = m->tet_v.as<int4>();
    d.tet_n = m->tet_n.as<int4>();
    d.kd = m->kd.as<double4>();
    for (int k = 0; k < 3; k++) {
        d.rmin[k] = h.kd_rect[k];
        d.rmax[k] = h.kd_rect[3 + k];
    }
    d.n_data = h.n_data;
    d.n_tets = T;
    {
        std::vector<int> start;
        std::vector<double4> pts;
        double g0[3], gh;
        int gn[3];
        build_grid(h, g0, gh, gn, start, pts);
        HIP_OR_FAIL(m->cell_start.ensure(sizeof(int) * start.size()));
        HIP_OR_FAIL(m->cell_pts.ensure(sizeof(double4) * pts.size()));
        HIP_OR_FAIL(hipMemcpy(m->cell_start.p, start.data(), sizeof(int) * start.size(),
                              hipMemcpyHostToDevice));
        HIP_OR_FAIL(hipMemcpy(m->cell_pts.p, pts.data(), sizeof(double4) * pts.size(),
                              hipMemcpyHostToDevice));
        d.cell_start = m->cell_start.as<int>();
        d.cell_pts = m->cell_pts.as<double4>();
        for (int k = 0; k < 3; k++) {
            d.g0[k] = g0[k];
            d.gn[k] = gn[k];
        }
        d.gh = gh;
        d.ginv = 1.0 / gh;
    }
    // bounding box of the data points: the Morton grid of the queries
    for (int k = 0; k < 3; k++) {
        m->bbox[k] = h.verts[k];
        m->bbox[3 + k] = h.verts[k];
    }
    for (int i = 0; i < h.n_data; i++)
        for (int k = 0; k < 3; k++) {
            m->bbox[k] = std::min(m->bbox[k], h.verts[3 * i + k]);
            m->bbox[3 + k] = std::max(m->bbox[3 + k], h.verts[3 * i + k]);
        }
    return 0;
}

int ensure_level_scratch(nat_mesh *m, int level) {
    // each of the (at most 16) arrays is rounded up to 256 bytes when carved
    const size_t per = nat::scratch_bytes_per_thread(m->caps[level]);
    const size_t need = per * static_cast<size_t>(m->nthr[level]) + 16 * 256;
    if (m->scratch[level].bytes >= need && m->scratch[level].p) return 0;
    HIP_OR_FAIL(m->scratch[level].ensure(need, /*zero=*/true));
    return 0;
}

int compute_rows(nat_mesh *m, const double *d_q, int M, hipStream_t st, long long *nnz_out) {
    m->last_M = -1;
    if (M == 0) {
        m->last_M = 0;
        m->last_nnz = 0;
        *nnz_out = 0;
        return 0;
    }
    const size_t Mz = static_cast<size_t>(M);
    // Morton bucketing: 3*bits-bit cells over the data bounding box
    int bits = 1;
    while (bits < 7 && (1ull << (3 * (bits + 1))) <= Mz * 2) bits++;
    const int nbuckets = 1 << (3 * bits);
    const size_t scan_elems =
        std::max(nat::scan_tmp_elems(M) + Mz + 1,
                 static_cast<size_t>(nbuckets) + 1 + nat::scan_tmp_elems(nbuckets));
    HIP_OR_FAIL(m->order.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->geom_order.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->keys.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->hist.ensure(sizeof(int) * nbuckets));
    HIP_OR_FAIL(m->scan.ensure(sizeof(long long) * scan_elems));
    HIP_OR_FAIL(m->row_off.ensure(sizeof(long long) * Mz));
    HIP_OR_FAIL(m->row_cnt.ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->row_total.ensure(sizeof(double) * Mz));
    HIP_OR_FAIL(m->rowptr64.ensure(sizeof(long long) * (Mz + 1)));
    HIP_OR_FAIL(m->ovf[0].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->ovf[1].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->counters.ensure(256));
    if (int rc = ensure_level_scratch(m, 0)) return rc;

    m->times = nat_timings{};
    mark(m, 0, st);
    HIP_OR_FAIL(nat::launch_morton_sort(d_q, M, m->bbox, bits, m->order.as<int>(),
                                        m->hist.as<int>(), m->keys.as<int>(),
                                        m->scan.as<long long>(), st));
    mark(m, 1, st);
    m->times.order_ms = elapsed(m, 0, 1);

    HIP_OR_FAIL(m->rec_off.ensure(sizeof(long long) * Mz));
    HIP_OR_FAIL(m->start.ensure(sizeof(int) * Mz));
    long long want = static_cast<long long>(std::ceil(m->nnz_per_row * M)) + 1024;
    long long want_rec = static_cast<long long>(std::ceil(m->units_per_row * M)) + 1024;
    for (int attempt = 0; attempt < 4; attempt++) {
        m->times.levels = 0;
        if (m->pool_cap < want) {
            HIP_OR_FAIL(m->pool_vals.ensure(sizeof(double) * want));
            HIP_OR_FAIL(m->pool_cols.ensure(sizeof(int) * want));
            m->pool_cap = want;
        }
        if (m->rec_cap < want_rec) {
            HIP_OR_FAIL(m->rec.ensure(sizeof(int4) * want_rec));
            m->rec_cap = want_rec;
        }
        // counters: [0..3] fused-level work, [4] topology work, [5] geometry
        // work, [8..11] overflow counts, u64 pool use at [16], records at [18]
        unsigned *ctr = m->counters.as<unsigned>();
        unsigned long long *pool_used = reinterpret_cast<unsigned long long *>(ctr + 16);
        unsigned long long *rec_used = reinterpret_cast<unsigned long long *>(ctr + 18);
        HIP_OR_FAIL(hipMemsetAsync(m->counters.p, 0, 256, st));
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
        nat::scratch_carve(m->scratch[0].p, m->caps[0], m->nthr[0], &sc0);
        nat::Topo topo;
        topo.rec = m->rec.as<int4>();
        topo.rec_cap = m->rec_cap;
        topo.rec_used = rec_used;
        topo.rec_off = m->rec_off.as<long long>();
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_locate(m->dev, m->caps[0], sc0, m->nthr[0] / 256, d_q,
                                       m->order.as<int>(), M, m->start.as<int>(), st));
        mark(m, 3, st);
        m->times.locate_ms = elapsed(m, 2, 3);
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_topo(m->dev, m->topo_blocks, d_q, m->order.as<int>(), M,
                                     m->start.as<int>(), ctr + 4, topo, ctr + 8,
                                     m->ovf[0].as<int>(), st));
        mark(m, 3, st);
        unsigned long long meta[2] = {0, 0};
        HIP_OR_FAIL(hipMemcpyAsync(meta, ctr + 16, sizeof(meta), hipMemcpyDeviceToHost, st));
        unsigned novf = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&novf, ctr + 8, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        m->times.topo_ms = elapsed(m, 2, 3);
        const long long units = static_cast<long long>(meta[1]);
        if (units > m->rec_cap) {
            want_rec = units + units / 8 + 1024;
            continue;
        }
        m->units_per_row = std::max(8.0, 1.125 * static_cast<double>(units) / M);
        m->last_units = units;
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_work_sort(m->order.as<int>(), M, topo, m->geom_order.as<int>(),
                                          st));
        HIP_OR_FAIL(nat::launch_geom(m->dev, m->nthr_geom / nat::kGeomThreads, d_q,
                                     m->geom_order.as<int>(), M, ctr + 5, topo, out, st));
        mark(m, 3, st);
        m->times.geom_ms = elapsed(m, 2, 3);
        m->times.rows_ms[0] = m->times.topo_ms + m->times.geom_ms;
        m->times.rows_items[0] = M;
        m->times.levels = 1;

        // larger levels: the queries that overflowed the previous one
        const int *items = m->ovf[0].as<int>();
        int n_items = static_cast<int>(novf);
        for (int level = 1; level < kLevels && n_items > 0; level++) {
            if (int rc = ensure_level_scratch(m, level)) return rc;
            nat::Scratch sc;
            nat::scratch_carve(m->scratch[level].p, m->caps[level], m->nthr[level], &sc);
            out.ovf_list = m->ovf[level & 1].as<int>();
            out.ovf_count = ctr + 8 + level;
            const int nblocks = (m->nthr[level] + 255) / 256;
            mark(m, 2, st);
            HIP_OR_FAIL(nat::launch_rows(m->dev, m->caps[level], sc, nblocks, d_q, items,
                                         n_items, ctr + level, out, st));
            mark(m, 3, st);
            unsigned nnext = 0;
            HIP_OR_FAIL(hipMemcpyAsync(&nnext, ctr + 8 + level, sizeof(unsigned),
                                       hipMemcpyDeviceToHost, st));
            HIP_OR_FAIL(hipStreamSynchronize(st));
            m->times.rows_ms[level] = elapsed(m, 2, 3);
            m->times.rows_items[level] = n_items;
            m->times.levels = level + 1;
            n_items = static_cast<int>(nnext);
            items = m->ovf[level & 1].as<int>();
        }
        if (n_items > 0)
            return fail(-4, "%d queries need more than %d cavity tets", n_items,
                        m->caps[kLevels - 1].cav);
        unsigned long long used = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&used, pool_used, sizeof(used), hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        if (static_cast<long long>(used) <= m->pool_cap) {
            m->last_M = M;
            m->last_nnz = static_cast<long long>(used);
            m->nnz_per_row = std::max(4.0, 1.125 * static_cast<double>(used) / M);
            *nnz_out = m->last_nnz;
            return 0;
        }
        want = static_cast<long long>(used) + 1024;
    }
    return fail(-5, "row pool did not converge");
}

// Saved topology: int64 header {magic, M, record units, N}, rec_off
// int64[M], the query order int32[M], the row totals float64[M] (8-byte
// aligned), then the records (16-byte aligned).
constexpr long long kTopoMagic = 0x4e41545450304c32LL;
size_t topo_order_off(long long M) { return 32 + 8 * static_cast<size_t>(M); }
size_t topo_total_off(long long M) {
    return (topo_order_off(M) + 4 * static_cast<size_t>(M) + 7) & ~static_cast<size_t>(7);
}
size_t topo_rec_off(long long M) {
    return (topo_total_off(M) + 8 * static_cast<size_t>(M) + 15) & ~static_cast<size_t>(15);
}

int grad_rows(nat_mesh *m, const char *blob, const double *d_q, int M, const int *d_rowptr,
              const double *d_w, const double *d_gy, const double *d_values, int D,
              double *d_dq, hipStream_t st) {
    long long hdr[4];
    HIP_OR_FAIL(hipMemcpyAsync(hdr, blob, sizeof(hdr), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    if (hdr[0] != kTopoMagic) return fail(-1, "d_topo is not a saved topology");
    if (hdr[1] != M)
        return fail(-1, "saved topology has %lld queries, not %d", hdr[1], M);
    if (hdr[3] != m->host.n_data) return fail(-1, "saved topology is of another mesh");
    const size_t Mz = static_cast<size_t>(M);
    HIP_OR_FAIL(m->ovf[0].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->ovf[1].ensure(sizeof(int) * Mz));
    HIP_OR_FAIL(m->counters.ensure(256));
    // counters: [0] record work, [1..3] fused-level work, [8..11] overflow counts
    unsigned *ctr = m->counters.as<unsigned>();
    HIP_OR_FAIL(hipMemsetAsync(m->counters.p, 0, 256, st));
    nat::Topo topo;
    topo.rec = reinterpret_cast<int4 *>(const_cast<char *>(blob) + topo_rec_off(M));
    topo.rec_cap = hdr[2];
    topo.rec_used = nullptr;
    topo.rec_off = reinterpret_cast<long long *>(const_cast<char *>(blob) + 32);
    const int *order = reinterpret_cast<const int *>(blob + topo_order_off(M));
    nat::GradIn g{d_gy, d_values, D, d_dq, d_rowptr, d_w,
                  reinterpret_cast<const double *>(blob + topo_total_off(M))};
    for (int l = 0; l < kLevels; l++) {
        m->times.grad_ms[l] = 0.0;
        m->times.grad_items[l] = 0;
    }
    mark(m, 2, st);
    HIP_OR_FAIL(nat::launch_grad_geom(m->dev, m->nthr_geom / nat::kGeomThreads, d_q, order, M,
                                      ctr, topo, g, m->ovf[0].as<int>(), ctr + 8, st));
    mark(m, 3, st);
    unsigned novf = 0;
    HIP_OR_FAIL(hipMemcpyAsync(&novf, ctr + 8, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    m->times.grad_ms[0] = elapsed(m, 2, 3);
    m->times.grad_items[0] = M;
    const int *items = m->ovf[0].as<int>();
    int n_items = static_cast<int>(novf);
    for (int level = 1; level < kLevels && n_items > 0; level++) {
        if (int rc = ensure_level_scratch(m, level)) return rc;
        nat::Scratch sc;
        nat::scratch_carve(m->scratch[level].p, m->caps[level], m->nthr[level], &sc);
        const int nblocks = (m->nthr[level] + 255) / 256;
        mark(m, 2, st);
        HIP_OR_FAIL(nat::launch_grad_rows(m->dev, m->caps[level], sc, nblocks, d_q, items,
                                          n_items, ctr + level, g, m->ovf[level & 1].as<int>(),
                                          ctr + 8 + level, st));
        mark(m, 3, st);
        unsigned nnext = 0;
        HIP_OR_FAIL(hipMemcpyAsync(&nnext, ctr + 8 + level, sizeof(unsigned),
                                   hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        m->times.grad_ms[level] = elapsed(m, 2, 3);
        m->times.grad_items[level] = n_items;
        n_items = static_cast<int>(nnext);
        items = m->ovf[level & 1].as<int>();
    }
    if (n_items > 0)
        return fail(-4, "%d queries need more than %d cavity tets", n_items,
                    m->caps[kLevels - 1].cav);
    return 0;
}

int emit_rows(nat_mesh *m, double *d_vals, int *d_cols, int *d_rowptr, hipStream_t st) {
    if (m->last_M < 0) return fail(-6, "nat_weights_emit without a preceding compute");
    if (m->last_nnz > 2147483647LL) return -1;
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
                                       m->pool_vals.as<double>(), m->pool_cols.as<int>(), M,
                                       d_vals, d_cols, d_rowptr, st));
    mark(m, 1, st);
    m->times.emit_ms = elapsed(m, 0, 1);
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
    if (device >= 0) {
        int ndev = 0;
        HIP_OR_FAIL(hipGetDeviceCount(&ndev));
        if (device >= ndev) return fail(-1, "device %d out of range (%d)", device, ndev);
    } else if (device != -1) {
        return fail(-1, "device %d invalid", device);
    }
    nat_mesh *m = new nat_mesh();
    m->device = device;
    const int rc = nat::build_packed_mesh(data_points, num_data_points, &m->host);
    if (rc) {
        delete m;
        if (rc == -2) return fail(-2, "mesh builder met a tet of non-positive orientation");
        if (rc == -3) return fail(-3, "mesh builder could not locate a point");
        return fail(rc, "mesh builder failed (%d)", rc);
    }
    if (device == -1) {  // host-only mesh: export and inspection, no queries
        *out = m;
        return 0;
    }
    DeviceGuard g(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete m;
        return fail(-10, "hipGetDeviceProperties failed");
    }
    m->num_cus = prop.multiProcessorCount;
    set_levels(m);
    const int urc = upload_mesh(m);
    if (urc) {
        m->release_all();
        delete m;
        return urc;
    }
    *out = m;
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
    info->n_data = m->host.n_data;
    info->n_tets = m->host.n_tets;
    info->kd_depth = m->host.kd_depth;
    info->coplanar = m->host.coplanar;
    info->cospherical = m->host.cospherical;
    info->edge_anomalies = m->host.edge_anomalies;
    info->device = m->device;
    info->num_cus = m->num_cus;
    info->device_bytes = m->device_bytes();
    info->workspace_bytes = m->workspace_bytes();
    return 0;
}

int nat_mesh_export(const nat_mesh *m, double *verts, int32_t *tet_v, int32_t *tet_n,
                    int32_t *kd_left, int32_t *kd_right, double *kd_pos, double *kd_rect) {
    if (!m) return fail(-1, "NULL mesh");
    const nat::PackedMesh &h = m->host;
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
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    long long nnz = 0;
    const int rc = compute_rows(m, d_queries, num_query_points,
                                static_cast<hipStream_t>(stream), &nnz);
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
    if (num_data_points != m->host.n_data)
        return fail(-2, "num_data_points %d does not match the mesh (%d)", num_data_points,
                    m->host.n_data);
    if (m->device < 0) return fail(-2, "host-only mesh has no device copy");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    const int M = num_query_points;
    const size_t Mz = static_cast<size_t>(M);
    hipStream_t st = nullptr;
    if (M > 0) {
        HIP_OR_FAIL(m->q.ensure(sizeof(double) * 3 * Mz));
        HIP_OR_FAIL(hipMemcpy(m->q.p, query_points, sizeof(double) * 3 * Mz,
                              hipMemcpyHostToDevice));
    }
    long long nnz = 0;
    int rc = compute_rows(m, m->q.as<double>(), M, st, &nnz);
    if (rc) return rc;
    if (nnz > 2147483647LL) return -1;
    const size_t nz = static_cast<size_t>(nnz);
    HIP_OR_FAIL(m->csr_vals.ensure(sizeof(double) * std::max<size_t>(nz, 1)));
    HIP_OR_FAIL(m->csr_cols.ensure(sizeof(int) * std::max<size_t>(nz, 1)));
    HIP_OR_FAIL(m->csr_rowptr.ensure(sizeof(int) * (Mz + 1)));
    rc = emit_rows(m, m->csr_vals.as<double>(), m->csr_cols.as<int>(), m->csr_rowptr.as<int>(),
                   st);
    if (rc) return rc;
    double *hv = static_cast<double *>(std::malloc(sizeof(double) * std::max<size_t>(nz, 1)));
    int *hc = static_cast<int *>(std::malloc(sizeof(int) * std::max<size_t>(nz, 1)));
    if (!hv || !hc) {
        std::free(hv);
        std::free(hc);
        return fail(-3, "host allocation of %zu entries failed", nz);
    }
    hipError_t e = hipMemcpy(weight_row_ptrs, m->csr_rowptr.p, sizeof(int) * (Mz + 1),
                             hipMemcpyDeviceToHost);
    if (e == hipSuccess && nz)
        e = hipMemcpy(hv, m->csr_vals.p, sizeof(double) * nz, hipMemcpyDeviceToHost);
    if (e == hipSuccess && nz)
        e = hipMemcpy(hc, m->csr_cols.p, sizeof(int) * nz, hipMemcpyDeviceToHost);
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
                 const void *d_values, int values_dtype, void *d_out, int out_dtype,
                 void *stream) {
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
                                    d_values, values_dtype, d_out, out_dtype,
                                    static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(-10, "spmm launch failed: %s", hipGetErrorString(e));
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
    const long long hdr[4] = {kTopoMagic, M, units, m->host.n_data};
    char *dst = static_cast<char *>(d_dst);
    HIP_OR_FAIL(hipMemcpyAsync(dst, hdr, sizeof(hdr), hipMemcpyHostToDevice, st));
    if (M > 0) {
        HIP_OR_FAIL(hipMemcpyAsync(dst + 32, m->rec_off.p, 8 * static_cast<size_t>(M),
                                   hipMemcpyDeviceToDevice, st));
        // the backward keeps the Morton order: its gathers (values, dL/dy)
        // favour spatial locality over grouping by cavity size
        HIP_OR_FAIL(hipMemcpyAsync(dst + topo_order_off(M), m->order.p,
                                   4 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        HIP_OR_FAIL(hipMemcpyAsync(dst + topo_total_off(M), m->row_total.p,
                                   8 * static_cast<size_t>(M), hipMemcpyDeviceToDevice, st));
        if (units > 0)
            HIP_OR_FAIL(hipMemcpyAsync(dst + topo_rec_off(M), m->rec.p,
                                       sizeof(int4) * static_cast<size_t>(units),
                                       hipMemcpyDeviceToDevice, st));
    }
    // the header comes from the stack: finish before returning
    HIP_OR_FAIL(hipStreamSynchronize(st));
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

int nat_topology_row_order(const void *d_topo, int num_query_points, const int32_t **out) {
    g_err.clear();
    if (!d_topo || !out) return fail(-1, "NULL argument");
    long long hdr[4];
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
    hipError_t e = nat::launch_spmm_t(num_rows, num_cols, dim, d_rowptr, d_cols, d_vals, d_gy,
                                      d_row_order, d_out, static_cast<hipStream_t>(stream));
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
