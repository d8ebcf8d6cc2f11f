// delaunay_kernels.hip -- the opt-in GPU Delaunay build: the kernels of one
// insertion round (delaunay.h holds the per-thread steps and the argument
// that a round equals a sequence of single insertions) and the host loop
// that runs rounds until every point is in.  Deterministic: candidates are
// taken in tet order, fresh tet slots at scanned offsets of the winners in
// candidate order, and every decision is an exact sign.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "delaunay.h"
#include "delaunay_build.h"
#include "mesh_builder.h"
#include "nat_internal.h"

namespace nat {
namespace {

using dl::State;

struct DevOps {
    __device__ static void amin(unsigned long long *p, unsigned long long v) { atomicMin(p, v); }
    __device__ static void add(unsigned *p, unsigned v) { atomicAdd(p, v); }
    __device__ static void fail(unsigned *p, unsigned v) { atomicOr(p, v); }
    __device__ static void add64(unsigned long long *p, unsigned long long v) {
        atomicAdd(p, v);
    }
};

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) dl_init_kernel(State s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.cap) {
        s.kround[i] = -1;
        s.mark[i] = 0;
    }
    if (i < s.n) s.loc[i] = 0;
    if (i == 0) {
        s.tv[0] = make_int4(s.n, s.n + 1, s.n + 2, s.n + 3);
        s.tn[0] = make_int4(-1, -1, -1, -1);
    }
}

__global__ void __launch_bounds__(kThreads) dl_reset_kernel(State s, unsigned live) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < live) dl::reset_tet(s, static_cast<int>(t));
    if (t == 0) *s.inserted = 0u;
}

__global__ void __launch_bounds__(kThreads) dl_offer_kernel(State s, int round, unsigned live) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < s.n) dl::offer<DevOps>(s, u, round, live);
}

__global__ void __launch_bounds__(kThreads) dl_flag_kernel(State s, unsigned live) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < live) dl::flag_tet(s, static_cast<int>(t));
}

__global__ void __launch_bounds__(kThreads) dl_pick_kernel(State s, unsigned live) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < live) dl::pick(s, static_cast<int>(t));
}

__global__ void __launch_bounds__(kThreads) dl_claim_kernel(State s, int round, unsigned live) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= s.slots) return;
    if (sl >= dl::n_cand(s, live)) {
        s.slot_c[sl] = 0;
        s.slot_o[sl] = 0;
        return;
    }
    dl::claim<DevOps>(s, sl, round, live);
}

__global__ void __launch_bounds__(kThreads) dl_win_kernel(State s, int round) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < s.slots) dl::win<DevOps>(s, sl, round);
}

__global__ void __launch_bounds__(kThreads) dl_faces_kernel(State s, int round) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < s.slots) dl::faces<DevOps>(s, sl, round);
}

__global__ void __launch_bounds__(kThreads) dl_commit_kernel(State s, int round, unsigned live) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < s.slots) dl::commit<DevOps>(s, sl, round, live);
}

__global__ void __launch_bounds__(kThreads) dl_relocate_kernel(State s, int round) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < s.n) dl::relocate<DevOps>(s, u, round);
}

__global__ void __launch_bounds__(kThreads) dl_audit_kernel(const double4 *pts, const int4 *tv,
                                                            const int4 *tn, int T, unsigned *err,
                                                            unsigned long long *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T) dl::audit_tet<DevOps>(pts, tv, tn, T, t, err, out);
}

unsigned blocks(long long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(size_t n) {
        return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

#define DL_HIP(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            if (msg) *msg = std::string(#x) + ": " + hipGetErrorString(e_);          \
            return -10;                                                              \
        }                                                                            \
    } while (0)

std::string error_text(unsigned e) {
    std::string s;
    auto add = [&](unsigned bit, const char *t) {
        if (e & bit) s += (s.empty() ? "" : "; ") + std::string(t);
    };
    add(dl::kErrCavity, "a cavity and its ring exceeded 2048 tets");
    add(dl::kErrRepeated, "a repeated point (empty cavity)");
    add(dl::kErrFlat, "a point on the plane of a cavity face (flat tet)");
    add(dl::kErrTets, "tet capacity exhausted");
    add(dl::kErrWalk, "a point location walk did not end");
    add(dl::kErrOutside, "a point outside the super tet");
    add(dl::kErrWide, "a predicate too wide for the exact evaluator");
    add(dl::kErrShape, "inconsistent cavity boundary");
    return s;
}

}  // namespace

int gpu_build_tets(const double *points, int n, GpuBuild *out, std::string *msg) {
    if (!points || n <= 0 || !out) {
        if (msg) *msg = "bad arguments";
        return -1;
    }
    if (n > (1 << 26)) {
        if (msg) *msg = "at most 2^26 points";
        return -1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> X(static_cast<size_t>(n + 4) * 3);
    std::copy(points, points + 3 * static_cast<size_t>(n), X.begin());
    super_vertices(points, n, X.data() + 3 * static_cast<size_t>(n));
    {
        const d3 a = {X[3 * n], X[3 * n + 1], X[3 * n + 2]};
        const d3 b = {X[3 * n + 3], X[3 * n + 4], X[3 * n + 5]};
        const d3 c = {X[3 * n + 6], X[3 * n + 7], X[3 * n + 8]};
        const d3 d = {X[3 * n + 9], X[3 * n + 10], X[3 * n + 11]};
        if (orient3d_exact_sign(a, b, c, d) != 1) {
            if (msg) *msg = "degenerate bounding box (the super tet has no volume)";
            return -3;
        }
    }
    std::vector<double4> P(static_cast<size_t>(n) + 4);
    for (int i = 0; i < n + 4; i++) P[i] = make_double4(X[3 * i], X[3 * i + 1], X[3 * i + 2], 0.0);

    const int cap = static_cast<int>(std::min<long long>(8LL * n + 4096, 2147483000LL));
    const int slots = std::max(1, std::min(16384, n));
    const long long face_cap = std::max<long long>(65536, 64LL * slots);
    Buf pts, tv, tn, best, owner, mark, kround, flag, flag_off, flag_tmp, loc, cand, slot_c,
        slot_o, slot_b, face_off, faces, need, need_off, slot_tmp, scr, err, inserted;
    DL_HIP(pts.alloc<double4>(P.size()));
    DL_HIP(tv.alloc<int4>(cap));
    DL_HIP(tn.alloc<int4>(cap));
    DL_HIP(best.alloc<unsigned long long>(cap));
    DL_HIP(owner.alloc<unsigned long long>(cap));
    DL_HIP(mark.alloc<int>(cap));
    DL_HIP(kround.alloc<int>(cap));
    DL_HIP(flag.alloc<int>(cap));
    DL_HIP(flag_off.alloc<long long>(static_cast<size_t>(cap) + 1));
    DL_HIP(flag_tmp.alloc<long long>(scan_tmp_elems(cap)));
    DL_HIP(loc.alloc<int>(n));
    DL_HIP(cand.alloc<int>(slots));
    DL_HIP(slot_c.alloc<int>(slots));
    DL_HIP(slot_o.alloc<int>(slots));
    DL_HIP(slot_b.alloc<int>(slots));
    DL_HIP(face_off.alloc<long long>(static_cast<size_t>(slots) + 1));
    DL_HIP(faces.alloc<int>(static_cast<size_t>(face_cap) * dl::kFaceInts));
    DL_HIP(need.alloc<int>(slots));
    DL_HIP(need_off.alloc<long long>(static_cast<size_t>(slots) + 1));
    DL_HIP(slot_tmp.alloc<long long>(scan_tmp_elems(slots)));
    DL_HIP(scr.alloc<int>(static_cast<size_t>(slots) * dl::kSlotInts));
    DL_HIP(err.alloc<unsigned>(1));
    DL_HIP(inserted.alloc<unsigned>(1));
    hipStream_t st = nullptr;
    DL_HIP(hipMemcpyAsync(pts.p, P.data(), P.size() * sizeof(double4), hipMemcpyHostToDevice, st));
    DL_HIP(hipMemsetAsync(err.p, 0, sizeof(unsigned), st));
    // the visited sets' tags start at 1: zeroed tags are empty
    DL_HIP(hipMemsetAsync(scr.p, 0, static_cast<size_t>(slots) * dl::kSlotInts * sizeof(int), st));

    State s{};
    s.pts = pts.as<double4>();
    s.n = n;
    s.tv = tv.as<int4>();
    s.tn = tn.as<int4>();
    s.cap = cap;
    s.loc = loc.as<int>();
    s.best = best.as<unsigned long long>();
    s.owner = owner.as<unsigned long long>();
    s.mark = mark.as<int>();
    s.kround = kround.as<int>();
    s.flag = flag.as<int>();
    s.flag_off = flag_off.as<long long>();
    s.cand = cand.as<int>();
    s.slots = slots;
    s.slot_c = slot_c.as<int>();
    s.slot_o = slot_o.as<int>();
    s.slot_b = slot_b.as<int>();
    s.face_off = face_off.as<long long>();
    s.faces = faces.as<int>();
    s.face_cap = face_cap;
    s.need = need.as<int>();
    s.need_off = need_off.as<long long>();
    s.scr = scr.as<int>();
    s.err = err.as<unsigned>();
    s.inserted = inserted.as<unsigned>();

    hipLaunchKernelGGL(dl_init_kernel, dim3(blocks(std::max(cap, n))), dim3(kThreads), 0, st, s);
    DL_HIP(hipGetLastError());
    unsigned live = 1;
    long long remaining = n;
    int round = 0;
    const int max_rounds = 1 << 20;
    while (remaining > 0) {
        round++;
        if (round > max_rounds) {
            if (msg) *msg = "GPU Delaunay build did not converge";
            return -3;
        }
        hipLaunchKernelGGL(dl_reset_kernel, dim3(blocks(live)), dim3(kThreads), 0, st, s, live);
        hipLaunchKernelGGL(dl_offer_kernel, dim3(blocks(n)), dim3(kThreads), 0, st, s, round,
                           live);
        hipLaunchKernelGGL(dl_flag_kernel, dim3(blocks(live)), dim3(kThreads), 0, st, s, live);
        DL_HIP(launch_scan_counts(s.flag, s.flag_off, static_cast<int>(live),
                                  flag_tmp.as<long long>(), st));
        hipLaunchKernelGGL(dl_pick_kernel, dim3(blocks(live)), dim3(kThreads), 0, st, s, live);
        hipLaunchKernelGGL(dl_claim_kernel, dim3(blocks(slots)), dim3(kThreads), 0, st, s, round,
                           live);
        hipLaunchKernelGGL(dl_win_kernel, dim3(blocks(slots)), dim3(kThreads), 0, st, s, round);
        DL_HIP(launch_scan_counts(s.slot_b, s.face_off, slots, slot_tmp.as<long long>(), st));
        hipLaunchKernelGGL(dl_faces_kernel, dim3(blocks(slots)), dim3(kThreads), 0, st, s,
                           round);
        DL_HIP(launch_scan_counts(s.need, s.need_off, slots, slot_tmp.as<long long>(), st));
        hipLaunchKernelGGL(dl_commit_kernel, dim3(blocks(slots)), dim3(kThreads), 0, st, s, round,
                           live);
        hipLaunchKernelGGL(dl_relocate_kernel, dim3(blocks(n)), dim3(kThreads), 0, st, s, round);
        DL_HIP(hipGetLastError());
        unsigned e = 0, ins = 0;
        long long fresh = 0;
        DL_HIP(hipMemcpyAsync(&e, s.err, sizeof(e), hipMemcpyDeviceToHost, st));
        DL_HIP(hipMemcpyAsync(&ins, s.inserted, sizeof(ins), hipMemcpyDeviceToHost, st));
        DL_HIP(hipMemcpyAsync(&fresh, s.need_off + slots, sizeof(fresh), hipMemcpyDeviceToHost,
                              st));
        DL_HIP(hipStreamSynchronize(st));
        if (e) {
            if (msg) *msg = "GPU Delaunay build: " + error_text(e) + " (use the host builder)";
            return (e & (dl::kErrRepeated | dl::kErrFlat)) ? -3 : -4;
        }
        if (ins == 0) {
            if (msg) *msg = "GPU Delaunay build made no progress";
            return -4;
        }
        live += static_cast<unsigned>(fresh);
        remaining -= ins;
    }
    out->rounds = round;
    out->n = n;
    out->n_tets = static_cast<int>(live);
    out->verts = std::move(X);
    out->tet_v.resize(static_cast<size_t>(live) * 4);
    out->tet_n.resize(static_cast<size_t>(live) * 4);
    DL_HIP(hipMemcpyAsync(out->tet_v.data(), s.tv, sizeof(int4) * live, hipMemcpyDeviceToHost,
                          st));
    DL_HIP(hipMemcpyAsync(out->tet_n.data(), s.tn, sizeof(int4) * live, hipMemcpyDeviceToHost,
                          st));
    DL_HIP(hipStreamSynchronize(st));
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int gpu_delaunay_audit(const double *verts, int nv, const int32_t *tet_v, const int32_t *tet_n,
                       int T, long long counts[4], std::string *msg) {
    Buf pts, tv, tn, out, err;
    std::vector<double4> P(nv);
    for (int i = 0; i < nv; i++)
        P[i] = make_double4(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2], 0.0);
    for (int i = 0; i < 4 * T; i++)
        if (tet_v[i] < 0 || tet_v[i] >= nv) {
            if (msg) *msg = "vertex slot out of range";
            return -1;
        }
    DL_HIP(pts.alloc<double4>(nv));
    DL_HIP(tv.alloc<int4>(T));
    DL_HIP(tn.alloc<int4>(T));
    DL_HIP(out.alloc<unsigned long long>(4));
    DL_HIP(err.alloc<unsigned>(1));
    hipStream_t st = nullptr;
    DL_HIP(hipMemcpyAsync(pts.p, P.data(), sizeof(double4) * nv, hipMemcpyHostToDevice, st));
    DL_HIP(hipMemcpyAsync(tv.p, tet_v, sizeof(int4) * T, hipMemcpyHostToDevice, st));
    DL_HIP(hipMemcpyAsync(tn.p, tet_n, sizeof(int4) * T, hipMemcpyHostToDevice, st));
    DL_HIP(hipMemsetAsync(out.p, 0, 4 * sizeof(unsigned long long), st));
    DL_HIP(hipMemsetAsync(err.p, 0, sizeof(unsigned), st));
    if (T > 0)
        hipLaunchKernelGGL(dl_audit_kernel, dim3(blocks(T)), dim3(kThreads), 0, st,
                           pts.as<double4>(), tv.as<int4>(), tn.as<int4>(), T, err.as<unsigned>(),
                           out.as<unsigned long long>());
    DL_HIP(hipGetLastError());
    unsigned long long h[4];
    unsigned e = 0;
    DL_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, st));
    DL_HIP(hipMemcpyAsync(&e, err.p, sizeof(e), hipMemcpyDeviceToHost, st));
    DL_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < 4; i++) counts[i] = static_cast<long long>(h[i]);
    if (e & dl::kErrWide) {
        if (msg) *msg = "a predicate too wide for the exact evaluator";
        return -4;
    }
    return 0;
}

}  // namespace nat
