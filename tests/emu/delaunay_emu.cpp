// delaunay_emu.cpp -- sequential host run of the GPU Delaunay build's
// per-thread steps (natinterp3d_amd/csrc/delaunay.h), one "thread" after the
// other, with the round loop of delaunay_kernels.hip.  Test infrastructure:
// it checks the build's logic and indexing on the CPU (built with the
// address sanitizer by tests/test_gpu_build.py), never runs in the product.
//
// usage: delaunay_emu in.bin out.bin [slots]
//   in.bin:  int32 n, then n x 3 float64
//   out.bin: int32 rc, err, rounds, T, then tet_v [T x 4], tet_n [T x 4]
//            (int32), then int64 audit counts [4]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "delaunay.h"
#include "mesh_builder.h"

using namespace nat;
using dl::HostOps;

static void scan(const int *in, long long *out, int n) {
    long long acc = 0;
    for (int i = 0; i < n; i++) {
        out[i] = acc;
        acc += in[i];
    }
    out[n] = acc;
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n <= 0) return 2;
    std::vector<double> X(static_cast<size_t>(n + 4) * 3);
    if (std::fread(X.data(), 8, 3 * static_cast<size_t>(n), f) != 3 * static_cast<size_t>(n))
        return 2;
    std::fclose(f);
    const int slots_arg = argc > 3 ? std::atoi(argv[3]) : 16384;
    super_vertices(X.data(), n, X.data() + 3 * static_cast<size_t>(n));
    std::vector<double4> P(static_cast<size_t>(n) + 4);
    for (int i = 0; i < n + 4; i++) P[i] = make_double4(X[3 * i], X[3 * i + 1], X[3 * i + 2], 0.0);

    const int cap = 8 * n + 4096;
    const int slots = std::max(1, std::min(slots_arg, n));
    std::vector<int4> tv(cap), tn(cap);
    std::vector<unsigned long long> best(cap), owner(cap);
    const long long face_cap = std::max<long long>(65536, 64LL * slots);
    std::vector<int> mark(cap, 0), kround(cap, -1), flag(cap), loc(n, 0), cand(slots),
        slot_c(slots), slot_o(slots), slot_b(slots), need(slots);
    std::vector<int> faces(static_cast<size_t>(face_cap) * dl::kFaceInts);
    std::vector<long long> flag_off(static_cast<size_t>(cap) + 1), need_off(slots + 1),
        face_off(slots + 1);
    std::vector<int> scr(static_cast<size_t>(slots) * dl::kSlotInts, 0);
    unsigned err = 0, inserted = 0;
    dl::State s{};
    s.pts = P.data();
    s.n = n;
    s.tv = tv.data();
    s.tn = tn.data();
    s.cap = cap;
    s.loc = loc.data();
    s.best = best.data();
    s.owner = owner.data();
    s.mark = mark.data();
    s.kround = kround.data();
    s.flag = flag.data();
    s.flag_off = flag_off.data();
    s.cand = cand.data();
    s.slots = slots;
    s.slot_c = slot_c.data();
    s.slot_o = slot_o.data();
    s.slot_b = slot_b.data();
    s.face_off = face_off.data();
    s.faces = faces.data();
    s.face_cap = face_cap;
    s.need = need.data();
    s.need_off = need_off.data();
    s.scr = scr.data();
    s.err = &err;
    s.inserted = &inserted;
    tv[0] = make_int4(n, n + 1, n + 2, n + 3);
    tn[0] = make_int4(-1, -1, -1, -1);

    unsigned live = 1;
    long long remaining = n;
    int round = 0, rc = 0;
    while (remaining > 0) {
        round++;
        inserted = 0;
        for (unsigned t = 0; t < live; t++) dl::reset_tet(s, t);
        for (int u = 0; u < n; u++) dl::offer<HostOps>(s, u, round, live);
        for (unsigned t = 0; t < live; t++) dl::flag_tet(s, t);
        scan(flag.data(), flag_off.data(), static_cast<int>(live));
        for (unsigned t = 0; t < live; t++) dl::pick(s, t);
        const int nc = dl::n_cand(s, live);
        for (int sl = 0; sl < slots; sl++) {
            if (sl >= nc) {
                slot_c[sl] = 0;
                slot_o[sl] = 0;
                continue;
            }
            dl::claim<HostOps>(s, sl, round, live);
        }
        for (int sl = 0; sl < slots; sl++) dl::win<HostOps>(s, sl, round);
        scan(slot_b.data(), face_off.data(), slots);
        for (int sl = 0; sl < slots; sl++) dl::faces<HostOps>(s, sl, round);
        scan(need.data(), need_off.data(), slots);
        for (int sl = 0; sl < slots; sl++) dl::commit<HostOps>(s, sl, round, live);
        for (int u = 0; u < n; u++) dl::relocate<HostOps>(s, u, round);
        if (err) {
            rc = -3;
            break;
        }
        if (inserted == 0) {
            rc = -4;
            break;
        }
        live += static_cast<unsigned>(need_off[slots]);
        remaining -= inserted;
        if (std::getenv("DL_EMU_VERBOSE"))
            std::fprintf(stderr, "round %d candidates %d inserted %u live %u remaining %lld\n",
                         round, nc, inserted, live, remaining);
    }
    unsigned long long audit[4] = {0, 0, 0, 0};
    if (rc == 0)
        for (unsigned t = 0; t < live; t++)
            dl::audit_tet<HostOps>(P.data(), tv.data(), tn.data(), static_cast<int>(live), t, &err,
                                   audit);
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    const int T = rc == 0 ? static_cast<int>(live) : 0;
    const int hdr[4] = {rc, static_cast<int>(err), round, T};
    std::fwrite(hdr, 4, 4, o);
    std::fwrite(tv.data(), 16, T, o);
    std::fwrite(tn.data(), 16, T, o);
    const long long a[4] = {static_cast<long long>(audit[0]), static_cast<long long>(audit[1]),
                            static_cast<long long>(audit[2]), static_cast<long long>(audit[3])};
    std::fwrite(a, 8, 4, o);
    std::fclose(o);
    return 0;
}
