// sphere_emu.cpp -- host run of the cavity search's certified in-sphere
// shortcut (natinterp3d_amd/csrc/sphere.h) against the plain FP64 insphere
// determinant (geom.h) it stands in for.  Test infrastructure only.
//
// usage: sphere_emu in.bin out.bin
//   in.bin:  int32 K, int32 Q, then K tets x 4 vertices x 3 float64 (as the
//            mesh stores them: the orientation rule of the device upload --
//            p0 and p1 swapped when orient3d(p0, p1, p2, p3) <= 0 -- is
//            applied here), then K x Q queries x 3 float64 (Q per tet)
//   out.bin: K x Q int8 shortcut results (1 inside, 0 outside, -1 not
//            certified), then K x Q int8 plain decisions (insphere > 0)
#include <cstdio>
#include <vector>

#include "sphere.h"

using namespace nat;

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int K = 0, Q = 0;
    if (std::fread(&K, 4, 1, f) != 1 || std::fread(&Q, 4, 1, f) != 1 || K < 0 || Q < 0) return 2;
    std::vector<double> T(static_cast<size_t>(K) * 12), E(static_cast<size_t>(K) * Q * 3);
    if (std::fread(T.data(), 8, T.size(), f) != T.size()) return 2;
    if (std::fread(E.data(), 8, E.size(), f) != E.size()) return 2;
    std::fclose(f);
    std::vector<signed char> fast(static_cast<size_t>(K) * Q), plain(static_cast<size_t>(K) * Q);
    for (int k = 0; k < K; k++) {
        const double *t = &T[static_cast<size_t>(k) * 12];
        d3 p[4];
        for (int i = 0; i < 4; i++) p[i] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
        if (orient3d(p[0], p[1], p[2], p[3]) <= 0) {
            const d3 x = p[0];
            p[0] = p[1];
            p[1] = x;
        }
        const SphereTet s = sphere_of(p[0], p[1], p[2], p[3], make_int4(-1, -1, -1, -1));
        for (int j = 0; j < Q; j++) {
            const size_t i = static_cast<size_t>(k) * Q + j;
            const d3 e = {E[3 * i], E[3 * i + 1], E[3 * i + 2]};
            fast[i] = static_cast<signed char>(sphere_side(s, e));
            plain[i] = insphere(p[0], p[1], p[2], p[3], e) > 0 ? 1 : 0;
        }
    }
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(fast.data(), 1, fast.size(), o);
    std::fwrite(plain.data(), 1, plain.size(), o);
    std::fclose(o);
    return 0;
}
