// step3_table_emu.cpp -- host run of the step-3 table's addressing
// (natinterp3d_amd/csrc/step3_table.h) and of the face-mask bits of the
// circumsphere record (sphere.h).  Test infrastructure only.
//
// usage: step3_table_emu
//   prints 256 lines "dm fcmask table_pos", then for a few tets the low
//   four bits of sphere_of's d1 (sphere_dm before the table writes a mask)
#include <cstdio>

#include "sphere.h"
#include "step3_table.h"

using namespace nat;

int main() {
    for (unsigned dm = 0; dm < 16; dm++)
        for (unsigned m = 0; m < 16; m++) std::printf("%u %u %u\n", dm, m, table_pos(m, dm));
    unsigned x = 12345u;
    auto rnd = [&]() {
        x = x * 1664525u + 1013904223u;
        return static_cast<double>(x >> 8) / 16777216.0;
    };
    for (int k = 0; k < 1000; k++) {
        d3 p[4];
        for (auto &v : p) v = {rnd(), rnd(), rnd()};
        if (orient3d(p[0], p[1], p[2], p[3]) <= 0) {
            const d3 t = p[0];
            p[0] = p[1];
            p[1] = t;
        }
        const SphereTet s = sphere_of(p[0], p[1], p[2], p[3], make_int4(-1, -1, -1, -1));
        std::printf("dm %u\n", sphere_dm(s));
    }
    return 0;
}
