// step3_table.h -- addressing of the step-3 table (DevMesh::old_vol).
//
// The table holds, per tet t and vertex VI, the old subcell volume of VI
// (natural_insertionfree.c:642-676) for each choice of vertex order of the
// three face circumcentres VI's subcell uses: each face's circumcentre in the
// tet's own ascending order or, when an earlier cavity tet holds the face, in
// that neighbour's order (the reference's face cache keeps the first
// caller's value, natural_insertionfree.c:233-253).  A cavity tet's face mask
// fcmask (bit K: face K is held by an earlier cavity tet) selects the values.
//
// A face's two circumcentres often have the same bits: cc_tri(a, b, c) and
// cc_tri(a, c, b) round identically (the cross products only change sign),
// so the two orders differ only when they start at different vertices.  The
// tet's mask dm marks the faces whose two values differ, and fcmask matters
// only through m = fcmask & dm.  The table keeps one 32-byte piece per m (the
// four vertices' values; bit VI of a mask does not change vertex VI's value)
// at position table_pos(fcmask, dm): the subsets of dm ordered by number of
// bits, then by value -- 16 pieces (512 bytes) per tet, of which the first
// 2^popcount(dm) are used.  The empty mask and the single bits, the masks of
// most cavity tets (the first has none, most others share one face with an
// earlier cavity tet), then sit in the tet's first 128-byte line.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace nat {

// Position of fcmask's piece among the subsets of dm (dm < 16):
// #{s subset of dm: popcount(s) < popcount(m)} plus the rank of m among the
// subsets of dm with its popcount in increasing value.  For p-subsets that
// rank is the combinatorial number system's: with b_1 < ... < b_p the indices
// of m's bits among dm's bits, sum_i C(b_i, i).
__host__ __device__ __forceinline__ unsigned table_pos(unsigned fcmask, unsigned dm) {
    dm &= 15u;
    const unsigned m = fcmask & dm;
    // C(n, r) for n = 0..3, r = 1..3 as 4-bit fields at (r - 1) * 16 + n * 4
    constexpr unsigned long long kBinom = 0x3210ull | (0x3100ull << 16) | (0x1000ull << 32);
    unsigned rank = 0, idx = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned in_dm = (dm >> k) & 1u;
        const unsigned in_m = (m >> k) & 1u;
        if (in_m) {
            rank += cnt < 3 ? static_cast<unsigned>((kBinom >> (cnt * 16 + idx * 4)) & 15ull) : 0u;
            cnt++;
        }
        idx += in_dm;
    }
    // #{s subset of dm: popcount(s) < cnt} = sum_{j < cnt} C(idx, j), idx = popcount(dm)
    unsigned below = cnt;  // idx <= 1, or cnt <= 1
    if (cnt >= 2) {
        if (idx == 2) below = 3;
        else if (idx == 3) below = cnt == 2 ? 4u : 7u;
        else if (idx == 4) below = cnt == 2 ? 5u : (cnt == 3 ? 11u : 15u);
    }
    return below + rank;
}

// index of vertex vi's value at table position pos
__host__ __device__ __forceinline__ size_t old_vol_index(int t, unsigned pos, int vi) {
    return static_cast<size_t>(t) * 64 + static_cast<size_t>(pos & 15u) * 4 +
           static_cast<size_t>(vi);
}

}  // namespace nat
