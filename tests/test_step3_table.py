"""Addressing of the step-3 table (csrc/step3_table.h), on the CPU.

table_pos(fcmask, dm) must place the piece of each reduced mask m = fcmask &
dm at its rank among the subsets of dm ordered by (number of bits, value):
checked against that definition by enumeration for all 256 (dm, fcmask)
pairs, together with what the layout is for -- the subsets of dm fill
positions 0 .. 2^popcount(dm) - 1 exactly once, bits of fcmask outside dm do
not move a piece, and the empty mask and single bits come first.  The
circumsphere record's d1 leaves its low four bits free for the table's mask
(sphere_dm is 0 before the table is built).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def emu_out(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("clang++ of the ROCm toolchain not found")
    exe = str(tmp_path_factory.mktemp("s3") / "step3_table_emu")
    cmd = [CLANG, "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I", os.path.join(REPO, "natinterp3d_amd", "csrc"),
           os.path.join(HERE, "emu", "step3_table_emu.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")


def positions(lines):
    pos = {}
    for ln in lines:
        f = ln.split()
        if len(f) == 3:
            pos[int(f[0]), int(f[1])] = int(f[2])
    return pos


def by_definition(m, dm):
    key = lambda s: (bin(s).count("1"), s)  # noqa: E731
    return sum(1 for s in range(16) if s & ~dm == 0 and key(s) < key(m & dm))


def test_positions_match_the_definition(emu_out):
    pos = positions(emu_out)
    assert len(pos) == 256
    for (dm, m), p in pos.items():
        assert p == by_definition(m, dm), (dm, m, p)


def test_subsets_fill_the_first_pieces_once(emu_out):
    pos = positions(emu_out)
    for dm in range(16):
        k = bin(dm).count("1")
        got = sorted(pos[dm, s] for s in range(16) if s & ~dm == 0)
        assert got == list(range(2 ** k)), dm
        for m in range(16):
            assert pos[dm, m] == pos[dm, m & dm]
        # empty mask first, then the single bits in bit order
        assert pos[dm, 0] == 0
        singles = [pos[dm, 1 << b] for b in range(4) if dm >> b & 1]
        assert singles == list(range(1, k + 1))


def test_a_planted_mistake_is_caught():
    """ranking the masks by value alone (no popcount order) breaks the
    first-line property the checks above rely on"""
    def by_value(m, dm):
        return sum(1 for s in range(16) if s & ~dm == 0 and s < (m & dm))
    bad = [dm for dm in range(16)
           if [by_value(1 << b, dm) for b in range(4) if dm >> b & 1]
           != list(range(1, bin(dm).count("1") + 1))]
    assert bad, "the planted mistake went unnoticed"


def test_circumsphere_record_leaves_the_mask_bits_free(emu_out):
    dms = [int(ln.split()[1]) for ln in emu_out if ln.startswith("dm ")]
    assert len(dms) == 1000 and all(d == 0 for d in dms)
