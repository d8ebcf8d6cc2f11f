"""The static error bound behind the robustness audit (nat_insphere_classify,
host side): whenever it certifies a sign, the FP64 determinant has the sign
of the exact determinant (rational arithmetic); exactly cospherical inputs
are never certified; well-separated inputs always are.

The exact signs of the robust predicate mode (nat_predicate_exact, the same
code the device runs for uncertified tests) against independent integer
determinants, Fractions and closed forms."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from natinterp3d_amd import _lib


def classify(pts):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 15)
    n = pts.shape[0]
    det = np.zeros(n)
    unc = np.zeros(n, np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = _lib.lib().nat_insphere_classify(pts.ctypes.data_as(dp), n, det.ctypes.data_as(dp),
                                          unc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == 0
    return det, unc.astype(bool)


def exact_sign(p):
    a, b, c, d, e = [tuple(Fraction(float(v)) for v in p[3 * i:3 * i + 3]) for i in range(5)]
    rows = []
    for x in (a, b, c, d):
        dx, dy, dz = x[0] - e[0], x[1] - e[1], x[2] - e[2]
        rows.append([dx, dy, dz, dx * dx + dy * dy + dz * dz])

    def det(mat):
        if len(mat) == 1:
            return mat[0][0]
        return sum((-1) ** j * mat[0][j] * det([r[:j] + r[j + 1:] for r in mat[1:]])
                   for j in range(len(mat)))
    v = det(rows)
    return (v > 0) - (v < 0)


def test_certified_signs_are_exact():
    rng = np.random.default_rng(3)
    cases = []
    # near-cospherical: four points on a sphere, the fifth on it too, perturbed
    for _ in range(600):
        ctr = rng.random(3)
        r = 0.1 + rng.random()
        u = rng.standard_normal((5, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        p = ctr + r * u
        p[4] += rng.choice([0.0, 1e-15, 1e-13, 1e-11, 1e-9]) * rng.standard_normal(3)
        cases.append(p.ravel())
    pts = np.array(cases)
    det, unc = classify(pts)
    checked = 0
    for i in np.flatnonzero(~unc):
        assert np.sign(det[i]) == exact_sign(pts[i]), f"case {i}: certified sign is wrong"
        checked += 1
    assert checked > 100 and unc.sum() > 10


def test_exactly_cospherical_is_uncertain():
    # corners of a unit cube: any four plus a fifth corner are cospherical
    cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    pts = np.array([np.concatenate([cube[0], cube[1], cube[2], cube[4], cube[7]]),
                    np.concatenate([cube[3], cube[5], cube[6], cube[0], cube[1]])])
    det, unc = classify(pts)
    assert (det == 0).all() and unc.all()


def test_separated_points_are_certified():
    rng = np.random.default_rng(4)
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    inside = rng.uniform(-0.2, 0.2, (200, 3))
    far = rng.uniform(3, 5, (200, 3)) * rng.choice([-1, 1], (200, 3))
    e = np.vstack([inside, far])
    pts = np.hstack([np.tile(tet.ravel(), (400, 1)), e])
    det, unc = classify(pts)
    assert not unc.any()
    assert (det[:200] != 0).all() and (np.sign(det[:200]) == np.sign(det[0])).all()


# ---- the exact decisions of the robust predicate mode (nat_predicate_exact)

def exact_host(pts, kind):
    w = 12 if kind == 0 else 15
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, w)
    n = pts.shape[0]
    sign = np.zeros(n, np.int32)
    unc = np.zeros(n, np.int32)
    ip = ctypes.POINTER(ctypes.c_int32)
    rc = _lib.lib().nat_predicate_exact(pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                        kind, sign.ctypes.data_as(ip), unc.ctypes.data_as(ip))
    assert rc == 0
    return sign, unc.astype(bool)


def near_degenerate(rng, n, kind):
    """Points on a sphere (insphere) or a plane (orient3d), the last one
    nudged by 0 to 1e-9, at scales 1e-3 .. 1e3 and offsets up to 1e3."""
    out = []
    for _ in range(n):
        scale = 10.0 ** rng.integers(-3, 4)
        off = rng.choice([0.0, 1.0, 1e3]) * rng.standard_normal(3)
        if kind == 1:
            u = rng.standard_normal((5, 3))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            p = off + scale * u
        else:
            b = rng.standard_normal((2, 3))
            w = rng.standard_normal((4, 2))
            p = off + scale * (w @ b)
        p[-1] += rng.choice([0.0, 1e-17, 1e-15, 1e-12, 1e-9]) * scale * rng.standard_normal(3)
        out.append(p.ravel())
    return np.array(out)


def int_sign(p, kind):
    from oracle import exact_predicates as X
    pts = [tuple(p[3 * i:3 * i + 3]) for i in range(4 if kind == 0 else 5)]
    return X.orient3d_sign(*pts) if kind == 0 else X.insphere_sign(*pts)


@pytest.mark.parametrize("kind", [0, 1])
def test_exact_signs_match_integer_determinants(kind):
    """The library's exact sign (1024-bit integers, the FP64 expression's
    order) against an independent cofactor expansion in Python integers, and
    against Fractions (exact_sign above) for insphere; the static filter
    never flags a test whose FP64 sign disagrees with it."""
    rng = np.random.default_rng(11 + kind)
    pts = near_degenerate(rng, 1500, kind)
    sign, unc = exact_host(pts, kind)
    assert not (sign == 2).any()
    want = np.array([int_sign(p, kind) for p in pts])
    assert np.array_equal(sign, want)
    if kind == 1:
        for i in range(0, len(pts), 15):
            assert sign[i] == exact_sign(pts[i])
        det, unc2 = classify(pts)
        assert np.array_equal(unc, unc2)
        assert (np.sign(det[~unc]) == sign[~unc]).all()
    assert unc.sum() > 50 and (~unc).sum() > 50


def test_exact_signs_on_wide_exponent_ranges():
    """Coordinates from 1e-30 to 1e3 in one test (spans the evaluator covers),
    subnormals, zeros; a span beyond its 1024 bits is reported as 2."""
    rng = np.random.default_rng(5)
    cases = []
    for _ in range(300):
        p = rng.standard_normal((5, 3)) * 10.0 ** rng.integers(-30, 4, (5, 3))
        p[rng.random((5, 3)) < 0.1] = 0.0
        cases.append(p.ravel())
    pts = np.array(cases)
    for kind in (0, 1):
        sub = pts[:, :12] if kind == 0 else pts
        sign, _ = exact_host(sub, kind)
        want = np.array([int_sign(p, kind) for p in sub])
        ok = sign != 2
        assert ok.sum() > 250
        assert np.array_equal(sign[ok], want[ok])
    # subnormal coordinates alone (multiples of 2^-1074) are on one grid
    sub = np.array([[0, 0, 0, 8, 0, 0, 0, 8, 0, 0, 0, 8, 2, 2, 2],
                    [0, 0, 0, 8, 0, 0, 0, 8, 0, 0, 0, 8, 9, 9, 9]], float) * 5e-324
    s, _ = exact_host(sub, 1)
    assert list(s) == [int_sign(p, 1) for p in sub] and 2 not in s
    # a subnormal beside 1.0 spans 1074 bits: unresolved
    mixed = np.array([[5e-324, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.25, 0.25, 0.25]])
    assert exact_host(mixed, 1)[0][0] == 2
    wide = np.array([[1e-300, 0, 0, 1e300, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.5, 0.5]])
    assert exact_host(wide, 1)[0][0] == 2
    assert exact_host(wide[:, :12], 0)[0][0] == 2


def test_exact_signs_on_closed_forms():
    """orient3d(e0, e1, e2, 0) = det(I) = +1 and -1 with e0, e1 swapped;
    any five cube corners are cospherical (insphere 0); the cube's centre is
    inside the circumsphere of a positively oriented corner tet, (2, 2, 2)
    outside it."""
    e = np.eye(3)
    z = np.zeros(3)
    pts = np.array([np.concatenate([e[0], e[1], e[2], z]), np.concatenate([e[1], e[0], e[2], z])])
    s, _ = exact_host(pts, 0)
    assert list(s) == [1, -1]
    cube = np.array([[x, y, z_] for x in (0, 1) for y in (0, 1) for z_ in (0, 1)], float)
    a, b, c, d = cube[0], cube[1], cube[2], cube[4]
    if int_sign(np.concatenate([a, b, c, d]), 0) < 0:
        a, b = b, a
    ins = np.array([np.concatenate([a, b, c, d, cube[7]]),
                    np.concatenate([a, b, c, d, [0.5, 0.5, 0.5]]),
                    np.concatenate([a, b, c, d, [2.0, 2.0, 2.0]])])
    s, _ = exact_host(ins, 1)
    assert list(s) == [0, 1, -1]


def test_integer_determinant_plant_is_caught():
    """A planted mistake in the checker's expansion (the lift of e dropped)
    disagrees with the library on near-cospherical tests."""
    from oracle import exact_predicates as X
    rng = np.random.default_rng(2)
    pts = near_degenerate(rng, 200, 1)
    sign, _ = exact_host(pts, 1)

    def planted(a, b, c, d, e):
        P = X._common_ints([a, b, c, d, e])
        rows = []
        for i in range(4):
            dx = [P[i][k] - P[4][k] for k in range(3)]
            rows.append(dx + [P[i][0] ** 2 + P[i][1] ** 2 + P[i][2] ** 2])
        return X._sign(X._det(rows))
    bad = [planted(*[tuple(p[3 * i:3 * i + 3]) for i in range(5)]) for p in pts]
    assert not np.array_equal(sign, np.array(bad))
