"""The static error bound behind the robustness audit (nat_insphere_classify,
host side): whenever it certifies a sign, the FP64 determinant has the sign
of the exact determinant (rational arithmetic); exactly cospherical inputs
are never certified; well-separated inputs always are."""
import ctypes
from fractions import Fraction

import numpy as np

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
