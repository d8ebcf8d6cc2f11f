"""Exact predicates and the exact-decision cavity search, in Python integers.

TEST INFRASTRUCTURE ONLY: the checker of the product's opt-in robust
predicate mode.  Imported by tests/ only; the product never imports it.

The reference decides every query-path test with the plain FP64
determinants (orient3dfast, predicates.c:1687-1710; inspherefast,
predicates.c:3301-3347) and never calls its adaptive ones
(predicates.c:3940-4265).  The robust mode replaces the sign of each test
the FP64 error bound does not certify by the sign of the exact determinant;
this module gives those exact signs independently -- every double is a
dyadic rational, so the coordinates of one test are put over a common
power-of-two denominator and the determinant is expanded in Python's
unbounded integers (cofactor expansion, not the FP64 expression's order) --
and restates the reference's cavity search (findCavity,
natural_insertionfree.c:486-551: depth-first from the start tet, a tet
joins when the query is strictly inside its circumsphere, its unvisited
neighbours are pushed) with those exact decisions.
"""
from fractions import Fraction

import numpy as np


def _common_ints(points):
    """Integers P[i][k] and a power-of-two D with points[i][k] = P[i][k] / D."""
    ratios = [[float(v).as_integer_ratio() for v in p] for p in points]
    den = 1
    for row in ratios:
        for _, d in row:
            den = max(den, d)
    return [[n * (den // d) for n, d in row] for row in ratios]


def _det(mat):
    if len(mat) == 1:
        return mat[0][0]
    total = 0
    for j, v in enumerate(mat[0]):
        if v:
            minor = [r[:j] + r[j + 1:] for r in mat[1:]]
            total += (-1) ** j * v * _det(minor)
    return total


def _sign(v):
    return (v > 0) - (v < 0)


def orient3d_sign(a, b, c, d):
    """Sign of the exact orient3d(a, b, c, d): det[a - d; b - d; c - d]."""
    P = _common_ints([a, b, c, d])
    rows = [[P[i][k] - P[3][k] for k in range(3)] for i in range(3)]
    return _sign(_det(rows))


def insphere_sign(a, b, c, d, e):
    """Sign of the exact insphere(a, b, c, d, e): det of the lifted rows
    [x - e, |x - e|^2] for x = a, b, c, d; > 0 when e is inside the
    circumsphere of the positively oriented (a, b, c, d)."""
    P = _common_ints([a, b, c, d, e])
    rows = []
    for i in range(4):
        dx = [P[i][k] - P[4][k] for k in range(3)]
        rows.append(dx + [dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]])
    return _sign(_det(rows))


def exact_value(points, kind):
    """The exact determinant as a Fraction (kind "orient3d" or "insphere")."""
    P = _common_ints(points)
    den = 1
    for v in points:
        for x in v:
            den = max(den, float(x).as_integer_ratio()[1])
    if kind == "orient3d":
        rows = [[P[i][k] - P[3][k] for k in range(3)] for i in range(3)]
        return Fraction(_det(rows), den ** 3)
    rows = []
    for i in range(4):
        dx = [P[i][k] - P[4][k] for k in range(3)]
        rows.append(dx + [dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]])
    return Fraction(_det(rows), den ** 5)


def exact_cavity(packed, q, start):
    """Cavity of query q from tet `start` with exact decisions: the tets in
    the order they join, each tet's orientation taken from the exact sign of
    its own orient3d (precomputeOrientSigns, natural_insertionfree.c:457-476,
    with the exact sign).  Returns (cavity tets, number of insphere tests)."""
    verts = packed["verts"]
    tet_v = packed["tet_v"]
    tet_n = packed["tet_n"]
    q = tuple(float(x) for x in q)
    V = lambda i: tuple(float(x) for x in verts[i])  # noqa: E731
    seen = set()
    stack = [int(start)]
    cavity = []
    tests = 0
    while stack:
        t = stack.pop()
        if t in seen:
            continue
        seen.add(t)
        p = [V(int(i)) for i in tet_v[t]]
        pos = orient3d_sign(*p) > 0
        s = insphere_sign(p[0], p[1], p[2], p[3], q) if pos else \
            insphere_sign(p[1], p[0], p[2], p[3], q)
        tests += 1
        if s > 0:
            cavity.append(t)
            for k in range(4):
                nb = int(tet_n[t][k])
                if nb >= 0 and nb not in seen:
                    stack.append(nb)
    return cavity, tests


def exact_neighbours(packed, q, start):
    """Natural neighbours of q under exact decisions: the data vertices of
    the exact cavity's tets (slots below n_data), as a set."""
    cav, _ = exact_cavity(packed, q, start)
    n = packed["verts"].shape[0] - 4
    out = set()
    for t in cav:
        for v in packed["tet_v"][t]:
            if int(v) < n:
                out.add(int(v))
    return out


def in_sphere_brute(packed, q, margin=1e-6):
    """Every tet whose circumsphere holds q strictly, by exact tests over the
    tets whose FP64 circumsphere comes within `margin` (relative to its
    radius) of q -- an independent check of the cavity search (on a
    Delaunay mesh the two sets are equal)."""
    verts = np.asarray(packed["verts"], np.float64)
    tv = np.asarray(packed["tet_v"], np.int64)
    P = verts[tv]                                        # [T, 4, 3]
    A = 2.0 * (P[:, 1:] - P[:, :1])                       # [T, 3, 3]
    rhs = (P[:, 1:] ** 2).sum(-1) - (P[:, :1] ** 2).sum(-1)
    ok = np.abs(np.linalg.det(A)) > 1e-300
    cc = np.zeros((len(tv), 3))
    cc[ok] = np.linalg.solve(A[ok], rhs[ok][..., None])[..., 0]
    r = np.linalg.norm(P[:, 0] - cc, axis=1)
    d = np.linalg.norm(np.asarray(q, np.float64)[None] - cc, axis=1)
    cand = np.flatnonzero(~ok | (d < r * (1.0 + margin)))
    out = []
    qt = tuple(float(x) for x in q)
    for t in cand:
        p = [tuple(float(x) for x in verts[i]) for i in tv[t]]
        o = orient3d_sign(*p)
        if o == 0:
            continue
        s = insphere_sign(p[0], p[1], p[2], p[3], qt) if o > 0 else \
            insphere_sign(p[1], p[0], p[2], p[3], qt)
        if s > 0:
            out.append(int(t))
    return sorted(out)
