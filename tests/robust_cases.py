"""Queries on the circumspheres of mesh tets, for the robust predicate mode.

For a tet T of the mesh, q = c + r u with c and r T's circumcentre and
radius (solved in extended precision, Cramer's rule) and u the direction from
c through the centroid of one of T's faces; q rounded to float64 lies within
a few ulps of T's circumsphere, beyond that face, so the cavity search tests
T from the tet across the face and the FP64 insphere sign is not certified.
"""
import numpy as np


def _det3(M):
    return (M[0, 0] * (M[1, 1] * M[2, 2] - M[1, 2] * M[2, 1])
            - M[0, 1] * (M[1, 0] * M[2, 2] - M[1, 2] * M[2, 0])
            + M[0, 2] * (M[1, 0] * M[2, 1] - M[1, 1] * M[2, 0]))


def borderline_queries(packed, n, seed=0, box=0.3):
    """n queries on circumspheres of tets of data vertices whose centroid is
    inside [-box, box]^3 (centred coordinates)."""
    V = np.asarray(packed["verts"], np.float64)
    tv = np.asarray(packed["tet_v"], np.int64)
    nd = V.shape[0] - 4
    P = V[tv]
    cen = P.mean(1)
    inner = np.flatnonzero((tv < nd).all(1) & (np.abs(cen) < box).all(1))
    rng = np.random.default_rng(seed)
    out = []
    for t in rng.choice(inner, n, replace=len(inner) < n):
        p = P[t].astype(np.longdouble)
        A = 2 * (p[1:] - p[0])
        rhs = (p[1:] ** 2).sum(1) - (p[0] ** 2).sum()
        D = _det3(A)
        cc = np.zeros(3, np.longdouble)
        for j in range(3):
            Mj = A.copy()
            Mj[:, j] = rhs
            cc[j] = _det3(Mj) / D
        r = np.sqrt(((p[0] - cc) ** 2).sum())
        fc = np.delete(p, rng.integers(4), 0).mean(0)
        u = fc - cc
        u /= np.sqrt((u ** 2).sum())
        out.append((cc + r * u).astype(np.float64))
    return np.array(out)
