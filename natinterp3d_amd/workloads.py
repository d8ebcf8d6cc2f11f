"""Seeded synthetic key/query/value sets used by the tests and the benchmark.

Each generator draws from ``np.random.default_rng(seed)`` in a fixed order;
keys are permuted after generation (the reference's insertion-order k-d tree
and list-head walks degenerate on sorted input).
"""
import numpy as np


def uniform_small(seed=0, n_keys=10_000, n_queries=100_000):
    """10k keys uniform in [0,1]^3, 100k queries in [0.05,0.95]^3, scalar
    float64 values."""
    rng = np.random.default_rng(seed)
    keys = rng.random((n_keys, 3))
    queries = 0.05 + 0.9 * rng.random((n_queries, 3))
    values = rng.standard_normal(n_keys)
    return keys, queries, values


def plain_at_scale(seed=1, n_keys=1_000_000, n_queries=10_000_000):
    """1M keys (90% uniform in [0,1]^3, 10% in 20 Gaussian blobs of sigma
    0.01), 10M queries uniform in [0.02,0.98]^3, float32 values [N,3]."""
    rng = np.random.default_rng(seed)
    n_u = n_keys * 9 // 10
    n_b = n_keys - n_u
    a = rng.random((n_u, 3))
    centers = rng.random((20, 3))
    b = centers[rng.integers(0, 20, n_b)] + 0.01 * rng.standard_normal((n_b, 3))
    keys = np.vstack([a, b])
    keys = keys[rng.permutation(len(keys))]
    queries = 0.02 + 0.96 * rng.random((n_queries, 3))
    values = rng.standard_normal((n_keys, 3)).astype(np.float32)
    return keys, queries, values


def _real_sph_harm(l, m, theta, phi):
    """Real spherical harmonic of degree l, order m at polar angle theta and
    azimuth phi."""
    from scipy.special import sph_harm_y
    y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return np.sqrt(2.0) * (-1.0) ** m * y.real
    if m < 0:
        return np.sqrt(2.0) * (-1.0) ** m * y.imag
    return y.real


def shell(seed=2, n_keys=20_000, n_queries=2_000_000, n_harmonics=8, max_degree=4):
    """Keys on a noisy closed shell: unit directions at radius
    r(u) = 1 + 0.1 * sum_k a_k Y_{l_k m_k}(u) (real harmonics, l_k <= 4,
    a_k ~ N(0,1)) plus N(0, 0.01) radial noise; queries inside 0.9 r(u)
    (uniform in volume along each direction); float64 values [N,3].
    Returns keys, queries, values and the radius function."""
    rng = np.random.default_rng(seed)
    ls = rng.integers(0, max_degree + 1, n_harmonics)
    ms = np.array([rng.integers(-l, l + 1) for l in ls])
    amp = rng.standard_normal(n_harmonics)

    def radius(u):
        theta = np.arccos(np.clip(u[:, 2], -1.0, 1.0))
        phi = np.arctan2(u[:, 1], u[:, 0])
        r = np.ones(len(u))
        for l, m, a in zip(ls, ms, amp):
            r += 0.1 * a * _real_sph_harm(int(l), int(m), theta, phi)
        return r

    u = rng.standard_normal((n_keys, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    keys = u * (radius(u) + 0.01 * rng.standard_normal(n_keys))[:, None]
    keys = keys[rng.permutation(n_keys)]
    uq = rng.standard_normal((n_queries, 3))
    uq /= np.linalg.norm(uq, axis=1, keepdims=True)
    rq = 0.9 * radius(uq) * rng.random(n_queries) ** (1.0 / 3.0)
    queries = uq * rq[:, None]
    values = rng.standard_normal((n_keys, 3))
    return keys, queries, values, radius


def batch_problems(seed=3, n_problems=256, keys_range=(500, 50_000),
                   queries_range=(1_000, 200_000), dim=8, shell_every=0,
                   shell_keys=(300, 3_000), shell_queries=(1_000, 10_000)):
    """Independent problems of uneven size: per problem, key and query counts
    log-uniform in the given ranges, points uniform in an axis-aligned box at
    the origin with sides U[0.5, 2], float32 values [n_keys, dim].  With
    shell_every = k > 0, every k-th problem is instead a hollow shell (see
    shell(); key and query counts log-uniform in shell_keys, shell_queries)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_problems):
        if shell_every > 0 and i % shell_every == shell_every - 1:
            nk = int(np.floor(np.exp(rng.uniform(*np.log(shell_keys)))))
            nq = int(np.floor(np.exp(rng.uniform(*np.log(shell_queries)))))
            keys, queries, _, _ = shell(seed=int(rng.integers(2**31)), n_keys=nk,
                                        n_queries=nq)
            values = rng.standard_normal((nk, dim)).astype(np.float32)
            out.append((keys, queries, values))
            continue
        nk = int(np.floor(np.exp(rng.uniform(np.log(keys_range[0]), np.log(keys_range[1])))))
        nq = int(np.floor(np.exp(rng.uniform(np.log(queries_range[0]),
                                             np.log(queries_range[1])))))
        side = rng.uniform(0.5, 2.0, 3)
        keys = rng.random((nk, 3)) * side
        queries = rng.random((nq, 3)) * side
        values = rng.standard_normal((nk, dim)).astype(np.float32)
        out.append((keys, queries, values))
    return out


def lattice_jitter(seed=4, n_side=100, n_queries=10_000_000, jitter=1e-9):
    """n_side^3 lattice scaled to [0,1]^3 with per-coordinate jitter
    U[-jitter, jitter] (near-cospherical tets), permuted; queries uniform in
    [0.01,0.99]^3; scalar float64 values."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_side) / (n_side - 1.0)
    keys = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    keys = keys + rng.uniform(-jitter, jitter, keys.shape)
    keys = keys[rng.permutation(len(keys))]
    queries = 0.01 + 0.98 * rng.random((n_queries, 3))
    values = rng.standard_normal(len(keys))
    return keys, queries, values
