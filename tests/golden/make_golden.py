"""Write the golden fixtures under tests/golden/ from the compiled reference
(oracle/_ref/libnatref.so, built from the reference's own C sources).

The reference's tests hold no stored vectors, so these fixtures are the
reference's outputs on the inputs its tests use (tests/test_natinterp3d.py)
plus the configuration-1 workload:

  small_cases.npz   keys, queries and the full reference CSR for each case
  meshes.npz        the reference mesh (packed arrays) for the small cases
  config1.json      SHA-256 digests of the reference CSR and mesh for
                    10,000 uniform keys / 100,000 queries (seed 0), plus
                    50 sampled rows in full

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from tests.golden.cases import CASES, config1  # noqa: E402


def centred(keys):
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    return keys - c, c


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    assert ref.available(), "build oracle/_ref first (make -C oracle ref)"
    small, meshes = {}, {}
    for name, (keys, queries) in CASES.items():
        kc, c = centred(keys)
        m = ref.RefMesh(kc)
        e = m.export()
        W = m.get_weights(np.ascontiguousarray(queries, np.float64) - c, num_threads=0)
        small[f"{name}__keys"] = np.asarray(keys, np.float64)
        small[f"{name}__queries"] = np.asarray(queries, np.float64)
        small[f"{name}__indptr"] = W.indptr
        small[f"{name}__indices"] = W.indices
        small[f"{name}__data"] = W.data
        for k, v in e.items():
            if k != "kd_data":
                meshes[f"{name}__{k}"] = v
        meshes[f"{name}__degeneracies"] = np.array(m.degeneracies(), np.int32)
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **small)
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **meshes)

    keys, queries = config1()
    kc, c = centred(keys)
    m = ref.RefMesh(kc)
    e = m.export()
    W = m.get_weights(queries - c, num_threads=8)
    rng = np.random.default_rng(123)
    rows = sorted(rng.choice(queries.shape[0], 50, replace=False).tolist())
    out = {
        "workload": "10,000 keys uniform in [0,1]^3 (default_rng(0)), 100,000 queries "
                    "0.05+0.9*U[0,1]^3 from the same generator",
        "n_tets": int(m.num_tets),
        "nnz": int(W.nnz),
        "sha256": {"indptr": sha(W.indptr), "indices": sha(W.indices), "data": sha(W.data)},
        "mesh_sha256": {k: sha(v) for k, v in e.items() if k != "kd_data"},
        "rows": {str(i): {"indices": W.indices[W.indptr[i]:W.indptr[i + 1]].tolist(),
                          "data": [float(x).hex() for x in
                                   W.data[W.indptr[i]:W.indptr[i + 1]]]} for i in rows},
    }
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote fixtures:", len(CASES), "small cases + config1 digests")


if __name__ == "__main__":
    main()
