"""Writes tests/golden/setup_hashes.json with the CPU ORACLE's setup of the 30^3
configurations (generator output, colors, permutation, color offsets, CSR pattern) as
short sha256 prefixes, and small oracle value/prec vectors for a 4^3 cube. Run from the
repository root after `make -C oracle` and `make -C mini-fem_amd`:

    python tests/golden/make_golden.py

The generator is the product's (the oracle has none); everything downstream of it is
computed by the oracle, so the product's setup is checked against these hashes.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "mini-fem_amd"), os.path.join(ROOT, "oracle")]

import minifem  # noqa: E402
import pyoracle  # noqa: E402

CONFIGS = {
    "lap_30_cube": dict(n=30, jitter=0.2, seed=1, shuffle=0),
    "ela_30_cube": dict(n=30, jitter=0.2, seed=2, shuffle=0),
    "ela_12_shuffled": dict(n=12, jitter=0.25, seed=3, shuffle=203),
}


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def main():
    out = {}
    for name, g in CONFIGS.items():
        coord, e2n, code, nnz = minifem.generate_cube(g["n"], g["jitter"], g["seed"], g["shuffle"])
        nN = coord.size // 3
        part, perm, c2e = pyoracle.coloring(e2n, nN)
        pe = pyoracle.permute(e2n, perm, 4)
        row, col = pyoracle.csr(pe, nN, nnz)
        out[name] = dict(g, coord=h(coord), elemToNode=h(e2n), code=h(code), colors=h(part),
                         perm=h(perm), colorToElem=c2e.tolist(), row=h(row), col=h(col),
                         nnz=int(row[-1]))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "setup_hashes.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    # small value vectors (4^3 cube, coefficients), lap and ela, COLORING order
    vec = {}
    for op in (0, 1):
        coord, e2n, code, nnz = minifem.generate_cube(4, 0.2, 7, 0)
        coef = minifem.generate_coef(op, e2n.size // 4, 107)
        m = pyoracle.OracleMesh(coord, e2n, code, nnz, op, coef=coef)
        vals = m.assemble()
        p, _ = m.prec(vals)
        vec[f"op{op}"] = dict(values=vals.tolist(), prec=p.tolist(), row=m.row.tolist(),
                              col=m.col.tolist())
    with open(os.path.join(os.path.dirname(path), "cube4_values.json"), "w") as f:
        json.dump(vec, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
