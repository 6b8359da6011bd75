"""Writes the golden fixtures in this folder: small seeded cases with their Netlib-order
FP32 result (the oracle, oracle/oracle.c) and FP64 result, so the oracle is pinned against
regressions and the GPU path is checked against stored vectors that need no CPU oracle run.

Cases: the 9 op pairs at 16 x 12 x 20 with padded leading dimensions, alpha = (1, 0.5),
beta = (0.5, -0.25); ragged 7 x 5 x 9 (N, N) and (C, T); beta == 0 with a NaN-filled C (C
must not be read, its padding rows must stay NaN); k == 1 with beta == 1.

Usage: python tests/golden/make_golden.py   (writes golden.npz and golden.json here)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def stored(t, rows, cols):
    return (rows, cols) if t == "N" else (cols, rows)


def case_list():
    out = []
    seed = 20260900
    for ta in "NTC":
        for tb in "NTC":
            out.append(dict(ta=ta, tb=tb, m=16, n=12, k=20, pad=(1, 2, 3), alpha=[1.0, 0.5], beta=[0.5, -0.25],
                            seed=seed, nan_c=False))
            seed += 3
    out.append(dict(ta="N", tb="N", m=7, n=5, k=9, pad=(0, 0, 0), alpha=[0.7, -0.3], beta=[0.2, 0.9], seed=seed,
                    nan_c=False))
    out.append(dict(ta="C", tb="T", m=7, n=5, k=9, pad=(2, 1, 1), alpha=[0.7, -0.3], beta=[0.2, 0.9],
                    seed=seed + 3, nan_c=False))
    out.append(dict(ta="T", tb="N", m=9, n=6, k=11, pad=(0, 1, 2), alpha=[1.25, 0.0], beta=[0.0, 0.0],
                    seed=seed + 6, nan_c=True))
    out.append(dict(ta="N", tb="C", m=10, n=4, k=1, pad=(0, 0, 0), alpha=[-0.5, 2.0], beta=[1.0, 0.0],
                    seed=seed + 9, nan_c=False))
    return out


def main():
    arrays, meta = {}, []
    for idx, c in enumerate(case_list()):
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        ar, ac = stored(ta, m, k)
        br, bc = stored(tb, k, n)
        lda, ldb, ldc = ar + c["pad"][0], br + c["pad"][1], m + c["pad"][2]
        A = oracle.fill(ar, ac, lda, c["seed"])
        B = oracle.fill(br, bc, ldb, c["seed"] + 1)
        C = oracle.fill(m, n, ldc, c["seed"] + 2)
        if c["nan_c"]:
            C[:] = np.complex64(complex(np.nan, np.nan))
        alpha, beta = complex(*c["alpha"]), complex(*c["beta"])
        out32 = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        out64 = oracle.zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        for name, arr in (("A", A), ("B", B), ("C", C), ("out32", out32), ("out64", out64)):
            arrays["%d_%s" % (idx, name)] = arr
        meta.append(dict(c, lda=lda, ldb=ldb, ldc=ldc))
    np.savez(os.path.join(HERE, "golden.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote %d cases" % len(meta))


if __name__ == "__main__":
    main()
