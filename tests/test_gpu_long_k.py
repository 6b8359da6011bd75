"""Very long k: m = n = 64, k = 2^22 + 7 (each operand 2.1 GB), where every k-direction
counter, chunk index and energy chain is at its longest: the one-launch kernel with k slices
(its default plan), the packed path with 16 K partitions (pinned), and bf16x3 (separate
products split in k; a leading dimension of 2^22 is too wide for reading the operands as
stored). FP32: error against FP64 (the GPU residual) at most twice the Netlib-order oracle's
own, and the reltol check against the oracle at the larger of 1e-5 and twice that error (at
this k the oracle's sequential FP32 sums are far from exact, so the verifier's per-output
allowance is not the bar); bf16x3: its normwise bound. Reruns are byte-identical."""
import sys

import numpy as np
import pytest

import oracle
from helpers import inputs, ptr, run_gpu, to_dev

pytestmark = pytest.mark.gpu

M = N = 64
K = (1 << 22) + 7
ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)


@pytest.fixture(scope="module")
def longk(gpu):
    import cgemm3m

    A, B, C = inputs("T", "N", M, N, K, K, K, M, 20261700)
    dA, dB = to_dev(A), to_dev(B)
    gold = oracle.cgemm("T", "N", M, N, K, ALPHA, A, K, B, K, BETA, C, M)
    dC0, dGold = to_dev(C), to_dev(gold)
    ref = cgemm3m.residual(gpu, "T", "N", M, N, K, ALPHA, ptr(dA), K, ptr(dB), K, BETA, ptr(dC0), ptr(dGold), M)
    print("oracle's own FP64 error", ref)
    sys.stdout.flush()
    return dict(A=A, B=B, C=C, dA=dA, dB=dB, dC0=dC0, gold=gold, ref=ref)


@pytest.mark.parametrize("form,pin,path", [("fp32", {}, 5), ("fp32", dict(tile=64, k_split=16), 2),
                                           ("bf16x3", {}, 4)], ids=["one-launch", "packed-16-parts", "bf16x3"])
def test_very_long_k(gpu, longk, form, pin, path):
    import cgemm3m

    d = longk
    cfg = cgemm3m.config(form, **pin)
    print("case long k", form, pin, "k =", K)
    sys.stdout.flush()
    st, out = run_gpu(gpu, "T", "N", M, N, K, ALPHA, d["A"], K, d["B"], K, BETA, d["C"], M, cfg, d["dA"], d["dB"])
    assert st == 0, st
    plan = gpu.last_plan(sync=True)
    print("plan", plan)
    assert plan["path"] == path, plan
    st2, out2 = run_gpu(gpu, "T", "N", M, N, K, ALPHA, d["A"], K, d["B"], K, BETA, d["C"], M, cfg, d["dA"], d["dB"])
    assert st2 == 0 and out2.tobytes() == out.tobytes(), "second run differs"
    assert np.isfinite(out).all()
    res = cgemm3m.residual(gpu, "T", "N", M, N, K, ALPHA, ptr(d["dA"]), K, ptr(d["dB"]), K, BETA, ptr(d["dC0"]),
                           ptr(to_dev(out)), M)
    ref = d["ref"]
    print("FP64 error: ours %s oracle %s" % (res, ref))
    if form == "fp32":
        assert res["normwise"] <= 2 * ref["normwise"], (res, ref)
        tol = max(1e-5, 2 * ref["normwise"])
        ok, worst = oracle.reltol_check(out, d["gold"], rtol=tol, atol_frac=tol)
        print("reltol at %.3g: worst %.3g" % (tol, worst))
        assert ok, worst
    else:
        assert res["normwise"] <= 6e-5, res
