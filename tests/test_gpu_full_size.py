"""The 8192^3 precision-forms workload at full size (complex64, N,N and C,T; uniform
[-1, 1] entries, alpha = (1, 0.5), beta = (0.5, -0.25)), checked through properties that do
not depend on size: byte-identical reruns of every form, the FP64 error of each form
against the plain FP32 path's (the bf16x3 target: max elementwise and normwise error at most
2x the plain path's), and, with integer-valued inputs that make every path exact, sampled
output columns equal to the Netlib-order oracle bit for bit."""
import os
import sys

import numpy as np
import pytest

import oracle
from helpers import ptr

pytestmark = pytest.mark.gpu

N = 8192
ALPHA, BETA = complex(1.0, 0.5), complex(0.5, -0.25)
FORMS = ["fp32", "bf16x3", "bf16x6", "fp16"]


def _inputs(torch, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = lambda: (torch.rand(2 * N * N, generator=g, device="cuda") * 2 - 1).view(torch.complex64)
    return u(), u(), u()


@pytest.mark.parametrize("ops", ["NN", "CT"])
def test_8192_forms_errors_and_reruns(gpu, ops):
    import torch

    import cgemm3m

    ta, tb = ops[0], ops[1]
    A, B, C0 = _inputs(torch, 20261800 + len(ops) + "NTC".index(ta))
    errs = {}
    for form in FORMS:
        cfg = cgemm3m.config(form)
        outs = []
        for _ in range(2):
            C = C0.clone()
            st = cgemm3m.cgemm3m(gpu, ta, tb, N, N, N, ALPHA, ptr(A), N, ptr(B), N, BETA, ptr(C), N, cfg)
            assert st == 0
            torch.cuda.synchronize()
            outs.append(C)
        same = torch.equal(outs[0].view(torch.int64), outs[1].view(torch.int64))
        assert same, "%s %s: reruns differ" % (form, ops)
        errs[form] = cgemm3m.residual(gpu, ta, tb, N, N, N, ALPHA, ptr(A), N, ptr(B), N, BETA, ptr(C0),
                                      ptr(outs[0]), N)
        print(form, ops, errs[form])
        sys.stdout.flush()
        del outs
    fp, bf = errs["fp32"], errs["bf16x3"]
    assert fp["normwise"] < 2e-5, fp
    # the target: bf16x3 at no worse than twice the plain path's error
    assert bf["max_rel"] <= 2.0 * fp["max_rel"], (bf, fp)
    assert bf["normwise"] <= 2.0 * fp["normwise"], (bf, fp)
    assert errs["bf16x6"]["normwise"] <= 2.0 * fp["normwise"], errs["bf16x6"]
    assert errs["fp16"]["normwise"] < 2e-3, errs["fp16"]


def test_8192_row_scaled_bf16x3_within_twice_fp32(gpu):
    """The range-stress variant: rows of op(A) scaled by 10^u, u ~ U[-3, 3] (N,N). bf16x3's
    max elementwise and normwise FP64 errors stay within twice the plain path's."""
    import torch

    import cgemm3m

    A, B, C0 = _inputs(torch, 20261850)
    g = torch.Generator(device="cuda")
    g.manual_seed(20261851)
    scale = torch.pow(10.0, (torch.rand(N, generator=g, device="cuda") * 6 - 3)).to(torch.complex64)
    A = (A.view(N, N) * scale[None, :]).reshape(-1)  # column-major: row i of A is A.view(N, N)[:, i]
    errs = {}
    for form in ("fp32", "bf16x3"):
        C = C0.clone()
        st = cgemm3m.cgemm3m(gpu, "N", "N", N, N, N, ALPHA, ptr(A), N, ptr(B), N, BETA, ptr(C), N,
                             cgemm3m.config(form))
        assert st == 0
        errs[form] = cgemm3m.residual(gpu, "N", "N", N, N, N, ALPHA, ptr(A), N, ptr(B), N, BETA, ptr(C0), ptr(C), N)
        print(form, "row-scaled", errs[form])
    fp, bf = errs["fp32"], errs["bf16x3"]
    assert bf["max_rel"] <= 2.0 * fp["max_rel"], (bf, fp)
    assert bf["normwise"] <= 2.0 * fp["normwise"], (bf, fp)


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_8192_integer_inputs_bitwise_sampled(gpu, form):
    """Integer real and imaginary parts in [-8, 8]: |partial sums| <= 16 * 16 * 8192 < 2^24,
    so every path is exact; 24 sampled columns equal the oracle's bit for bit."""
    import torch

    import cgemm3m

    g = torch.Generator(device="cuda")
    g.manual_seed(20261900)
    ints = lambda: torch.randint(-8, 9, (2 * N * N,), generator=g, device="cuda").float().view(torch.complex64)
    A, B, C0 = ints(), ints(), ints()
    alpha, beta = complex(2, -1), complex(1, 1)
    C = C0.clone()
    st = cgemm3m.cgemm3m(gpu, "C", "N", N, N, N, alpha, ptr(A), N, ptr(B), N, beta, ptr(C), N, cgemm3m.config(form))
    assert st == 0
    torch.cuda.synchronize()
    cols = np.random.default_rng(5).choice(N, 24, replace=False)
    Ah = A.cpu().numpy()
    out = C.view(N, N)[torch.as_tensor(cols, device="cuda")].cpu().numpy()
    Bs = B.view(N, N)[torch.as_tensor(cols, device="cuda")].cpu().numpy().reshape(-1)
    Cs = C0.view(N, N)[torch.as_tensor(cols, device="cuda")].cpu().numpy().reshape(-1)
    threads = min(16, len(os.sched_getaffinity(0)))
    gold = oracle.cgemm("C", "N", N, len(cols), N, alpha, Ah, N, Bs, N, beta, Cs, N,
                        threads=threads).reshape(len(cols), N)
    got = out.view(np.float32) + np.float32(0.0)
    want = gold.view(np.float32) + np.float32(0.0)
    same = got.tobytes() == want.tobytes()
    assert same, "%s: %d sampled outputs differ" % (form, int(np.count_nonzero(out != gold)))


def test_grouped_512_batch(gpu):
    """The grouped configuration at full size: 512 problems, m, n uniform over multiples of 16
    in [32, 512], k over multiples of 64 in [64, 2048], ops uniform over {N,T,C}^2, alpha = 1,
    beta = 0. Every member's FP64 normwise error is within the FP32 bound, a rerun is
    byte-identical, and replays of a captured graph equal the eager call byte for byte."""
    import torch

    import cgemm3m

    rng = np.random.default_rng(4243)
    g = torch.Generator(device="cuda")
    g.manual_seed(4243)
    probs, keep = [], []
    for _ in range(512):
        ta, tb = str(rng.choice(list("NTC"))), str(rng.choice(list("NTC")))
        m, n, k = int(rng.integers(2, 33)) * 16, int(rng.integers(2, 33)) * 16, int(rng.integers(1, 33)) * 64
        A = (torch.rand(2 * m * k, generator=g, device="cuda") * 2 - 1).view(torch.complex64)
        B = (torch.rand(2 * k * n, generator=g, device="cuda") * 2 - 1).view(torch.complex64)
        C = torch.zeros(m * n, dtype=torch.complex64, device="cuda")
        keep.append((A, B, C))
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=1.0, A=ptr(A), lda=m if ta == "N" else k, B=ptr(B),
                          ldb=k if tb == "N" else n, beta=0.0, C=ptr(C), ldc=m))
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    first = [c.clone() for (_, _, c) in keep]
    zero = torch.zeros(512 * 512, dtype=torch.complex64, device="cuda")
    worst = 0.0
    for p, (A, B, C) in zip(probs, keep):
        r = cgemm3m.residual(gpu, p["ta"], p["tb"], p["m"], p["n"], p["k"], 1.0, p["A"], p["lda"], p["B"], p["ldb"],
                             0.0, ptr(zero), ptr(C), p["ldc"])
        worst = max(worst, r["normwise"])
        assert r["normwise"] < 4e-6, (p["ta"], p["tb"], p["m"], p["n"], p["k"], r)
    print("grouped 512: worst member normwise error %.3g" % worst)
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    again = sum(not torch.equal(c.view(torch.int64), f.view(torch.int64)) for (_, _, c), f in zip(keep, first))
    assert again == 0, "%d members differ on a rerun" % again
    stream = torch.cuda.current_stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gpu.set_stream(torch.cuda.current_stream())
        cgemm3m.grouped(gpu, probs)
    gpu.set_stream(stream)
    for rep in range(2):
        for (_, _, c) in keep:
            c.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        bad = sum(not torch.equal(c.view(torch.int64), f.view(torch.int64)) for (_, _, c), f in zip(keep, first))
        assert bad == 0, "replay %d: %d members differ" % (rep, bad)
    del graph

