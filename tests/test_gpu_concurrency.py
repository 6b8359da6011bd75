"""Two handles on two streams at once: every call's temporary storage (workspace, arrival
counters, descriptor slots, flag bitmaps) belongs to its handle, so calls of different
handles may overlap on the GPU. Each of several plans, and a grouped call of 12 members,
runs on both handles, interleaved with no synchronisation between the streams, and every
output must equal, byte for byte, the same call made alone on one handle (which itself
passes the oracle elsewhere)."""
import sys

import pytest

from helpers import inputs, ptr, to_dev

pytestmark = pytest.mark.gpu

ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)
# (precision, ta, tb, m, n, k): packed FP32 with the correction, the one-launch kernel with k
# slices, a K split, bf16x3 separate products, FP16 block-scaled
CASES = [("fp32", "N", "C", 640, 384, 2048), ("fp32", "T", "N", 64, 48, 8192), ("fp32", "N", "N", 256, 256, 16384),
         ("bf16x3", "C", "T", 768, 512, 1024), ("fp16", "N", "T", 512, 640, 1536)]


def _problem(c, seed):
    prec, ta, tb, m, n, k = c
    lda, ldb, ldc = (m if ta == "N" else k), (k if tb == "N" else n), m
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, seed)
    return dict(prec=prec, ta=ta, tb=tb, m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, dA=to_dev(A), dB=to_dev(B),
                C=C)


def _launch(h, p, dC):
    import cgemm3m

    st = cgemm3m.cgemm3m(h, p["ta"], p["tb"], p["m"], p["n"], p["k"], ALPHA, ptr(p["dA"]), p["lda"], ptr(p["dB"]),
                         p["ldb"], BETA, ptr(dC), p["ldc"], cgemm3m.config(p["prec"]), check=False)
    assert st == 0, (p["prec"], p["m"], p["n"], p["k"], st)


def _grouped_members(seed):
    import numpy as np

    rng = np.random.default_rng(seed)
    members = []
    for i in range(12):
        ta, tb = str(rng.choice(list("NTC"))), str(rng.choice(list("NTC")))
        m, n, k = int(rng.integers(16, 300)), int(rng.integers(16, 300)), int(rng.integers(8, 1500))
        members.append(_problem(("fp32", ta, tb, m, n, k), seed + 11 * i))
    return members


def _grouped(h, members, outs):
    import cgemm3m

    probs = [dict(ta=p["ta"], tb=p["tb"], m=p["m"], n=p["n"], k=p["k"], alpha=ALPHA, A=ptr(p["dA"]), lda=p["lda"],
                  B=ptr(p["dB"]), ldb=p["ldb"], beta=BETA, C=ptr(dC), ldc=p["ldc"]) for p, dC in zip(members, outs)]
    cgemm3m.grouped(h, probs)


def test_two_handles_two_streams_interleaved(gpu):
    import torch

    import cgemm3m

    probs = [_problem(c, 20261300 + 7 * i) for i, c in enumerate(CASES)]
    members = _grouped_members(20261400)
    # alone: one handle, one stream, one call at a time
    solo = []
    for p in probs:
        dC = to_dev(p["C"])
        _launch(gpu, p, dC)
        torch.cuda.synchronize()
        solo.append(dC.cpu().numpy())
    gC = [to_dev(p["C"]) for p in members]
    _grouped(gpu, members, gC)
    torch.cuda.synchronize()
    solo_grouped = [c.cpu().numpy() for c in gC]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    h1, h2 = cgemm3m.Handle(stream=s1), cgemm3m.Handle(stream=s2)
    try:
        for rnd in range(3):
            print("round", rnd)
            sys.stdout.flush()
            outs = {}
            # inputs were written on the default stream before; both streams start after it
            torch.cuda.synchronize()
            grouped_outs = {}
            for i, p in enumerate(probs):
                for name, h, s in (("h1", h1, s1), ("h2", h2, s2)):
                    with torch.cuda.stream(s):
                        dC = to_dev(p["C"])  # allocated and filled on this stream
                    outs[(i, name)] = dC
                    _launch(h, p, dC)
                    if i == 1:  # a grouped call between the single calls
                        with torch.cuda.stream(s):
                            g_out = [to_dev(q["C"]) for q in members]
                        grouped_outs[name] = g_out
                        _grouped(h, members, g_out)
            torch.cuda.synchronize()
            for name, g_out in grouped_outs.items():
                for j, (c, ref) in enumerate(zip(g_out, solo_grouped)):
                    same = c.cpu().numpy().tobytes() == ref.tobytes()
                    assert same, "round %d grouped member %d on %s differs from the call made alone" % (rnd, j, name)
            for (i, name), dC in outs.items():
                same = dC.cpu().numpy().tobytes() == solo[i].tobytes()
                assert same, "round %d case %s on %s differs from the call made alone" % (rnd, CASES[i], name)
    finally:
        h1.close()
        h2.close()
