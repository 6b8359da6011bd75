"""Single calls of every plan kind captured into a hipGraph (torch.cuda.CUDAGraph on a side
stream) and replayed: each replay, with eager calls on the same handle and synchronisations
between the replays, must give the eager call's bytes. Counters and flags that must start at
zero are cleared inside the call by fill kernels (no memset nodes), or left zero by the
kernels that use them; a fresh handle captures before any eager call has set up its
buffers."""
import sys

import pytest

from helpers import inputs, ptr, to_dev

pytestmark = pytest.mark.gpu

ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)
# (label, precision, ta, tb, m, n, k, pinned config fields, plan path or None)
PLANS = [("packed fp32", "fp32", "N", "T", 1024, 768, 1536, dict(tile=128, k_split=1), 1),
         ("k split", "fp32", "C", "N", 256, 256, 16384, dict(tile=128, k_split=3), 2),
         ("table k split", "fp32", "N", "N", 2048, 2048, 1024, {}, 2),
         ("one launch, k slices", "fp32", "T", "N", 64, 48, 8192, {}, 5),
         ("one launch", "fp32", "N", "C", 100, 80, 200, {}, 5),
         ("bf16x3 separate products", "bf16x3", "C", "T", 768, 512, 1024, {}, 4),
         ("bf16x3 read as stored", "bf16x3", "N", "N", 256, 256, 32768, {}, 4),
         ("fp16 block-scaled", "fp16", "N", "T", 512, 640, 1536, {}, None),
         ("bf16x6", "bf16x6", "T", "C", 300, 200, 700, {}, None)]


def _setup(p, seed):
    label, prec, ta, tb, m, n, k, pin, path = p
    lda, ldb = (m if ta == "N" else k), (k if tb == "N" else n)
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, m, seed)
    return dict(prec=prec, pin=pin, ta=ta, tb=tb, m=m, n=n, k=k, lda=lda, ldb=ldb, dA=to_dev(A), dB=to_dev(B),
                dC=to_dev(C), dC0=to_dev(C))


def _call(h, d):
    import cgemm3m

    st = cgemm3m.cgemm3m(h, d["ta"], d["tb"], d["m"], d["n"], d["k"], ALPHA, ptr(d["dA"]), d["lda"], ptr(d["dB"]),
                         d["ldb"], BETA, ptr(d["dC"]), d["m"], cgemm3m.config(d["prec"], **d["pin"]), check=False)
    assert st == 0, st


@pytest.mark.parametrize("fresh", [True, False], ids=["fresh-handle", "used-handle"])
@pytest.mark.parametrize("plan", PLANS, ids=[p[0].replace(" ", "-").replace(",", "") for p in PLANS])
def test_single_call_plan_hipgraph_replay(gpu, plan, fresh):
    import torch

    import cgemm3m

    print("case", plan, "fresh" if fresh else "used")
    sys.stdout.flush()
    d = _setup(plan, 20261500 + plan[4] + plan[6])
    other = _setup(plan, 20261600 + plan[4])  # eager work between the replays
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    hE = cgemm3m.Handle(stream=s)
    _call(hE, d)
    torch.cuda.synchronize()
    eager = d["dC"].cpu().numpy().copy()
    if plan[8] is not None:
        assert hE.last_plan(sync=True)["path"] == plan[8], hE.last_plan()
    hG = cgemm3m.Handle(stream=s) if fresh else hE
    try:
        d["dC"].copy_(d["dC0"])
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            _call(hG, d)
        for rep in range(3):
            other["dC"].copy_(other["dC0"])
            d["dC"].copy_(d["dC0"])
            torch.cuda.synchronize()
            _call(hG, other)  # eager call on the handle whose buffers the graph uses
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            same = d["dC"].cpu().numpy().tobytes() == eager.tobytes()
            assert same, "%s: replay %d differs from the eager call" % (plan[0], rep)
    finally:
        if fresh:
            hG.close()
        hE.close()
