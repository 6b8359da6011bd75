"""Grouped batch of uneven complex GEMMs: one packing launch + persistent product launches,
checked per problem against the Netlib oracle, for byte-identical reruns, and replayed from
a hipGraph (torch.cuda.CUDAGraph capture on torch's capture stream)."""
import sys

import numpy as np
import pytest

import oracle
from helpers import inputs, ptr, region, to_dev

pytestmark = pytest.mark.gpu


def make_problems(count, seed, general_scalars=False):
    rng = np.random.default_rng(seed)
    probs, host = [], []
    for i in range(count):
        ta, tb = str(rng.choice(list("NTC"))), str(rng.choice(list("NTC")))
        m = int(rng.integers(2, 33)) * 16
        n = int(rng.integers(2, 33)) * 16
        k = int(rng.integers(1, 33)) * 64
        lda = (m if ta == "N" else k) + int(rng.integers(0, 3))
        ldb = (k if tb == "N" else n) + int(rng.integers(0, 3))
        ldc = m + int(rng.integers(0, 3))
        alpha, beta = (complex(0.7, -0.3), complex(0.2, 0.9)) if general_scalars else (1.0, 0.0)
        A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, seed * 1000 + 3 * i)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, A=ptr(dA), lda=lda, B=ptr(dB), ldb=ldb,
                          beta=beta, C=ptr(dC), ldc=ldc))
        host.append(dict(A=A, B=B, C=C, dA=dA, dB=dB, dC=dC))
    return probs, host


def check_all(probs, host, label):
    for i, (p, hst) in enumerate(zip(probs, host)):
        out = hst["dC"].cpu().numpy()
        gold = oracle.cgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], hst["A"], p["lda"], hst["B"],
                            p["ldb"], p["beta"], hst["C"], p["ldc"])
        ok, worst = oracle.reltol_check(region(out, p["m"], p["n"], p["ldc"]), region(gold, p["m"], p["n"], p["ldc"]))
        verdict, stats = oracle.verify(out, gold, hst["C"], p["m"], p["n"], p["ldc"])
        assert ok and verdict, (label, i, p["ta"], p["tb"], p["m"], p["n"], p["k"], worst, stats)


@pytest.mark.parametrize("general", [False, True])
def test_grouped_vs_oracle(gpu, general):
    import torch

    import cgemm3m

    probs, host = make_problems(48, 7 + general, general)
    print("grouped problems", [(p["ta"] + p["tb"], p["m"], p["n"], p["k"]) for p in probs[:8]], "...")
    sys.stdout.flush()
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    first = [h["dC"].cpu().numpy().copy() for h in host]
    check_all(probs, host, "grouped")
    for h in host:
        h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    for i, (h, f) in enumerate(zip(host, first)):
        same = h["dC"].cpu().numpy().tobytes() == f.tobytes()
        assert same, "grouped rerun differs in problem %d" % i


@pytest.mark.parametrize("tile", [0, 16])
def test_grouped_one_launch_members(gpu, tile):
    """Members on the one-launch kernel (pinned with variant 5; tile 16 is its single-wave
    form): the grouped call gives the bytes of per-problem calls with the same plan, and
    meets the oracle."""
    import torch

    import cgemm3m

    probs, host = make_problems(12, 51 + tile, True)
    cfg = cgemm3m.config(variant=5, tile=tile, k_split=1, correction=-1)
    cgemm3m.grouped(gpu, probs, cfg=cfg)
    torch.cuda.synchronize()
    grouped = [h["dC"].cpu().numpy().copy() for h in host]
    check_all(probs, host, "grouped-one-launch")
    for i, (p, h) in enumerate(zip(probs, host)):
        h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
        cgemm3m.cgemm3m(gpu, p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                        p["ldb"], p["beta"], p["C"], p["ldc"], cfg)
        torch.cuda.synchronize()
        plan = gpu.last_plan()
        assert plan["path"] == 5, plan
        same = h["dC"].cpu().numpy().tobytes() == grouped[i].tobytes()
        assert same, "member %d differs from its one-problem call" % i


@pytest.mark.parametrize("tile,parts", [(128, 3), (64, 2)])
def test_grouped_k_split_members(gpu, tile, parts):
    """Members pinned to a K split: the quad launch writes each partition's planes and the
    grouped merge combines them in partition order; every member equals its one-problem call
    with the same plan byte for byte and meets the oracle."""
    import torch

    import cgemm3m

    probs, host = make_problems(10, 71 + tile + parts, True)
    cfg = cgemm3m.config(tile=tile, k_split=parts, correction=-1)
    cgemm3m.grouped(gpu, probs, cfg=cfg)
    torch.cuda.synchronize()
    grouped = [h["dC"].cpu().numpy().copy() for h in host]
    check_all(probs, host, "grouped-k-split")
    one = cgemm3m.config(tile=tile, k_split=parts, correction=-1, variant=3 if tile == 128 else 1)
    for i, (p, h) in enumerate(zip(probs, host)):
        h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
        cgemm3m.cgemm3m(gpu, p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                        p["ldb"], p["beta"], p["C"], p["ldc"], one)
        torch.cuda.synchronize()
        plan = gpu.last_plan()
        assert plan["parts"] == min(parts, (p["k"] + 31) // 32 * 32 // 16), plan
        same = h["dC"].cpu().numpy().tobytes() == grouped[i].tobytes()
        assert same, "member %d (%s%s %d x %d x %d) differs from its one-problem call" % (
            i, p["ta"], p["tb"], p["m"], p["n"], p["k"])


def test_grouped_edge_problems(gpu):
    """Empty, k = 0 and alpha = 0 members are handled inside one grouped call."""
    import torch

    import cgemm3m

    probs, host = make_problems(5, 21, True)
    probs[1]["m"] = 0
    probs[2]["k"] = 0
    probs[2]["beta"] = complex(0.5, -0.25)
    probs[3]["alpha"] = 0.0
    probs[3]["beta"] = 0.0
    keep1 = host[1]["dC"].cpu().numpy().copy()
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    same = host[1]["dC"].cpu().numpy().tobytes() == keep1.tobytes()
    assert same
    for i in (0, 2, 3, 4):
        p, hst = probs[i], host[i]
        gold = oracle.cgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], hst["A"], p["lda"], hst["B"],
                            p["ldb"], p["beta"], hst["C"], p["ldc"])
        ok, stats = oracle.verify(hst["dC"].cpu().numpy(), gold, hst["C"], p["m"], p["n"], p["ldc"])
        assert ok, (i, stats)


def test_grouped_invalid_member_launches_nothing(gpu):
    import torch

    import cgemm3m

    probs, host = make_problems(3, 31)
    probs[2]["lda"] = 1  # invalid leading dimension in the last member
    before = [h["dC"].cpu().numpy().copy() for h in host]
    st = cgemm3m.grouped(gpu, probs, check=False)
    torch.cuda.synchronize()
    assert st == 3
    for h, b in zip(host, before):
        same = h["dC"].cpu().numpy().tobytes() == b.tobytes()
        assert same


def test_grouped_hipgraph_replay(gpu):
    import torch

    import cgemm3m

    probs, host = make_problems(24, 41)
    side = torch.cuda.Stream()
    h = cgemm3m.Handle(stream=side)
    with torch.cuda.stream(side):
        cgemm3m.grouped(h, probs)  # eager call sizes the workspace and stages descriptors
    torch.cuda.synchronize()
    eager = [x["dC"].cpu().numpy().copy() for x in host]
    check_all(probs, host, "eager")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.set_stream(torch.cuda.current_stream())
        cgemm3m.grouped(h, probs)
    torch.cuda.synchronize()
    failures = []
    for rep in range(3):
        for x in host:
            # eager work between replays: restore C (padding rows included, which must stay untouched)
            x["dC"].copy_(torch.from_numpy(x["C"]).cuda())
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        bad = []
        for i, (x, e) in enumerate(zip(host, eager)):
            got = x["dC"].cpu().numpy()
            if got.tobytes() != e.tobytes():
                p = probs[i]
                reg = region(got, p["m"], p["n"], p["ldc"])
                zero_cols = int((np.abs(reg).sum(axis=0) == 0).sum())
                bad.append((i, p["ta"] + p["tb"], p["m"], p["n"], p["k"], int((got != e).sum()), zero_cols,
                            float(np.abs(got - e).max())))
        print("replay", rep, "differing (idx, ops, m, n, k, elements, zero columns, max diff):", bad[:10])
        sys.stdout.flush()
        failures += bad
    # the same call eagerly again, after the graph
    h.set_stream(side)
    for x in host:
        x["dC"].copy_(torch.from_numpy(x["C"]).cuda())
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        cgemm3m.grouped(h, probs)
    torch.cuda.synchronize()
    again = [i for i, (x, e) in enumerate(zip(host, eager)) if x["dC"].cpu().numpy().tobytes() != e.tobytes()]
    print("eager after graph differs in", again)
    h.close()
    assert not again
    assert not failures, "%d problem replays differ" % len(failures)


@pytest.mark.parametrize("user_ws", [True, False])
def test_grouped_first_call_captured(gpu, user_ws):
    """The first grouped call on a fresh handle is itself captured, with the caller's
    workspace or the library's default one (allocated during the capture). The batch holds
    one-launch members (32 x 32) and packed 128-tile members (512 x 512), so the capture
    forks the second stream; every replay matches an eager call on another handle byte for
    byte. With the default workspace, an eager call that outgrows it runs between two
    replays: the captured buffer must stay valid."""
    import ctypes

    import torch

    import cgemm3m

    probs, host = make_problems(10, 61)
    for i, p in enumerate(probs):
        big = i % 2 == 0
        ta, tb = p["ta"], p["tb"]
        m = n = 512 if big else 32
        k = 1024 if big else 128
        lda, ldb, ldc = (m if ta == "N" else k), (k if tb == "N" else n), m
        A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, 9000 + 3 * i)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs[i] = dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=1.0, A=ptr(dA), lda=lda, B=ptr(dB), ldb=ldb, beta=0.0,
                        C=ptr(dC), ldc=ldc)
        host[i] = dict(A=A, B=B, C=C, dA=dA, dB=dB, dC=dC)
    ref = cgemm3m.Handle()
    cgemm3m.grouped(ref, probs)
    torch.cuda.synchronize()
    eager = [x["dC"].cpu().numpy().copy() for x in host]
    ref.close()
    check_all(probs, host, "eager")
    for x in host:
        x["dC"].copy_(torch.from_numpy(x["C"]).cuda())
    torch.cuda.synchronize()
    h = cgemm3m.Handle()
    need = cgemm3m.grouped_workspace_size(h, probs)
    ws = torch.empty(max(need, 256), dtype=torch.uint8, device="cuda")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        h.set_stream(torch.cuda.current_stream())  # (resets the workspace to the default)
        if user_ws:
            st = h.L.aclblasSetWorkspace(h.h, ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws.numel()))
            assert st == 0
        cgemm3m.grouped(h, probs)
    torch.cuda.synchronize()
    for rep in range(2):
        if rep == 1 and not user_ws:
            # 4096^3 on the packed path needs far more than the default workspace holds
            big = 4096
            X = torch.zeros(big * big, dtype=torch.complex64, device="cuda")
            Y = torch.zeros(big * big, dtype=torch.complex64, device="cuda")
            Z = torch.zeros(big * big, dtype=torch.complex64, device="cuda")
            h.set_stream(torch.cuda.current_stream())
            cgemm3m.cgemm3m(h, "N", "N", big, big, big, 1.0, ptr(X), big, ptr(Y), big, 0.0, ptr(Z), big)
            torch.cuda.synchronize()
            assert h.last_plan()["workspace"] > need
            del X, Y, Z
        for x in host:
            x["dC"].copy_(torch.from_numpy(x["C"]).cuda())
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        bad = [i for i, (x, e) in enumerate(zip(host, eager)) if x["dC"].cpu().numpy().tobytes() != e.tobytes()]
        print("replay", rep, "differing problems:", bad)
        sys.stdout.flush()
        assert not bad
    del graph
    h.close()


def test_single_call_hipgraph_replay(gpu):
    """aclblasCgemm3m (FP32, fused path with correction) captured once, replayed 3 times."""
    import torch

    import cgemm3m

    m, n, k = 384, 256, 1536
    A, B, C = inputs("N", "T", m, n, k, m, n, m, 777)
    dA, dB, dC, dC0 = to_dev(A), to_dev(B), to_dev(C), to_dev(C)
    side = torch.cuda.Stream()
    h = cgemm3m.Handle(stream=side)
    with torch.cuda.stream(side):
        cgemm3m.cgemm3m(h, "N", "T", m, n, k, 1.0, ptr(dA), m, ptr(dB), n, 0.5, ptr(dC), m)
    torch.cuda.synchronize()
    eager = dC.cpu().numpy().copy()
    graph = torch.cuda.CUDAGraph()
    dC.copy_(dC0)
    with torch.cuda.graph(graph):
        h.set_stream(torch.cuda.current_stream())
        cgemm3m.cgemm3m(h, "N", "T", m, n, k, 1.0, ptr(dA), m, ptr(dB), n, 0.5, ptr(dC), m)
    for rep in range(3):
        dC.copy_(dC0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = dC.cpu().numpy()
        ndiff = int((got != eager).sum())
        print("single replay", rep, "differing elements", ndiff)
        assert ndiff == 0
    h.close()


def test_grouped_many_distinct_pointer_calls(gpu):
    """More grouped calls with new buffers than the handle keeps descriptor slots (32): each
    call's descriptors name new pointers, so slots are reused once their last call is done.
    Every call must give the bytes of the first call on the same data."""
    import torch

    import cgemm3m

    shapes = [("N", "N", 96, 80, 128), ("C", "T", 130, 70, 200), ("T", "N", 200, 160, 64)]
    data = []
    for i, (ta, tb, m, n, k) in enumerate(shapes):
        lda, ldb = (m if ta == "N" else k), (k if tb == "N" else n)
        data.append((ta, tb, m, n, k, lda, ldb) + inputs(ta, tb, m, n, k, lda, ldb, m, 9800 + 3 * i))
    reference = None
    for call in range(40):
        keep, probs = [], []
        for ta, tb, m, n, k, lda, ldb, A, B, C in data:
            dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
            keep.append(dC)
            keep += [dA, dB]
            probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=complex(0.7, -0.3), A=ptr(dA), lda=lda, B=ptr(dB),
                              ldb=ldb, beta=complex(0.2, 0.9), C=ptr(dC), ldc=m))
        cgemm3m.grouped(gpu, probs)
        torch.cuda.synchronize()
        outs = [keep[3 * i].cpu().numpy() for i in range(len(data))]
        if reference is None:
            reference = outs
            for (ta, tb, m, n, k, lda, ldb, A, B, C), out in zip(data, outs):
                gold = oracle.cgemm(ta, tb, m, n, k, complex(0.7, -0.3), A, lda, B, ldb, complex(0.2, 0.9), C, m)
                ok, worst = oracle.reltol_check(region(out, m, n, m), region(gold, m, n, m))
                assert ok, worst
        for i, (o, r) in enumerate(zip(outs, reference)):
            assert o.tobytes() == r.tobytes(), "call %d member %d differs" % (call, i)
        del keep


def test_create_reserves_default_workspace(gpu):
    """aclblasCreate reserves the 32 MiB default workspace on the device (cann_ops_blas.h:32-46);
    a call that fits in it allocates nothing more."""
    import torch

    import cgemm3m

    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    h = cgemm3m.Handle()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 >= 32 << 20, (free0, free1)
    m = n = k = 512
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 9900)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    cfg = cgemm3m.config(tile=64, k_split=1)
    assert cgemm3m.workspace_size(h, "N", "N", m, n, k, 1, 0, cfg) <= 32 << 20
    free2 = torch.cuda.mem_get_info()[0]
    cgemm3m.cgemm3m(h, "N", "N", m, n, k, 1.0, ptr(dA), m, ptr(dB), k, 0.0, ptr(dC), m, cfg)
    torch.cuda.synchronize()
    assert h.last_plan()["path"] == 1
    free3 = torch.cuda.mem_get_info()[0]
    assert free3 == free2, (free2, free3)
    h.close()


@pytest.mark.parametrize("pin", [dict(), dict(tile=128, k_split=3), dict(tile=64, k_split=1, identity=2),
                                 dict(tile=64, k_split=2, correction=1)])
def test_grouped_raw_reads_match_packed_planes(gpu, pin):
    """The persistent kernel reads its members' operands as stored and forms the plane values in
    registers; the form that packs them into planes with one launch first (variant 7) must give
    the same bytes: same plane values, same MFMA order per accumulator (all 9 ops, uneven sizes,
    K splits, both 3M identities, the correction on)."""
    import torch

    import cgemm3m

    probs, host = make_problems(40, 91 + len(pin), True)
    outs = []
    for variant in (0, 7):
        for h in host:
            h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
        cgemm3m.grouped(gpu, probs, cfg=cgemm3m.config(variant=variant, **pin))
        torch.cuda.synchronize()
        outs.append([h["dC"].cpu().numpy().copy() for h in host])
    for i, (r, p) in enumerate(zip(*outs)):
        assert r.tobytes() == p.tobytes(), "member %d (%s%s %d x %d x %d) differs" % (
            i, probs[i]["ta"], probs[i]["tb"], probs[i]["m"], probs[i]["n"], probs[i]["k"])
    check_all(probs, host, "grouped-raw")


def test_grouped_edge_fragments_match_full_blocks(gpu):
    """Blocks at a member's last block row or column with at most 32 outputs left run on one
    32-wide fragment (32 x 64, 64 x 32 and 32 x 32 wave shapes); the packed form (variant 7)
    computes whole 64 x 64 blocks. Both must give the same bytes: sizes on either side of each
    32 and 64 boundary, odd sizes, k tails, all nine op pairs, the correction on and off,
    general scalars; then the oracle's verdict per member."""
    import torch

    import cgemm3m

    sizes = [(1, 1, 7), (31, 33, 100), (32, 64, 64), (33, 65, 129), (63, 31, 1100), (64, 96, 1024),
             (65, 97, 1500), (96, 32, 250), (97, 161, 2049), (160, 33, 1030), (200, 224, 300), (17, 250, 1111)]
    rng = np.random.default_rng(77)
    probs, host = [], []
    for i, (m, n, k) in enumerate(sizes):
        ta, tb = "NTC"[i % 3], "NTC"[(i // 3) % 3]
        lda = (m if ta == "N" else k) + int(rng.integers(0, 3))
        ldb = (k if tb == "N" else n) + int(rng.integers(0, 3))
        ldc = m + 1
        A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, 5000 + i)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=complex(0.7, -0.3), A=ptr(dA), lda=lda, B=ptr(dB),
                          ldb=ldb, beta=complex(0.2, 0.9), C=ptr(dC), ldc=ldc))
        host.append(dict(A=A, B=B, C=C, dA=dA, dB=dB, dC=dC))
    for correction in (0, -1):
        outs = []
        for variant in (0, 7):
            for h in host:
                h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
            # tile 64 pinned: every member on the persistent kernel (none on the one-launch one)
            cgemm3m.grouped(gpu, probs, cfg=cgemm3m.config(variant=variant, correction=correction, tile=64))
            torch.cuda.synchronize()
            outs.append([h["dC"].cpu().numpy().copy() for h in host])
        for i, (r, p) in enumerate(zip(*outs)):
            same = r.tobytes() == p.tobytes()
            assert same, ("member differs", correction, i, probs[i]["ta"], probs[i]["tb"], sizes[i])
        check_all(probs, host, "edge-fragments correction=%d" % correction)


def test_grouped_k_split_rule_follows_batch_size(gpu):
    """Corrected members (k >= 1024) drop the tuned table's K split only in a batch that fills
    the GPU without it. Seen through the strided entry, whose overlapping outputs (strideC = 0)
    run batch after batch with the member plan, recorded as the last plan: 2 batches of
    160 x 144 x 2048 (9 blocks of 64 x 64 each) keep the table's partitions, 700 (6300 blocks)
    run unsplit. A small grouped batch of that shape still meets the oracle."""
    import torch

    import cgemm3m

    m, n, k = 160, 144, 2048
    A = to_dev(inputs("N", "T", m, n, k, m, n, m, 4040)[0])
    B = to_dev(inputs("N", "T", m, n, k, m, n, m, 4041)[1])
    parts = {}
    for batch in (2, 700):
        C = torch.zeros(m * n, dtype=torch.complex64, device="cuda")
        cgemm3m.strided_batched(gpu, "N", "T", m, n, k, 1.0, ptr(A), m, 0, ptr(B), n, 0, 0.0, ptr(C), m, 0, batch)
        torch.cuda.synchronize()
        plan = gpu.last_plan()
        assert plan["correction"] == 1, plan
        parts[batch] = plan["parts"]
    assert parts[2] > 1 and parts[700] == 1, parts

    probs, host = [], []
    for i in range(2):
        A, B, C = inputs("N", "T", m, n, k, m, n, m, 4050 + i)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs.append(dict(ta="N", tb="T", m=m, n=n, k=k, alpha=complex(0.7, -0.3), A=ptr(dA), lda=m, B=ptr(dB),
                          ldb=n, beta=complex(0.2, 0.9), C=ptr(dC), ldc=m))
        host.append(dict(A=A, B=B, C=C, dA=dA, dB=dB, dC=dC))
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    check_all(probs, host, "small batch, K split kept")


def test_grouped_more_members_than_grid_y(gpu):
    """A batch with more corrected members than a launch's grid y holds (65535): 70000 members
    of 16 x 16 x 1024 on the persistent kernel (tile 64 pinned), all reading the same A and B,
    each writing its own C. The bitmap-clear and recompute launches take the members in chunks;
    every member must equal the first byte for byte, and the first must meet the oracle."""
    import torch

    import cgemm3m

    count, m, n, k = 70000, 16, 16, 1024
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 777)
    dA, dB = to_dev(A), to_dev(B)
    dC = torch.zeros(count * m * n, dtype=torch.complex64, device="cuda")
    base = dC.data_ptr()
    probs = [dict(ta="N", tb="N", m=m, n=n, k=k, alpha=complex(0.7, -0.3), A=ptr(dA), lda=m, B=ptr(dB), ldb=k,
                  beta=0.0, C=base + i * m * n * 8, ldc=m) for i in range(count)]
    cgemm3m.grouped(gpu, probs, cfg=cgemm3m.config(tile=64, k_split=1))
    torch.cuda.synchronize()
    out = dC.cpu().numpy().reshape(count, m * n)
    first = out[0].tobytes()
    same = all(out[i].tobytes() == first for i in range(count))
    assert same, "members differ"
    gold = oracle.cgemm("N", "N", m, n, k, complex(0.7, -0.3), A, m, B, k, 0.0, np.zeros(m * n, np.complex64), m)
    ok, stats = oracle.verify(out[0], gold, np.zeros(m * n, np.complex64), m, n, m)
    assert ok, stats
