"""CPU tests of the C ABI: the library loads, exports every declared symbol, and the
argument contract (status codes and their order) holds before any device work.

Mirrors the reference's ED rows (ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.csv:307-338)
and NullHandle (cgemm3m_test.cpp:166-173); validation order is cgemm_host.cpp:589-634, 701-730.
"""
import ctypes
import os
import re

import pytest

import cgemm3m
from cgemm3m import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAKE = ctypes.c_void_p(0x1000)  # never dereferenced: every case below fails or returns before use


def test_header_symbols_exported():
    L = _lib.load()
    header = open(os.path.join(ROOT, "include", "aclblas_cgemm3m.h")).read()
    declared = set(re.findall(r"\b(aclblas\w+)\s*\(", header))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_handle_lifecycle():
    L = _lib.load()
    assert L.aclblasCreate(None) == _lib.HANDLE_IS_NULLPTR
    h = ctypes.c_void_p()
    assert L.aclblasCreate(ctypes.byref(h)) == _lib.SUCCESS and h.value
    assert L.aclblasCreate(ctypes.byref(h)) == _lib.INVALID_VALUE  # *handle must be null
    s = ctypes.c_void_p(1)
    assert L.aclblasGetStream(h, ctypes.byref(s)) == _lib.SUCCESS and not s.value
    assert L.aclblasGetStream(h, None) == _lib.INVALID_VALUE
    assert L.aclblasSetStream(None, None) == _lib.HANDLE_IS_NULLPTR
    assert L.aclblasSetWorkspace(h, None, 16) == _lib.INVALID_VALUE
    assert L.aclblasSetWorkspace(h, FAKE, 0) == _lib.INVALID_VALUE
    assert L.aclblasSetWorkspace(None, FAKE, 16) == _lib.HANDLE_IS_NULLPTR
    v = ctypes.c_int()
    assert L.aclblasGetVersion(None, ctypes.byref(v)) == _lib.SUCCESS and v.value == 10000
    assert L.aclblasGetVersion(h, None) == _lib.INVALID_VALUE
    assert L.aclblasDestroy(h) == _lib.SUCCESS
    assert L.aclblasDestroy(None) == _lib.HANDLE_IS_NULLPTR


@pytest.fixture
def h():
    hd = cgemm3m.Handle()
    logs = []
    cb = _lib.LOG_CB(lambda msg: logs.append(msg.decode()))
    hd.L.aclblasSetLogCallback(cb)
    hd.logs = logs
    yield hd
    hd.L.aclblasSetLogCallback(_lib.LOG_CB())
    hd.close()


def call(h, ta="N", tb="N", m=16, n=16, k=16, alpha=1, A=FAKE, lda=16, B=FAKE, ldb=16, beta=0, C=FAKE, ldc=16,
         entry="aclblasCgemm3m"):
    if entry == "aclblasCgemm3m":
        return cgemm3m.cgemm3m(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, check=False)
    op = lambda t: _lib.OP[t] if isinstance(t, str) else t
    return getattr(h.L, entry)(h.h, op(ta), op(tb), m, n, k, _lib.complex_arg(alpha), A, lda, B, ldb,
                               _lib.complex_arg(beta), C, ldc)


def test_null_handle():
    L = _lib.load()
    one, zero = _lib.Complex(1, 0), _lib.Complex(0, 0)
    st = L.aclblasCgemm3m(None, 111, 111, 0, 0, 0, ctypes.byref(one), None, 1, None, 1, ctypes.byref(zero), None, 1)
    assert st == _lib.HANDLE_IS_NULLPTR


@pytest.mark.parametrize("kw,expect", [
    (dict(m=0, lda=1, ldc=1), 0),                          # m0_quick_return
    (dict(n=0), 0),                                        # n0_quick_return
    (dict(m=0, n=0, k=0, lda=1, ldb=1, ldc=1), 0),         # m0n0k0
    (dict(m=0, A=None, lda=1, ldc=1), 0),                  # m0_a_null
    (dict(n=0, B=None), 0),                                # n0_b_null
    (dict(m=0, C=None, lda=1, ldc=1, beta=1), 0),          # m0_c_null
    (dict(k=0, A=None, B=None, ldb=1, beta=1), 0),         # k0 with beta 1: nothing to do
    (dict(alpha=0, A=None, beta=1), 0),                    # alpha0_a_null, beta 1
    (dict(alpha=0, A=None, B=None, beta=1), 0),
    (dict(alpha=None), 3),                                 # null_alpha
    (dict(beta=None), 3),                                  # null_beta
    (dict(A=None), 3),                                     # null_A
    (dict(B=None), 3),                                     # null_B
    (dict(C=None, beta=1), 3),                             # null_C
    (dict(ta=999), 3),                                     # invalid_transa (INVALID_ENUM remapped)
    (dict(tb=999), 3),                                     # invalid_transb
    (dict(lda=8), 3),                                      # invalid_lda_N
    (dict(ta="T", lda=8), 3),                              # invalid_lda_T
    (dict(ldb=8), 3),                                      # invalid_ldb_N
    (dict(tb="T", ldb=8), 3),                              # invalid_ldb_T
    (dict(ldc=8), 3),                                      # invalid_ldc
    (dict(m=-1, lda=1, ldc=1), 3),                         # neg_m
    (dict(n=-1), 3),                                       # neg_n
    (dict(k=-1, ldb=1), 3),                                # neg_k
    (dict(m=-1, A=None, lda=1, ldc=1), 3),                 # neg_m_null_A
    (dict(ta="T", m=20, k=4, lda=3, ldc=20), 3),           # T: lda >= k
    (dict(ta="T", m=20, k=4, lda=4, ldc=20, A=None), 3),   # lda ok, then null A
    (dict(tb="C", n=20, k=4, ldb=19, ldc=16), 3),          # C: ldb >= n
    (dict(ta="N", m=0, n=0, lda=0, ldc=1), 3),             # lda >= max(1, m)
])
@pytest.mark.parametrize("entry", ["aclblasCgemm3m", "aclblasCgemm3mNoAlloc"])
def test_status_contract(h, kw, expect, entry):
    # the no-allocation entry keeps aclblasCgemm3m's argument contract
    assert call(h, entry=entry, **kw) == expect


def test_validation_order(h):
    """Enum errors come before dimension errors, which come before pointer errors."""
    assert call(h, ta=999, m=-1) == 3
    assert call(h, m=-1, alpha=None) == 3
    assert call(None if False else h, m=16, lda=1, C=None) == 3
    assert any("lda=1" in x for x in h.logs)


def test_config_validation(h):
    for cfg in (cgemm3m.config(precision=9), cgemm3m.config(tile=96), cgemm3m.config(k_split=17),
                cgemm3m.config(correction=3), cgemm3m.config(identity=5), cgemm3m.config(variant=6), cgemm3m.config(variant=4),
                cgemm3m.config("bf16x3", variant=5), cgemm3m.config(variant=5, k_split=17),
                cgemm3m.config(variant=5, identity=3), cgemm3m.config("fp16", variant=3),
                cgemm3m.config(tile=16), cgemm3m.config(tile=16, variant=3),
                cgemm3m.config("fp16", variant=4, k_split=2), cgemm3m.config("bf16x3", variant=6, k_split=2),
                    cgemm3m.config("bf16x3", variant=4, tile=64),
                cgemm3m.config("bf16x3", variant=3), cgemm3m.config(tile=64, variant=3)):
        st = cgemm3m.cgemm3m(h, "N", "N", 16, 16, 16, 1, FAKE, 16, FAKE, 16, 0, FAKE, 16, cfg, check=False)
        assert st == _lib.INVALID_VALUE


def test_workspace_size_query(h):
    n = cgemm3m.workspace_size(h, "N", "N", 1024, 1024, 1024, 1, 0)
    # 3 A planes + 3 B planes of 1024 x 1024 FP32 at least
    assert n >= 6 * 1024 * 1024 * 4
    # the one-launch small-problem kernel reads A and B directly: no workspace
    assert cgemm3m.workspace_size(h, "N", "N", 100, 100, 100, 1, 0) == 0
    big = cgemm3m.workspace_size(h, "N", "N", 8192, 8192, 8192, 1, 0.5)
    assert big >= 6 * 8192 * 8192 * 4
    bf = cgemm3m.workspace_size(h, "N", "N", 8192, 8192, 8192, 1, 0, cgemm3m.config("bf16x3"))
    assert bf >= 6 * 8192 * 8192 * 4
    assert cgemm3m.workspace_size(h, "N", "N", 0, 5, 5, 1, 0) == 0
    # padded extents and kernel counters are 32-bit: dimensions near INT_MAX are refused
    # with INVALID_VALUE instead of overflowing (an m x 1 x 1 product with m = 2^31 - 1 is
    # a valid call whose operands fit in HBM)
    top = 2**31 - 1
    for dims in ((top, 1, 1), (1, top, 1), (1, 1, top), (64, 64, top - 100)):
        with pytest.raises(cgemm3m.Cgemm3mError) as err:
            cgemm3m.workspace_size(h, "N", "N", *dims, 1, 0)
        assert err.value.status == _lib.INVALID_VALUE, dims
    assert cgemm3m.workspace_size(h, "N", "N", 1, 1, top - 4096, 1, 0) >= 0


def test_header_is_plain_c99(tmp_path):
    """include/aclblas_cgemm3m.h compiles as strict C99 (a C caller needs nothing else)."""
    import subprocess

    src = tmp_path / "h.c"
    src.write_text('#include "aclblas_cgemm3m.h"\nint main(void) { aclblasCgemm3mConfig c = {0}; (void)c; '
                   'return (int)sizeof(aclblasComplex) - 8; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "h")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(tmp_path / "h")]).returncode == 0


def sb_call(h, ta="N", tb="N", m=16, n=16, k=16, alpha=1, A=FAKE, lda=16, sa=256, B=FAKE, ldb=16, sb=256, beta=0,
            C=FAKE, ldc=16, sc=256, batch=2):
    return cgemm3m.strided_batched(h, ta, tb, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, batch,
                                   check=False)


@pytest.mark.parametrize("kw,expect", [
    (dict(ta=999), 10),                                    # INVALID_ENUM here (README: not remapped)
    (dict(tb=999, m=-1), 10),                              # enum before dimensions
    (dict(batch=-1), 3),
    (dict(m=-1), 3), (dict(k=-1, ldb=1), 3),
    (dict(lda=8), 3), (dict(ldc=8), 3),
    (dict(alpha=None), 3), (dict(beta=None), 3),
    (dict(batch=0, A=None, B=None, C=None), 0),            # quick return after the checks
    (dict(batch=0, alpha=None), 3),                        # ... which still cover scalar pointers
    (dict(m=0, lda=1, ldc=1, A=None, C=None), 0),
    (dict(n=0, B=None, C=None), 0),
    (dict(C=None), 3),
    (dict(A=None), 3), (dict(B=None), 3),
])
def test_strided_batched_status_contract(h, kw, expect):
    """Status rows of aclblasCgemm3mStridedBatched (ops-blas/blas/gemm/README.md, its
    constraints list); every case fails or returns before touching memory."""
    assert sb_call(h, **kw) == expect


def test_strided_batched_null_handle():
    L = _lib.load()
    one = _lib.Complex(1, 0)
    st = L.aclblasCgemm3mStridedBatched(None, 111, 111, 1, 1, 1, ctypes.byref(one), None, 1, 0, None, 1, 0,
                                        ctypes.byref(one), None, 1, 0, 1)
    assert st == _lib.HANDLE_IS_NULLPTR


def cg_call(h, ta="N", tb="N", m=16, n=16, k=16, alpha=1, A=FAKE, lda=16, B=FAKE, ldb=16, beta=0, C=FAKE, ldc=16):
    return cgemm3m.cgemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, check=False)


@pytest.mark.parametrize("kw,expect", [
    (dict(ta=999), 10), (dict(tb=999), 10),                # INVALID_ENUM, not remapped (cgemm_host.cpp:713)
    (dict(ta=999, m=-1), 10),
    (dict(m=-1), 3), (dict(lda=8), 3), (dict(ldc=8), 3), (dict(alpha=None), 3), (dict(beta=None), 3),
    (dict(m=0, lda=1, ldc=1, A=None, C=None), 0),
    (dict(C=None), 3),
    (dict(A=None), 3), (dict(B=None), 3),
    (dict(alpha=0, A=None), 3),                            # k > 0: A required even when alpha == 0
    (dict(k=0, A=None, B=None, ldb=1, beta=1), 0),         # k == 0, beta == 1: nothing to do
    (dict(precision_ok=True), None),
])
def test_cgemm_status_contract(h, kw, expect):
    """aclblasCgemm's contract (ops-blas/blas/gemm/README.md, aclblasCgemm; shared validation
    cgemm_host.cpp:589-634, :687-730 with forceThreeProducts == false)."""
    if kw.get("precision_ok"):
        # the four-product form is FP32-only and LDS-free: other pins are rejected
        for cfg in (cgemm3m.config("bf16x3", identity=3), cgemm3m.config(identity=3, variant=1),
                    cgemm3m.config(identity=4)):
            st = cgemm3m.cgemm3m(h, "N", "N", 16, 16, 16, 1, FAKE, 16, FAKE, 16, 0, FAKE, 16, cfg, check=False)
            assert st == _lib.INVALID_VALUE
        return
    assert cg_call(h, **kw) == expect


def _installed_classes():
    from cgemm3m import autotune

    text = open(autotune.TABLE).read()
    classes = text[text.index("kTunedTable[]"):]
    cls = {}
    for a, b, g, t, p in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", classes):
        if int(t):
            cls[(int(a), int(b), int(g))] = (int(t), int(p))
    return cls


def _planned(h, ta, tb, m, n, k):
    """The plan the library records for this call; without a device the call stops right
    after planning (NOT_INITIALIZED) and launches nothing."""
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    st = cgemm3m.cgemm3m(h, ta, tb, m, n, k, 1, FAKE, lda, FAKE, ldb, 0, FAKE, m, check=False)
    assert st == 1, st  # NOT_INITIALIZED: the plan is made and recorded before the device is needed
    return h.last_plan()


def _expect(plan, tile, parts, k):
    if tile >= 64:  # packed tiles; K partitions capped by the 16-deep stages
        assert plan["path"] in (1, 2) and plan["tile"] == tile, plan
        assert plan["parts"] == min(parts, (k + 31) // 32 * 2), plan
    else:  # the one-launch kernel; parts = k slices (capped by the 16-deep chunks)
        assert plan["path"] == 5 and plan["parts"] == min(parts, (k + 15) // 16), plan


def test_planner_follows_tuned_classes(h):
    """The planner's table lookup against the installed table: a shape of every one-problem
    class gets its class's plan, under every op pair (the table holds classes only)."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("needs a host without a GPU (the call must stop after planning)")
    from cgemm3m import autotune

    checked = 0
    for (mnl, kl, g), (t, p) in sorted(_installed_classes().items()):
        if g:
            continue
        m, n, k = 2 ** (mnl // 2), 2 ** (mnl - mnl // 2), 2 ** kl  # a shape of that class
        assert autotune.bin_of(m, n, k) == (mnl, kl)
        for ops in ("NT", "CN", "TC"):
            _expect(_planned(h, ops[0], ops[1], m, n, k), t, p, k)
        checked += 1
    assert checked >= 20, checked


def test_planner_random_arguments(h):
    """Random sizes, ops and configs (valid and not) through the host planner and the grouped
    layout: only documented statuses come back, and every plain call on valid arguments gets
    a plan (NOT_INITIALIZED without a device means the plan was made)."""
    import random

    import torch

    if torch.cuda.is_available():
        pytest.skip("needs a host without a GPU (the calls must stop after planning)")
    rng = random.Random(11)

    def dim():
        r = rng.random()
        return rng.choice([0, 1, 2, 15, 16, 17, 33]) if r < 0.1 else rng.randint(1, 3000 if r < 0.85 else 70000)

    for _ in range(1500):
        m, n, k = dim(), dim(), dim()
        if m * n > 1 << 28:
            n = max(1, (1 << 28) // m)
        ta, tb = rng.choice("NTC"), rng.choice("NTC")
        plain = rng.random() < 0.3
        cfg = None if plain else cgemm3m.config(
            rng.choice(["fp32", "bf16x3", "bf16x6", "fp16"]), tile=rng.choice([0, 0, 16, 32, 64, 128, 7]),
            k_split=rng.choice([0, 0, 1, 2, 3, 4, 16, 17]), correction=rng.choice([0, -1, 1, 2, 3]),
            identity=rng.choice([0, 1, 2, 3]), variant=rng.choice([0, 0, 1, 2, 3, 4, 5, 6, 7]))
        lda = max(1, m if ta == "N" else k)
        ldb = max(1, k if tb == "N" else n)
        st = cgemm3m.cgemm3m(h, ta, tb, m, n, k, complex(0.5, 1), FAKE, lda, FAKE, ldb, 0.5, FAKE, max(1, m), cfg,
                             check=False)
        assert st in (_lib.SUCCESS, 1, _lib.INVALID_VALUE), (st, m, n, k, ta, tb, cfg and list(cfg._fields_))
        if plain:
            assert st == (_lib.SUCCESS if m == 0 or n == 0 else 1), (st, m, n, k)
    for _ in range(150):
        probs = [dict(ta=rng.choice("NTC"), tb=rng.choice("NTC"), m=rng.randint(0, 600), n=rng.randint(0, 600),
                      k=rng.randint(0, 2500)) for _ in range(rng.randint(1, 40))]
        assert cgemm3m.grouped_workspace_size(h, probs) >= 0

