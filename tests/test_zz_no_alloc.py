"""aclblasCgemm3mNoAlloc: aclblasCgemm3m whose temporary storage is a fixed per-GPU device array.

It has no counterpart in the reference; it must give exactly what aclblasCgemm3m gives on the
same inputs. Two calls in a row on one handle, the second smaller, so the second runs on what
the first left in the array (the K-split partials and energies of the first; the arrival
counters and flag bitmap of the second must be cleared by the call itself). A call whose plan
needs more than the array returns NOT_SUPPORTED and leaves C alone.
"""
import ctypes
import sys

import numpy as np
import pytest

import helpers
import oracle

pytestmark = pytest.mark.gpu

# (ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, seed, expected path)
FIRST = ("N", "T", 300, 260, 5000, 307, 263, 305, (0.7, -0.3), (0.2, 0.9), 41, 2)   # K split + correction
SECOND = ("N", "T", 40, 36, 2600, 43, 39, 41, (1.0, 0.5), (0.5, -0.25), 42, 5)     # one launch, k slices
FUSED = ("C", "N", 1024, 1024, 512, 515, 514, 1030, (1.0, 0.0), (0.0, 0.0), 43, 1)  # fused tile, no split


def _entry(h, name, case, dA, dB, dC):
    from cgemm3m import _lib

    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta = case[:10]
    fn = getattr(h.L, name)
    return fn(h.h, _lib.OP[ta], _lib.OP[tb], m, n, k, _lib.complex_arg(alpha), helpers.ptr(dA), lda,
              helpers.ptr(dB), ldb, _lib.complex_arg(beta), helpers.ptr(dC), ldc)


def _run_pair(h, case):
    import torch

    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, seed, path = case
    print("no-alloc case", case)
    sys.stdout.flush()
    A, B, C = helpers.inputs(ta, tb, m, n, k, lda, ldb, ldc, seed)
    dA, dB = helpers.to_dev(A), helpers.to_dev(B)
    dRef, dOut = helpers.to_dev(C), helpers.to_dev(C)
    assert _entry(h, "aclblasCgemm3m", case, dA, dB, dRef) == 0
    torch.cuda.synchronize()
    assert _entry(h, "aclblasCgemm3mNoAlloc", case, dA, dB, dOut) == 0
    plan = h.last_plan(sync=True)
    torch.cuda.synchronize()
    assert plan["path"] == path, plan
    ref, out = dRef.cpu().numpy(), dOut.cpu().numpy()
    assert ref.tobytes() == out.tobytes(), "no-alloc output differs from aclblasCgemm3m's"
    gold = oracle.cgemm(ta, tb, m, n, k, complex(*alpha), A, lda, B, ldb, complex(*beta), C, ldc)
    helpers.check_against_oracle(out, gold, C, m, n, ldc, "no-alloc %s%s %dx%dx%d" % (ta, tb, m, n, k))


def test_no_alloc_two_calls_second_smaller(gpu):
    _run_pair(gpu, FIRST)
    _run_pair(gpu, SECOND)


def test_no_alloc_fused_tile(gpu):
    _run_pair(gpu, FUSED)


def test_no_alloc_rejects_a_plan_beyond_the_array(gpu):
    import torch

    from cgemm3m import _lib

    m = n = k = 4096  # packed planes alone take about 400 MiB
    need = ctypes.c_size_t()
    assert gpu.L.aclblasCgemm3mGetWorkspaceSize(gpu.h, 111, 111, m, n, k, _lib.complex_arg(1), _lib.complex_arg(0),
                                               None, ctypes.byref(need)) == 0
    assert need.value > 64 << 20
    dA = torch.zeros(m * k, dtype=torch.complex64, device="cuda")
    dB = torch.zeros(k * n, dtype=torch.complex64, device="cuda")
    dC = torch.full((m * n,), 3 + 4j, dtype=torch.complex64, device="cuda")
    case = ("N", "N", m, n, k, m, k, m, (1.0, 0.0), (0.0, 0.0))
    assert _entry(gpu, "aclblasCgemm3mNoAlloc", case, dA, dB, dC) == _lib.NOT_SUPPORTED
    torch.cuda.synchronize()
    assert bool((dC == (3 + 4j)).all())
    del dA, dB, dC
    # the entry still works after a rejection
    _run_pair(gpu, SECOND)


def test_no_alloc_generated_sweep(gpu):
    """Every 9th case of the generated sweep (sizes 1 to 1024, every op pair, scalar sets with
    alpha = 0, padded leading dimensions), one after another on the same array: each output
    byte-identical to aclblasCgemm3m's on the same inputs. A case whose plan needs more than
    the 64 MiB array must be refused with NOT_SUPPORTED and leave C alone."""
    import torch

    import cgemm3m
    from cgemm3m import _lib
    from suite_cases import suite_cases

    served = 0
    for c in [x for x in suite_cases() if x["cat"] == "sweep"][::9]:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        case = (ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], (c["alpha"].real, c["alpha"].imag),
                (c["beta"].real, c["beta"].imag))
        A, B, C = helpers.inputs(ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], c["seed"])
        dA, dB = helpers.to_dev(A), helpers.to_dev(B)
        dRef, dOut = helpers.to_dev(C), helpers.to_dev(C)
        need = cgemm3m.workspace_size(gpu, ta, tb, m, n, k, c["alpha"], c["beta"])
        assert _entry(gpu, "aclblasCgemm3m", case, dA, dB, dRef) == 0
        st = _entry(gpu, "aclblasCgemm3mNoAlloc", case, dA, dB, dOut)
        torch.cuda.synchronize()
        if need > (64 << 20):
            assert st == _lib.NOT_SUPPORTED, (case, need, st)
            assert dOut.cpu().numpy().tobytes() == C.tobytes(), "a refused call wrote C"
            continue
        assert st == 0, (case, need, st)
        served += 1
        same = dRef.cpu().numpy().tobytes() == dOut.cpu().numpy().tobytes()
        assert same, "no-alloc output differs from aclblasCgemm3m's: %s" % (case,)
    assert served >= 60, served
