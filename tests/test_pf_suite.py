"""The performance-suite harness's summary step (no GPU), on a synthetic kernel trace and a
two-case file of this test's own: a case's time is the mean over the 10 sampled calls of the
sum of each call's kernel durations (warm-up calls dropped), cases split at the non-library
separator kernels, ratio = baseline time / measured time."""
import csv
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import pf_suite  # noqa: E402


def _write_cases(path):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["id", "transa", "transb", "m", "n", "k", "lda", "ldb", "ldc", "gpu_ms"])
        w.writerow(["case-a", "ACLBLAS_OP_N", "ACLBLAS_OP_T", 64, 32, 16, 64, 32, 64, 0.030])
        w.writerow(["case-b", "ACLBLAS_OP_C", "ACLBLAS_OP_N", 512, 512, 512, 512, 512, 512, 0.100])


def _write_trace(path):
    """case-a: 15 calls of one kernel, 10 us each except the warm-ups (50 us); case-b: 15 calls
    of two kernels (20 + 30 us), preceded by the handle's one-time pool clear."""
    rows, t = [], 1000000

    def add(name, us):
        nonlocal t
        rows.append({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + int(us * 1000)})
        t += int(us * 1000) + 5000

    add("at::native::fill_kernel", 1)
    for i in range(15):
        add("void c3m::gemm3m_small_kernel(c3m::SmallArgs)", 50 if i < 5 else 10)
    add("at::native::fill_kernel", 1)
    add("c3m::fill_u64_kernel(unsigned long long*, long long)", 3)
    for i in range(15):
        add("c3m::pack_kernel(c3m::PackArgs, c3m::PackArgs, int)", 20)
        add("void c3m::gemm3m_f32_direct_kernel<128, 128>(c3m::GemmArgs)", 30)
    add("at::native::fill_kernel", 1)
    rows.reverse()  # the trace file is not in time order
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerows(rows)


def test_summary_times_and_ratios(tmp_path):
    cases, trace, out = tmp_path / "cases.csv", tmp_path / "trace.csv", tmp_path / "out.md"
    _write_cases(cases)
    _write_trace(trace)
    pf_suite.summary(str(trace), str(cases), str(out))
    text = out.read_text()
    assert "| case-a | NT | 64 | 32 | 16 | 30.0 | 10.0 | 3.00 | 1 (gemm3m_small_kernel) |" in text
    assert "| case-b | CN | 512 | 512 | 512 | 100.0 | 50.0 | 2.00 | 2 (gemm3m_f32_direct_kernel+pack_kernel) |" in text
    assert "| ratio min | 2.000 |" in text and "| ratio max | 3.000 |" in text
    assert "| total GPU baseline / total MI355X time | 2.167 |" in text  # 130 us / 60 us


def test_summary_refuses_a_trace_that_does_not_split_into_the_cases(tmp_path):
    cases, trace, out = tmp_path / "cases.csv", tmp_path / "trace.csv", tmp_path / "out.md"
    _write_cases(cases)
    with open(trace, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerow({"Kernel_Name": "c3m::pack_kernel", "Start_Timestamp": 1, "End_Timestamp": 2})
    with pytest.raises(SystemExit):
        pf_suite.summary(str(trace), str(cases), str(out))
