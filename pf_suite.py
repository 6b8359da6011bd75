#!/usr/bin/env python3
"""The reference's performance suite on one MI355X, as a held-out check of the planner: the 200
TC_PF cases (alpha = 1, beta = 0) with their shapes, ops, leading dimensions and the reference's
GPU baseline times, read from the reference's own data file
ops-blas/test/gemm/cgemm3m/arch35/gpu_baseline.csv, whose path is given on the command line
(the file is not part of this repository; the planner's table is tuned on the tuner's own
generated shapes, not on these).

The reference's method (ops-blas/test/gemm/cgemm3m/README.md:95-115, cgemm3m_profile.h:32-71):
5 warm-up and 10 sampled aclblasCgemm3m calls per case, each followed by a stream
synchronisation; a call's time is the SUM of its kernels' device durations from the
profiler's kernel trace (launch gaps excluded); a case's time is the mean over the 10
samples; the reported ratio is GPU_time / device_time (the reference's gate on its own
device: every case >= 0.4, mean >= 0.5). Here the same two steps:

  rocprofv3 --kernel-trace --output-format csv -d OUT -o run -- python3 pf_suite.py run CASES.csv
  python3 pf_suite.py summary OUT/run_kernel_trace.csv CASES.csv profiles/pf_suite_vs_gpu_baseline.md

`run` separates cases with one non-library kernel (a torch fill), so `summary` can split the
trace: per case, the library's kernels form 15 calls of equal kernel count.
"""
import csv
import math
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
WARMUP, SAMPLES = 5, 10


def load_cases(path):
    """Cases of the reference's gpu_baseline.csv (columns id, transa, transb, m, n, k, lda, ldb,
    ldc, gpu_ms; ops written as ACLBLAS_OP_N / _T / _C)."""
    with open(path) as f:
        rows = list(csv.DictReader(f))
    for r in rows:
        for key in ("m", "n", "k", "lda", "ldb", "ldc"):
            r[key] = int(r[key])
        for key in ("transa", "transb"):
            r[key] = r[key].strip()[-1]
        r["gpu_ms"] = float(r["gpu_ms"])
    return rows


def run(path):
    sys.path.insert(0, os.path.join(ROOT, "ascend-open-source-contributions_amd"))
    import torch

    import cgemm3m

    cases = load_cases(path)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261204)
    h = cgemm3m.Handle(stream=torch.cuda.current_stream())
    P = cgemm3m.torch_ptr
    marker = torch.zeros(64, device=dev)
    for idx, c in enumerate(cases):
        ta, tb, m, n, k = c["transa"], c["transb"], c["m"], c["n"], c["k"]
        a_cols = k if ta == "N" else m
        b_cols = n if tb == "N" else k
        A = (torch.rand(2 * c["lda"] * a_cols, generator=gen, device=dev) * 2 - 1).view(torch.complex64)
        B = (torch.rand(2 * c["ldb"] * b_cols, generator=gen, device=dev) * 2 - 1).view(torch.complex64)
        C = torch.zeros(c["ldc"] * n, dtype=torch.complex64, device=dev)
        torch.cuda.synchronize()
        marker.fill_(float(idx))  # case separator in the kernel trace
        for _ in range(WARMUP + SAMPLES):
            cgemm3m.cgemm3m(h, ta, tb, m, n, k, 1.0, P(A), c["lda"], P(B), c["ldb"], 0.0, P(C), c["ldc"])
            torch.cuda.synchronize()
        if idx % 20 == 19:
            print("pf suite: %d / %d cases" % (idx + 1, len(cases)), flush=True)
        del A, B, C
    marker.fill_(-1.0)
    torch.cuda.synchronize()
    h.close()
    print("pf suite: done", flush=True)


def summary(trace_csv, path, out_md):
    cases = load_cases(path)
    groups, cur = [], None
    with open(trace_csv) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))  # file order is not time order
        for r in rows:
            lib = "c3m::" in r["Kernel_Name"]
            dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6  # ms
            if lib:
                if cur is None:
                    cur = []
                    groups.append(cur)
                cur.append((r["Kernel_Name"], dur))
            else:
                cur = None
    if len(groups) != len(cases):
        raise SystemExit("trace holds %d library kernel groups, expected %d cases" % (len(groups), len(cases)))
    calls = WARMUP + SAMPLES
    lines, ratios = [], []
    for c, g in zip(cases, groups):
        if len(g) % calls == 1 and "fill_u64_kernel" in g[0][0]:
            g = g[1:]  # the handle's one-time clear of its arrival-counter pool (a warm-up call)
        if len(g) % calls:
            raise SystemExit("case %s: %d kernels is not a multiple of %d calls" % (c["id"], len(g), calls))
        per = len(g) // calls
        times = [sum(d for _, d in g[i * per:(i + 1) * per]) for i in range(calls)]
        ours = sum(times[WARMUP:]) / SAMPLES
        ratio = c["gpu_ms"] / ours
        ratios.append(ratio)
        kernels = "+".join(sorted({name.split("(")[0].replace("void ", "").replace("c3m::", "").split("<")[0]
                                   for name, _ in g[:per]}))
        lines.append((c, ours, ratio, per, kernels))
    gmean = math.exp(sum(math.log(x) for x in ratios) / len(ratios))
    trace_dir = os.path.dirname(os.path.abspath(trace_csv))
    out = ["# The reference's 200 performance cases on one MI355X (held out from tuning)", "",
           "Cases, GPU baseline times and method from the reference (`ops-blas/test/gemm/cgemm3m/arch35/"
           "gpu_baseline.csv`, `ops-blas/test/gemm/cgemm3m/README.md:95-115`): alpha = 1, beta = 0; per case 5 "
           "warm-up + 10 sampled `aclblasCgemm3m` calls, each followed by a stream synchronisation; a call's time is "
           "the sum of its kernels' device durations (rocprofv3 kernel trace); the case time is the mean of the 10 "
           "samples; ratio = GPU baseline time / MI355X time (the reference's own device gate is ratio >= 0.4 per "
           "case, mean >= 0.5).", "",
           "Held out: the planner's table (`csrc/tuned_table.h`) holds shape classes only, tuned on the tuner's "
           "own seeded shape set (`autotune.generated_shapes`), none of these cases.", "",
           "Commands (MI355X, one GPU; CASES = the reference's `ops-blas/test/gemm/cgemm3m/arch35/gpu_baseline.csv`):",
           "", "```",
           "cd /tmp && TMPDIR=/tmp rocprofv3 --kernel-trace --output-format csv -d %s -o run -- "
           "python3 /root/repo/pf_suite.py run CASES" % trace_dir,
           "python3 /root/repo/pf_suite.py summary %s CASES %s" % (os.path.abspath(trace_csv), os.path.abspath(out_md)),
           "```", "",
           "| | value |", "|---|---|",
           "| cases | %d |" % len(ratios),
           "| ratio min | %.3f |" % min(ratios),
           "| ratio mean (arithmetic) | %.3f |" % (sum(ratios) / len(ratios)),
           "| ratio geometric mean | %.3f |" % gmean,
           "| ratio max | %.3f |" % max(ratios),
           "| cases with ratio >= 1 (MI355X faster) | %d |" % sum(1 for x in ratios if x >= 1.0),
           "| cases with ratio >= 0.4 | %d |" % sum(1 for x in ratios if x >= 0.4),
           "| total GPU baseline / total MI355X time | %.3f |" % (sum(c["gpu_ms"] for c in cases) /
                                                                 sum(o for _, o, _, _, _ in lines)),
           "", "| id | ops | m | n | k | GPU baseline (us) | MI355X (us) | ratio | kernels per call |",
           "|---|---|---|---|---|---|---|---|---|"]
    for c, ours, ratio, per, kernels in lines:
        out.append("| %s | %s%s | %d | %d | %d | %.1f | %.1f | %.2f | %d (%s) |" % (
            c["id"], c["transa"], c["transb"], c["m"], c["n"], c["k"], c["gpu_ms"] * 1e3, ours * 1e3, ratio, per,
            kernels))
    with open(out_md, "w") as f:
        f.write("\n".join(out) + "\n")
    print("\n".join(out[:24]))


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "run":
        run(sys.argv[2])
    elif len(sys.argv) == 5 and sys.argv[1] == "summary":
        summary(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        raise SystemExit("usage: python pf_suite.py run CASES.csv | summary TRACE.csv CASES.csv OUT.md\n"
                         "(CASES.csv: the reference's ops-blas/test/gemm/cgemm3m/arch35/gpu_baseline.csv)")


if __name__ == "__main__":
    main()
