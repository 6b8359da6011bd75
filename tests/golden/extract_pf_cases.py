"""Writes pf_cases.csv here: the reference's 200 performance cases (TC_PF, alpha = 1, beta = 0)
with the reference GPU baseline time of each, taken verbatim from the reference's data file
ops-blas/test/gemm/cgemm3m/arch35/gpu_baseline.csv (columns id, transa, transb, m, n, k, lda,
ldb, ldc, gpu_ms). The reference times each case as the sum of its kernels' device durations
per API call, mean of 10 calls after 5 warm-up calls (ops-blas/test/gemm/cgemm3m/README.md:
95-115, cgemm3m_profile.h:32-71). Run in a checkout that has the reference at /root/reference.
"""
import csv
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/ops-blas/test/gemm/cgemm3m/arch35/gpu_baseline.csv"

with open(SRC) as f:
    rows = list(csv.DictReader(f))
with open(os.path.join(HERE, "pf_cases.csv"), "w", newline="") as f:
    w = csv.writer(f)
    w.writerow(["id", "transa", "transb", "m", "n", "k", "lda", "ldb", "ldc", "gpu_ms"])
    for r in rows:
        w.writerow([r["id"], r["transa"].replace("ACLBLAS_OP_", ""), r["transb"].replace("ACLBLAS_OP_", ""), r["m"],
                    r["n"], r["k"], r["lda"], r["ldb"], r["ldc"], r["gpu_ms"]])
print("wrote", len(rows), "cases")
