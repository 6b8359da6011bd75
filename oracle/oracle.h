/* oracle.h -- TEST INFRASTRUCTURE ONLY (CPU checker for aclblasCgemm3m). See oracle.c. */
#ifndef CGEMM3M_ORACLE_H_
#define CGEMM3M_ORACLE_H_
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* Matrices are interleaved complex64 (float pairs), column-major; op values 111/112/113. */
int oracle_cgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, float* C, int ldc, int threads);
int oracle_zgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, const float* Cin, int ldc, double* out, int ldo,
                 int threads);
int oracle_verify_component(const float* actual, const float* golden, long long count, double* stats);
int oracle_verify(const float* actual, const float* golden, const float* initial, int m, int n, int ldc,
                  double* stats);
float oracle_uniform(uint64_t seed, uint64_t idx);
void oracle_fill(uint64_t seed, uint64_t start, uint64_t count, float* out);
#ifdef __cplusplus
}
#endif
#endif
