/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY. CPU restatement of the arithmetic the reference
 * checks aclblasCgemm3m against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker; the product never calls it.
 *
 * Functions and what each follows:
 *   oracle_cgemm      Netlib reference-BLAS 3.10.0 cgemm loop order (the third-party oracle
 *                     the reference pins: ops-blas/test/gemm/cgemm3m/README.md:68-81, call
 *                     site ops-blas/test/gemm/cgemm3m/cgemm3m_golden.h:46-48), wrapped in the
 *                     golden's own k==0 / alpha==0 handling (cgemm3m_golden.h:28-45).
 *                     Pinned by tests/test_oracle.py: brute force at tiny sizes, closed-form
 *                     ordering probes per branch, integer-exact inputs, the reference KATs.
 *   oracle_zgemm      FP64 complex GEMM of the FP32 inputs (error reporting).
 *                     Pinned against exact rational sums at tiny sizes.
 *   oracle_verify     The reference verifier's verdict (cgemm3m_verify.h:22-84): per real
 *                     component, finite match if |a-g| <= (1+|g|)/8192, pass if >=99% match
 *                     and max error <= max(0.01, 32 ulp(g at the max-error point)); NaN and
 *                     Inf equivalent, opposite infinities not; padding rows byte-identical.
 *                     Pinned by the reference's verifier KATs (cgemm3m_test.cpp:93-164).
 *   oracle_fill       Counter-based uniform[-1,1) generator (splitmix64 finaliser), the same
 *                     stream as oracle/__init__.py's numpy version. Pinned by cross-check.
 *
 * Build: -O2 -ffp-contract=off, no fast-math, so every complex product is the naive
 * (ar*br - ai*bi, ar*bi + ai*br) with one rounding per operation, as gfortran emits for
 * COMPLEX arithmetic at -O2 on x86-64 without FMA.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    float r, i;
} cf;

static inline cf cmul(cf a, cf b)
{
    cf z;
    z.r = a.r * b.r - a.i * b.i;
    z.i = a.r * b.i + a.i * b.r;
    return z;
}
static inline cf cadd(cf a, cf b)
{
    cf z;
    z.r = a.r + b.r;
    z.i = a.i + b.i;
    return z;
}
static inline cf cconj(cf a)
{
    a.i = -a.i;
    return a;
}

enum { OP_N = 111, OP_T = 112, OP_C = 113 };

/* Columns j are independent in every branch of the Netlib loop nests, so splitting j
 * across threads gives bitwise the same result as the serial loop. */
int oracle_cgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, float* C, int ldc, int threads)
{
    const cf al = {alpha[0], alpha[1]};
    const cf be = {beta[0], beta[1]};
    const cf* a = (const cf*)A;
    const cf* b = (const cf*)B;
    cf* c = (cf*)C;
    const int nota = transa == OP_N, conja = transa == OP_C;
    const int notb = transb == OP_N, conjb = transb == OP_C;
    const int beta_zero = be.r == 0.0f && be.i == 0.0f;
    const int beta_one = be.r == 1.0f && be.i == 0.0f;
    if (m == 0 || n == 0) return 0;
    /* cgemm3m_golden.h:28-45: k == 0 or alpha == 0 -> beta == 1 untouched, beta == 0 zeros,
     * else beta*C; Netlib's own alpha == 0 branch computes the same values. */
    if (k == 0 || (al.r == 0.0f && al.i == 0.0f)) {
        if (beta_one) return 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) {
                cf* v = &c[(size_t)j * ldc + i];
                if (beta_zero) {
                    v->r = 0.0f;
                    v->i = 0.0f;
                } else {
                    *v = cmul(be, *v);
                }
            }
        return 0;
    }
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int j = 0; j < n; ++j) {
        cf* cj = &c[(size_t)j * ldc];
        if (nota) {
            /* C(:,j) := beta*C(:,j) (zero when beta == 0, untouched when beta == 1), then
             * for l: temp = alpha*op(B)(l,j); C(:,j) += temp*A(:,l). */
            if (beta_zero) {
                for (int i = 0; i < m; ++i) cj[i].r = cj[i].i = 0.0f;
            } else if (!beta_one) {
                for (int i = 0; i < m; ++i) cj[i] = cmul(be, cj[i]);
            }
            for (int l = 0; l < k; ++l) {
                cf bl;
                if (notb)
                    bl = b[(size_t)j * ldb + l];
                else
                    bl = b[(size_t)l * ldb + j];
                if (conjb) bl = cconj(bl);
                const cf temp = cmul(al, bl);
                const cf* al_col = &a[(size_t)l * lda];
                for (int i = 0; i < m; ++i) cj[i] = cadd(cj[i], cmul(temp, al_col[i]));
            }
        } else {
            /* temp = sum_l op(A)(i,l)*op(B)(l,j) in l order; C = alpha*temp (+ beta*C). */
            for (int i = 0; i < m; ++i) {
                cf temp = {0.0f, 0.0f};
                const cf* ai = &a[(size_t)i * lda];
                for (int l = 0; l < k; ++l) {
                    cf x = ai[l];
                    if (conja) x = cconj(x);
                    cf y;
                    if (notb)
                        y = b[(size_t)j * ldb + l];
                    else
                        y = b[(size_t)l * ldb + j];
                    if (conjb) y = cconj(y);
                    temp = cadd(temp, cmul(x, y));
                }
                if (beta_zero)
                    cj[i] = cmul(al, temp);
                else
                    cj[i] = cadd(cmul(al, temp), cmul(be, cj[i]));
            }
        }
    }
    return 0;
}

typedef struct {
    double r, i;
} cd;

int oracle_zgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, const float* Cin, int ldc, double* out, int ldo,
                 int threads)
{
    const cd al = {alpha[0], alpha[1]};
    const cd be = {beta[0], beta[1]};
    const int beta_zero = beta[0] == 0.0f && beta[1] == 0.0f;
    const cf* a = (const cf*)A;
    const cf* b = (const cf*)B;
    const cf* c = (const cf*)Cin;
    cd* o = (cd*)out;
    if (m == 0 || n == 0) return 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            double sr = 0.0, si = 0.0;
            for (int l = 0; l < k; ++l) {
                cf x = transa == OP_N ? a[(size_t)l * lda + i] : a[(size_t)i * lda + l];
                cf y = transb == OP_N ? b[(size_t)j * ldb + l] : b[(size_t)l * ldb + j];
                double xr = x.r, xi = transa == OP_C ? -(double)x.i : (double)x.i;
                double yr = y.r, yi = transb == OP_C ? -(double)y.i : (double)y.i;
                sr += xr * yr - xi * yi;
                si += xr * yi + xi * yr;
            }
            cd v;
            v.r = al.r * sr - al.i * si;
            v.i = al.r * si + al.i * sr;
            if (!beta_zero) {
                const cf cv = c[(size_t)j * ldc + i];
                v.r += be.r * cv.r - be.i * cv.i;
                v.i += be.r * cv.i + be.i * cv.r;
            }
            o[(size_t)j * ldo + i] = v;
        }
    }
    return 0;
}

/* One real component of the reference verifier. */
typedef struct {
    long long finite, matched, nonfinite, nonfinite_failed;
    double max_error, max_error_limit;
    float max_error_golden;
} comp_t;

static void comp_init(comp_t* r)
{
    memset(r, 0, sizeof(*r));
    r->max_error_limit = 0.01;
}

static void comp_add(comp_t* r, float actual, float golden)
{
    if (!isfinite(actual) || !isfinite(golden)) {
        r->nonfinite++;
        const int both = !isfinite(actual) && !isfinite(golden);
        const int equiv = both && (isnan(actual) || isnan(golden) || actual == golden);
        r->nonfinite_failed += equiv ? 0 : 1;
        return;
    }
    r->finite++;
    const double err = fabs((double)actual - (double)golden);
    r->matched += err <= (1.0 + fabs((double)golden)) / 8192.0 ? 1 : 0;
    if (err > r->max_error) {
        r->max_error = err;
        r->max_error_golden = golden;
        int e = golden == 0.0f ? -126 : ilogbf(fabsf(golden));
        if (e < -126) e = -126;
        const double lim = 32.0 * ldexp(1.0, e - 23);
        r->max_error_limit = lim > 0.01 ? lim : 0.01;
    }
}

static int comp_passed(const comp_t* r)
{
    return r->nonfinite_failed == 0 && r->max_error <= r->max_error_limit &&
           (r->finite == 0 || r->matched * 100 >= r->finite * 99);
}

int oracle_verify_component(const float* actual, const float* golden, long long count, double* stats)
{
    comp_t r;
    comp_init(&r);
    for (long long t = 0; t < count; ++t) comp_add(&r, actual[t], golden[t]);
    if (stats) {
        stats[0] = (double)r.finite;
        stats[1] = (double)r.matched;
        stats[2] = (double)r.nonfinite;
        stats[3] = (double)r.nonfinite_failed;
        stats[4] = r.max_error;
        stats[5] = r.max_error_limit;
        stats[6] = r.max_error_golden;
    }
    return comp_passed(&r);
}

int oracle_verify(const float* actual, const float* golden, const float* initial, int m, int n, int ldc,
                  double* stats)
{
    comp_t re, im;
    comp_init(&re);
    comp_init(&im);
    int padding_ok = 1;
    for (int col = 0; col < n; ++col)
        for (int row = 0; row < ldc; ++row) {
            const size_t idx = ((size_t)col * ldc + row) * 2;
            if (row < m) {
                comp_add(&re, actual[idx], golden[idx]);
                comp_add(&im, actual[idx + 1], golden[idx + 1]);
            } else if (memcmp(&actual[idx], &initial[idx], 2 * sizeof(float)) != 0) {
                padding_ok = 0;
            }
        }
    if (stats) {
        stats[0] = comp_passed(&re);
        stats[1] = comp_passed(&im);
        stats[2] = padding_ok;
        stats[3] = re.max_error;
        stats[4] = im.max_error;
        stats[5] = (double)re.matched / (re.finite ? (double)re.finite : 1.0);
        stats[6] = (double)im.matched / (im.finite ? (double)im.finite : 1.0);
    }
    return comp_passed(&re) && comp_passed(&im) && padding_ok;
}

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Value of stream `seed` at counter `idx`: 24 random bits u, (float)u * 2^-23 - 1, exactly
 * representable and uniform on [-1, 1) in steps of 2^-23. */
float oracle_uniform(uint64_t seed, uint64_t idx)
{
    const uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ULL + (idx + 1) * 0xD1B54A32D192ED03ULL);
    return (float)(z >> 40) * 0x1p-23f - 1.0f;
}

void oracle_fill(uint64_t seed, uint64_t start, uint64_t count, float* out)
{
    for (uint64_t t = 0; t < count; ++t) out[t] = oracle_uniform(seed, start + t);
}
