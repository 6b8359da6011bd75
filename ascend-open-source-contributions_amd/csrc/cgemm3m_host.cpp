// cgemm3m_host.cpp -- aclblasCgemm3m: validation, planner, workspace carving, launches.
//
// Reference: entry ops-blas/blas/gemm/arch35/cgemm3m_host.cpp:17-23 and the shared
// dispatcher cgemm_host.cpp (validation :589-634, :687-699; quick returns :636-663,
// :701-730; planner :130-285; workspace :287-360, :665-685; launches :362-551).
// The planner here is sized for MI355X (256 CUs, 160 KiB LDS, 288 GB HBM) and is a pure
// function of (shape, scalars, op, CU count, config), so the same call always takes the
// same path and reruns are byte-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "host_internal.h"
#include "launch.h"

namespace c3m {

static const char* kOp = "aclblasCgemm3m";
// Product-kernel variant used when the caller does not pin one. FP32 128x128 tiles: the
// LDS-free four-wave kernel (products kernel at 8192^3 on MI355X, complex TFLOP/s: 186 vs
// 162 for eight waves through LDS); 16-bit formats: eight waves per 128x128 tile (bf16x3 390
// vs 336 for four waves, bf16x6 232 vs 212); 64x64 tiles have one form.
static int default_variant(int format, int bm)
{
    if (format == FMT_F32) return bm == 128 ? 2 : 0;
    return 1;
}

StageTimer::StageTimer(aclblasHandle_t h, int stage) : h_(h), stage_(stage)
{
    if (h_ != nullptr && h_->timing && hipEventCreate(&start_) == hipSuccess) {
        (void)hipEventRecord(start_, h_->stream);
    } else {
        start_ = nullptr;
    }
}

StageTimer::~StageTimer()
{
    if (start_ == nullptr) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) == hipSuccess && hipEventRecord(stop, h_->stream) == hipSuccess) {
        h_->timed.push_back({stage_, start_, stop});
    } else {
        (void)hipEventDestroy(start_);
        if (stop != nullptr) (void)hipEventDestroy(stop);
    }
}

static bool is_op(aclblasOperation_t t) { return t == ACLBLAS_OP_N || t == ACLBLAS_OP_T || t == ACLBLAS_OP_C; }
bool is_zero(const aclblasComplex& z) { return z.real == 0.0f && z.imag == 0.0f; }
static bool is_one(const aclblasComplex& z) { return z.real == 1.0f && z.imag == 0.0f; }

// Same checks, order and status codes as ValidateCgemmParams (cgemm_host.cpp:589-634),
// with the 3m remap of INVALID_ENUM to INVALID_VALUE (:711-714).
aclblasStatus_t validate(aclblasHandle_t h, aclblasOperation_t ta, aclblasOperation_t tb, int m, int n, int k,
                         const aclblasComplex* alpha, int lda, int ldb, const aclblasComplex* beta, int ldc,
                         const char* op)
{
    if (h == nullptr) {
        log_error(op, "no handle was given (create one with aclblasCreate)");
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (!is_op(ta)) {
        log_error(op, "op(A) code %d is not N (111), T (112) or C (113)", (int)ta);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (!is_op(tb)) {
        log_error(op, "op(B) code %d is not N (111), T (112) or C (113)", (int)tb);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (m < 0 || n < 0 || k < 0) {
        log_error(op, "negative dimension in (m, n, k) = (%d, %d, %d)", m, n, k);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const int ldaMin = ta != ACLBLAS_OP_N ? std::max(1, k) : std::max(1, m);
    if (lda < ldaMin) {
        log_error(op, "lda=%d is below the stored rows of A (%d)", lda, ldaMin);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const int ldbMin = tb != ACLBLAS_OP_N ? std::max(1, n) : std::max(1, k);
    if (ldb < ldbMin) {
        log_error(op, "ldb=%d is below the stored rows of B (%d)", ldb, ldbMin);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (ldc < std::max(1, m)) {
        log_error(op, "ldc=%d is below the rows of C (%d)", ldc, std::max(1, m));
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (alpha == nullptr) {
        log_error(op, "no alpha scalar was given");
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (beta == nullptr) {
        log_error(op, "no beta scalar was given");
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// ValidateCgemmPointers (cgemm_host.cpp:687-699), called with k = 0 when alpha == 0.
aclblasStatus_t validate_pointers(int k, const void* A, const void* B, const void* C, const char* op)
{
    if (C == nullptr) {
        log_error(op, "C is a null pointer but the output is not empty");
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (k > 0 && (A == nullptr || B == nullptr)) {
        log_error(op, "k = %d needs both A and B, and one of them is a null pointer", k);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static bool mul_ok(size_t a, size_t b, size_t& out)
{
    if (a != 0 && b > std::numeric_limits<size_t>::max() / a) return false;
    out = a * b;
    return true;
}

static bool add_region(size_t bytes, size_t& total, size_t& off)
{
    if (total > std::numeric_limits<size_t>::max() - kAlign) return false;
    off = round_up(total, kAlign);
    if (bytes > std::numeric_limits<size_t>::max() - off) return false;
    total = off + bytes;
    return true;
}

bool make_plan(const Problem& p, const aclblasCgemm3mConfig* cfg, int cus, Plan& pl, const char* op)
{
    if (cus <= 0) cus = 256;
    pl = Plan();
    pl.precision = cfg ? cfg->precision : ACLBLAS_CGEMM3M_FP32;
    if (p.m == 0 || p.n == 0) {
        pl.path = 0;
        return true;
    }
    if (p.k == 0 || is_zero(p.alpha)) {
        pl.path = is_one(p.beta) ? 0 : 3;
        return true;
    }
    // Padded extents and the kernels' element and chunk counts are 32-bit integers: every
    // dimension must stay below INT_MAX after padding (m, n to 256; k to the deepest stage
    // group).
    constexpr int kMaxDim = std::numeric_limits<int>::max() - 4096;
    if (p.m > kMaxDim || p.n > kMaxDim || p.k > kMaxDim) {
        log_error(op, "m, n and k above %d are not supported", kMaxDim);
        return false;
    }
    switch (pl.precision) {
        case ACLBLAS_CGEMM3M_FP32: pl.format = FMT_F32; pl.pieces = 1; pl.bk = 16; break;
        case ACLBLAS_CGEMM3M_BF16X3: pl.format = FMT_BF16; pl.pieces = 2; pl.bk = 16; break;
        case ACLBLAS_CGEMM3M_BF16X6: pl.format = FMT_BF16; pl.pieces = 3; pl.bk = 16; break;
        case ACLBLAS_CGEMM3M_FP16_SCALED: pl.format = FMT_F16; pl.pieces = 1; pl.bk = 16; break;
        default: log_error(op, "invalid precision %d", pl.precision); return false;
    }
    pl.identity = (cfg && (cfg->identity == ID_STANDARD || cfg->identity == ID_FOUR)) ? cfg->identity : ID_FUSED;
    if (cfg && cfg->identity != 0 && cfg->identity != ID_FUSED && cfg->identity != ID_STANDARD &&
        cfg->identity != ID_FOUR) {
        log_error(op, "invalid identity %d", cfg->identity);
        return false;
    }
    if (pl.identity == ID_FOUR && (pl.format != FMT_F32 || (cfg->variant != 0 && cfg->variant != 3))) {
        log_error(op, "the four-product form runs on FP32 products in the LDS-free kernel only");
        return false;
    }
    const long long m = p.m, n = p.n, k = p.k;
    auto tiles_for = [&](int t) { return ((m + t - 1) / t) * ((n + t - 1) / t); };
    int tile = 0;
    if (cfg && cfg->tile != 0) {
        // tile 16 names the one-launch kernel's single-wave form (variant 5)
        if (cfg->tile != 64 && cfg->tile != 128 && !(cfg->tile == 16 && cfg->variant == 5)) {
            log_error(op, "invalid tile %d", cfg->tile);
            return false;
        }
        tile = cfg->tile == 16 ? 64 : cfg->tile;
    } else {
        // The tuned table holds FP32-product measurements; other formats use the occupancy rule.
        // (An entry naming the one-launch kernel, tile 32 or 16, is used further down.) The
        // 16-bit fused kernels keep 128 x 128 tiles (split in k below) down to a quarter of
        // the CUs' worth of tiles: FP16 1024 x 1024 x 4096 0.168 ms on 64 x 64 tiles against
        // 0.112 ms on 128 x 128 tiles over four k partitions.
        int tt = 0, tp = 0;
        if (pl.format == FMT_F32 && tuned_lookup(m, n, k, false, &tt, &tp) && tt >= 64)
            tile = tt;
        else if (pl.format == FMT_F32)
            tile = tiles_for(128) >= cus ? 128 : 64;
        else
            tile = tiles_for(128) >= cus / 4 ? 128 : 64;
    }
    pl.bm = pl.bn = tile;
    pl.mPad = (int)round_up((size_t)m, 128);
    pl.nPad = (int)round_up((size_t)n, 128);
    pl.kPad = (int)round_up((size_t)k, 32);
    pl.kSteps = pl.kPad / pl.bk;
    const long long tiles = ((m + pl.bm - 1) / pl.bm) * ((n + pl.bn - 1) / pl.bn);
    if (cfg && cfg->kSplit != 0) {
        // up to 128 partitions for the BF16 separate products (variants 4 and 7), 16 elsewhere
        const int maxSplit = (pl.format == FMT_BF16 && (cfg->variant == 4 || cfg->variant == 7)) ? 128 : 16;
        if (cfg->kSplit < 1 || cfg->kSplit > maxSplit) {
            log_error(op, "invalid kSplit %d", cfg->kSplit);
            return false;
        }
        pl.parts = std::min(cfg->kSplit, pl.kSteps);
    } else if (int tt = 0, tp = 0; pl.format == FMT_F32 && (!cfg || cfg->tile == 0) &&
                                   tuned_lookup(m, n, k, false, &tt, &tp) && tt >= 64) {
        pl.parts = std::max(1, std::min(tp, pl.kSteps));
    } else {
        // Too few output tiles to fill the CUs: split k, keeping >= 256 k per partition. FP32
        // (classes the tuned table does not list): the split, at most 16, that minimises
        // rounds x (fixed cost + k / parts), rounds = workgroups / CUs rounded up and the fixed
        // cost of a tile taken as 64 k. A few workgroups more than the CUs cost a whole round;
        // the rule takes a smaller split that fits one round, or a larger one whose extra
        // rounds are full. On 21 seeded random shapes of such classes it ran 1.35x faster
        // (geometric mean) than the smallest split filling every CU (408 x 153 x 33967: 13
        // partitions, 273 workgroups, 400 us; 12, 252, 291 us; 549 x 1122 x 7275: 2, 324, 474
        // us; 3, 486, 356 us), and never slower than the largest split within one round.
        // 16-bit formats on 64 x 64 tiles
        // (two workgroups fit a CU, and the long k loops are latency-bound at one): about two
        // per CU, at most 64 partitions (256 x 256 x 65536 bf16x3: 1.57 ms at 16 partitions).
        // 16-bit formats on 128 x 128 tiles (one workgroup per CU): the largest split that still
        // runs in one round, since a few workgroups more than the CUs make a second, nearly
        // empty round. FP16 on MI355X, ms per call: 1024 x 1024 x 16384 0.434 at 4 partitions
        // (256 workgroups) against 0.447 at 8 (512); 768 x 768 x 8192 0.132 at 7 (252) against
        // 0.159 at 15 (540).
        pl.parts = 1;
        if (tiles < cus) {
            long long parts;
            if (pl.format == FMT_F32) {
                const long long pmax = std::min<long long>(16, std::max(1, pl.kSteps * pl.bk / 256));
                double best = 0.0;
                parts = 1;
                for (long long p = 1; p <= pmax; ++p) {
                    const double rounds = (double)((tiles * p + cus - 1) / cus);
                    const double t = rounds * (64.0 + (double)k / (double)p);
                    if (p == 1 || t < best * 0.999) {
                        best = t;
                        parts = p;
                    }
                }
            } else if (tile == 64)
                parts = std::min<long long>((2 * cus + tiles - 1) / tiles, 64);
            else
                parts = std::min<long long>(cus / tiles, 16);
            parts = std::min<long long>(parts, std::max(1, pl.kSteps * pl.bk / 256));
            pl.parts = (int)std::max<long long>(parts, 1);
        }
    }
    if (tiles * pl.parts > std::numeric_limits<int>::max()) {
        log_error(op, "problem too large for one launch");
        return false;
    }
    pl.path = pl.parts > 1 ? 2 : 1;
    if (cfg && (cfg->variant < 0 || cfg->variant > 7 ||
                (cfg->variant == 7 && pl.format != FMT_BF16) ||
                (cfg->variant == 6 && pl.format == FMT_F32) ||
                (cfg->variant == 5 && (pl.format != FMT_F32 || pl.identity == ID_FOUR)) ||
                (cfg->variant == 3 && (pl.format != FMT_F32 || (pl.bm != 128 && pl.identity != ID_FOUR))) ||
                (cfg->variant == 4 && pl.format == FMT_F32))) {
        log_error(op, "invalid variant %d for this precision and tile", cfg->variant);
        return false;
    }
    pl.variant = (cfg && cfg->variant != 0) ? cfg->variant - 1 : default_variant(pl.format, pl.bm);
    // Separate products (16-bit formats): 256 x 256 tiles of each real product over the whole k
    // range, then merge_kernel combines the three FP32 product planes. Chosen by default when
    // nothing is pinned, the fused plan has no K split and the 3 x (256 x 256 tiles) fill at
    // least half the CUs (8192^3 on MI355X, complex TFLOP/s fused vs separate: bf16x3 395 vs
    // 577, bf16x6 232 vs 319; 2048^3: 291 vs 323, 173 vs 198).
    //
    // BF16 products with too few 256 x 256 tiles for the CUs split k instead (tall-skinny and
    // small outputs): the most partitions that keep the workgroups within one per CU (a
    // partial second round costs a whole round: bf16x3 1024 x 2048 x 4096 0.280 ms at 3
    // partitions, 288 workgroups, against 0.229 ms at 2, 192), each partition at least 256
    // deep, at most 128 partitions; merge_kernel sums each product's partitions in index
    // order first.
    const bool pinned = cfg && (cfg->tile != 0 || cfg->kSplit != 0 || cfg->variant != 0);
    const long long tiles256 = ((m + 255) / 256) * ((n + 255) / 256);
    long long splitParts = 1;
    if (pl.format == FMT_BF16 && !pinned && 3 * tiles256 < cus / 2)
        splitParts = std::max<long long>(1, std::min<long long>({cus / (3 * tiles256), 128, pl.kSteps * pl.bk / 256}));
    //
    // With few outputs (at most 16 tiles of 128 x 128) the BF16 products read the operands as
    // stored instead (variant 7 pins it): no pack launch, the plane values and BF16 pieces
    // formed in registers, the same bits. About one workgroup per CU over the K partitions.
    // Every lane offset of that kernel must fit in 31 bits (72 leading dimensions of 8 bytes).
    const long long tiles128 = ((m + 127) / 128) * ((n + 127) / 128);
    const bool rawFits = (long long)std::max(p.lda, p.ldb) * 8 * 72 < (1LL << 31);
    if (cfg && cfg->variant == 7 && !rawFits) {
        log_error(op, "a leading dimension of %d is too wide for variant 7", std::max(p.lda, p.ldb));
        return false;
    }
    const bool raw = (cfg && cfg->variant == 7) || (splitParts > 1 && rawFits && tiles128 <= 16);
    const bool products = (cfg && (cfg->variant == 4 || cfg->variant == 6)) ||
                          (pl.format != FMT_F32 && !pinned && pl.parts == 1 && 3 * tiles256 >= cus / 2) ||
                          splitParts > 1 || raw;
    if (products) {
        const bool pinnedSplit =
            cfg && (cfg->variant == 4 || cfg->variant == 7) && cfg->kSplit > 1 && pl.format == FMT_BF16;
        if (cfg && ((cfg->kSplit > 1 && !pinnedSplit) || (cfg->tile != 0 && cfg->tile != 128))) {
            log_error(op, "the separate-products form has its own tile; K split only for BF16 products in the "
                          "register form");
            return false;
        }
        pl.path = 4;
        pl.raw = raw;
        pl.bm = pl.bn = raw ? 128 : 256;
        if (raw && !pinnedSplit)
            splitParts = std::max<long long>(1, std::min<long long>({cus / tiles128, 128, pl.kSteps * pl.bk / 256}));
        pl.parts = pinnedSplit ? std::min(cfg->kSplit, pl.kSteps) : (int)splitParts;
        pl.mPad = (int)round_up((size_t)m, raw ? 128 : 256);
        pl.nPad = (int)round_up((size_t)n, raw ? 128 : 256);
        // kernel: the register-prefetch form by default; variant 6 pins the LDS-DMA form (byte-
        // identical, measured slower: bf16x3 at 8192^3 7.52 ms against 6.94 ms), whose unrolled
        // stage loop wants k in whole groups of its unroll depth (16-deep stages)
        pl.variant = raw ? 6 : ((cfg && cfg->variant == 6) ? 5 : 3);
        if (pl.variant == 5) {
            pl.kPad = (int)round_up((size_t)k, 16 * (size_t)lowp_products_lds_unroll(pl.pieces));
            pl.kSteps = pl.kPad / pl.bk;
        }
    }
    if (cfg && (cfg->correction < -1 || cfg->correction > 2)) {
        log_error(op, "invalid correction flag %d", cfg->correction);
        return false;
    }
    // pl.correction: 0 off, 1 FP64 recompute, 2 recompute in the reference BLAS's FP32 order
    if (cfg && cfg->correction != 0)
        pl.correction = cfg->correction > 0 ? cfg->correction : 0;
    else
        pl.correction = k >= 1024;
    // One-launch small-problem kernel: no pack launch. Pinned with variant 5 (kSplit = k
    // slices per tile, tile 16 = one wave per tile instead of four), or chosen when nothing is
    // pinned if the tuned table names it (tile 32 four waves, tile 16 one wave; parts =
    // slices) for this shape class or, for classes it does not list, the output has at most
    // 256 of its 16x16 tiles. Workspace: none, unless the correction is on (one flagged-output
    // counter per tile) or k is sliced (arrival counters, partial products and energies per
    // slice).
    {
        int tt = 0, tp = 0;
        const bool listed = tuned_lookup(m, n, k, false, &tt, &tp);
        const long long tiles16 = ((m + 15) / 16) * ((n + 15) / 16);
        const bool fits = listed ? (tt == 32 || tt == 16) : tiles16 <= 256;
        const bool autoSmall = pl.format == FMT_F32 && !pinned && pl.identity != ID_FOUR && fits;
        if ((cfg && cfg->variant == 5) || autoSmall) {
            const int correction = pl.correction;
            const long long chunks = (k + 15) / 16;
            long long slices = 1;
            if (cfg && cfg->variant == 5 && cfg->kSplit > 0)
                slices = cfg->kSplit;
            else if (listed && tp > 0)
                slices = tp;
            else  // about two workgroups per CU, each slice at least 8 chunks (two per wave)
                slices = std::min<long long>({(2LL * cus + tiles16 - 1) / tiles16, chunks / 8, 16});
            slices = std::max<long long>(1, std::min(slices, chunks));
            pl = Plan{};
            pl.precision = ACLBLAS_CGEMM3M_FP32;
            pl.identity = (cfg && cfg->identity == ID_STANDARD) ? ID_STANDARD : ID_FUSED;
            pl.path = 5;
            pl.bm = pl.bn = 16;
            pl.parts = (int)slices;
            pl.variant = 4;
            pl.correction = correction;
            pl.smallTiles = tiles16;
            // waves per tile: one for the single-wave form (table tile 16, or pinned tile 16)
            const bool single = (cfg && cfg->variant == 5) ? cfg->tile == 16 : (listed && tt == 16);
            pl.smallWaves = single ? 1 : 4;
            size_t total = 0, bytes = 0, n2 = 0;
            if (slices > 1) {
                // whole 8-byte words: the counters are cleared as 64-bit words
                if (!mul_ok((size_t)(tiles16 + 1) / 2, 8, bytes) || !add_region(bytes, total, pl.offArrive))
                    return false;
                if (!mul_ok((size_t)tiles16, (size_t)slices, n2) || !mul_ok(n2, 768 * 4, bytes) ||
                    !add_region(bytes, total, pl.offPartial))
                    return false;
                if (correction && (!mul_ok(n2, 32 * 8, bytes) || !add_region(bytes, total, pl.offEPA))) return false;
            }
            if (correction && (!mul_ok((size_t)tiles16, 4, bytes) || !add_region(bytes, total, pl.offCounter)))
                return false;
            pl.total = total;
            return true;
        }
    }

    // Workspace regions.
    const size_t elem = pl.format == FMT_F32 ? 4 : 2;
    size_t t1, t2;
    if (!mul_ok((size_t)pl.mPad, (size_t)pl.kPad, t1) || !mul_ok(t1, elem, pl.aPlaneBytes)) return false;
    if (!mul_ok((size_t)pl.nPad, (size_t)pl.kPad, t2) || !mul_ok(t2, elem, pl.bPlaneBytes)) return false;
    pl.planesA = pl.planesB = 3 * pl.pieces;
    if (pl.identity == ID_FOUR) pl.planesB = 2;  // (Br, Bi)
    if (pl.raw) pl.aPlaneBytes = pl.bPlaneBytes = 0;  // no planes: the products read A and B as stored
    size_t total = 0, bytes = 0;
    if (!mul_ok(pl.aPlaneBytes, (size_t)pl.planesA, bytes) || !add_region(bytes, total, pl.offA)) return false;
    if (!mul_ok(pl.bPlaneBytes, (size_t)pl.planesB, bytes) || !add_region(bytes, total, pl.offB)) return false;
    if (pl.parts > 1 || pl.path == 4) {
        size_t mn;
        if (!mul_ok((size_t)pl.mPad, (size_t)pl.nPad, mn)) return false;
        // K split: re and im planes per partition; separate products: one plane per product
        pl.partStride = pl.path == 4 ? mn : 2 * mn;
        if (!mul_ok(pl.partStride * 4, (size_t)(pl.path == 4 ? 3 * pl.parts : pl.parts), bytes) ||
            !add_region(bytes, total, pl.offPartial))
            return false;
    }
    pl.kTiles = (pl.kPad + 63) / 64;
    if (pl.correction) {
        const size_t outs = (size_t)m * (size_t)n;
        pl.bitmapWords = (long long)((outs + 63) / 64);
        // energy partials: two halves per (k tile, row), the half-sum order
        if (!add_region(2 * (size_t)pl.kTiles * pl.mPad * 8, total, pl.offEPA)) return false;
        if (!add_region(2 * (size_t)pl.kTiles * pl.nPad * 8, total, pl.offEPB)) return false;
        if (!add_region((size_t)pl.mPad * 8, total, pl.offEA)) return false;
        if (!add_region((size_t)pl.nPad * 8, total, pl.offEB)) return false;
        if (!add_region((size_t)pl.bitmapWords * 8, total, pl.offBitmap)) return false;
        if (!add_region(256, total, pl.offCounter)) return false;
    }
    if (pl.format == FMT_F16) {
        const size_t kb = (size_t)(pl.kPad + 127) / 128;
        if (!add_region(3 * (size_t)(pl.mPad / 128) * kb * 4, total, pl.offScaleA)) return false;
        if (!add_region(3 * (size_t)(pl.nPad / 128) * kb * 4, total, pl.offScaleB)) return false;
        if (!add_region(3 * (size_t)(pl.mPad / 128) * 4, total, pl.offSumA)) return false;
        if (!add_region(3 * (size_t)(pl.nPad / 128) * 4, total, pl.offSumB)) return false;
    }
    pl.total = total;
    return true;
}

static PackArgs pack_args(const Problem& p, const Plan& pl, char* ws, int role)
{
    PackArgs a{};
    const bool isA = role == 0;
    const aclblasOperation_t t = isA ? p.ta : p.tb;
    a.src = isA ? (const void*)p.A : (const void*)p.B;
    a.ld = isA ? p.lda : p.ldb;
    a.mn = isA ? p.m : p.n;
    a.kk = p.k;
    a.mnPad = isA ? pl.mPad : pl.nPad;
    a.kPad = pl.kPad;
    // op(A)(i,l): N reads A[i + l*lda] (i adjacent); op(B)(l,j): T/C reads B[j + l*ldb] (j adjacent).
    a.mnContig = isA ? (t == ACLBLAS_OP_N) : (t != ACLBLAS_OP_N);
    a.conj = t == ACLBLAS_OP_C;
    a.role = role;
    a.identity = pl.identity;
    a.format = pl.format;
    a.pieces = pl.pieces;
    a.nplanes = isA ? pl.planesA / pl.pieces : pl.planesB / pl.pieces;
    a.planes = ws + (isA ? pl.offA : pl.offB);
    a.planeBytes = isA ? pl.aPlaneBytes : pl.bPlaneBytes;
    a.energyPartial = pl.correction ? reinterpret_cast<double*>(ws + (isA ? pl.offEPA : pl.offEPB)) : nullptr;
    a.scaleExp = pl.format == FMT_F16 ? reinterpret_cast<const int*>(ws + (isA ? pl.offScaleA : pl.offScaleB))
                                      : nullptr;
    return a;
}

aclblasStatus_t enqueue(aclblasHandle_t h, const Problem& p, const Plan& pl, char* ws, const char* op, bool noAlloc)
{
    hipStream_t s = h->stream;
    // With noAlloc the call creates nothing: no counter pool, no timing events.
    aclblasHandle_t th = noAlloc ? nullptr : h;
    hipError_t e = hipSuccess;
    h->lastCounter = nullptr;
    if (pl.path == 0) return ACLBLAS_STATUS_SUCCESS;
    if (pl.path == 5) {
        const double amag =
            std::sqrt((double)p.alpha.real * p.alpha.real + (double)p.alpha.imag * p.alpha.imag);
        unsigned int* slots = pl.correction ? reinterpret_cast<unsigned int*>(ws + pl.offCounter) : nullptr;
        const bool sliced = pl.parts > 1;
        // Arrival counters must be zero at launch; the kernel leaves them zero. The handle's
        // pool is cleared once, when created on an eager call; otherwise (more tiles than the
        // pool holds, or a capture before the pool exists) the workspace's counters are
        // cleared by a fill launch first.
        unsigned int* arrive = nullptr;
        bool clear = false;
        if (sliced) {
            if (!noAlloc && pl.smallTiles <= kArrivePool && h->arrivePool == nullptr && !stream_capturing(s)) {
                void* pool = nullptr;
                if (hipMalloc(&pool, (size_t)kArrivePool * sizeof(unsigned int)) == hipSuccess) {
                    e = launch_fill_u64(static_cast<unsigned long long*>(pool), kArrivePool / 2, s);
                    if (e == hipSuccess)
                        h->arrivePool = static_cast<unsigned int*>(pool);
                    else
                        (void)hipFree(pool);
                } else {
                    (void)hipGetLastError();  // fall back to the workspace's counters
                }
            }
            if (!noAlloc && pl.smallTiles <= kArrivePool && h->arrivePool != nullptr) {
                arrive = h->arrivePool;
            } else {
                arrive = reinterpret_cast<unsigned int*>(ws + pl.offArrive);
                clear = true;
            }
        }
        SmallArgs a{p.A, p.lda, (int)p.ta, p.B, p.ldb, (int)p.tb, p.m, p.n, p.k, pl.identity,
                    p.alpha.real, p.alpha.imag, p.beta.real, p.beta.imag, is_zero(p.beta) ? 1 : 0, p.C, p.ldc,
                    pl.correction, amag * 0x1p-24, 0x1p-8, slots, pl.parts, pl.smallWaves,
                    sliced ? reinterpret_cast<float*>(ws + pl.offPartial) : nullptr,
                    sliced && pl.correction ? reinterpret_cast<double*>(ws + pl.offEPA) : nullptr, arrive};
        StageTimer t(th, ST_GEMM);
        if (e == hipSuccess && clear)
            e = launch_fill_u64(reinterpret_cast<unsigned long long*>(arrive), (pl.smallTiles + 1) / 2, s);
        if (e == hipSuccess) e = launch_gemm_small(a, s);
        if (slots != nullptr) {
            h->lastCounter = slots;
            h->lastCounterSlots = pl.smallTiles;
        }
    } else if (pl.path == 3) {
        ScaleCArgs a{p.m, p.n, p.beta.real, p.beta.imag, is_zero(p.beta) ? 1 : 0, p.C, p.ldc};
        StageTimer t(th, ST_SCALE);
        e = launch_scale_c(a, s);
    } else {
        const PackArgs pa = pack_args(p, pl, ws, 0);
        const PackArgs pb = pack_args(p, pl, ws, 1);
        if (pl.format == FMT_F16 && !pl.raw) {
            StageTimer t(th, ST_PACK);
            auto scale_args = [](const PackArgs& x) {
                return ScaleArgs{x.src, x.ld, x.mn, x.kk, x.mnPad, x.kPad, x.mnContig, x.conj, x.role,
                                 x.identity, const_cast<int*>(x.scaleExp)};
            };
            if (e == hipSuccess) e = launch_scale_exponents(scale_args(pa), scale_args(pb), s);
            ScaleSummaryArgs sm{const_cast<int*>(pa.scaleExp), 3 * (pl.mPad / 128), const_cast<int*>(pb.scaleExp),
                                3 * (pl.nPad / 128),
                                (pl.kPad + 127) / 128, reinterpret_cast<int*>(ws + pl.offSumA),
                                reinterpret_cast<int*>(ws + pl.offSumB)};
            if (e == hipSuccess) e = launch_scale_summary(sm, s);
        }
        if (e == hipSuccess && !pl.raw) {
            StageTimer t(th, ST_PACK);
            e = launch_pack_pair(pa, pb, s);
        } else if (e == hipSuccess && pl.correction) {  // products read as stored: energies only
            StageTimer t(th, ST_PACK);
            e = launch_energy_partial_pair(pa, pb, s);
        }
        double* eA = nullptr;
        double* eB = nullptr;
        unsigned long long* bitmap = nullptr;
        unsigned int* counter = nullptr;
        if (pl.correction) {
            eA = reinterpret_cast<double*>(ws + pl.offEA);
            eB = reinterpret_cast<double*>(ws + pl.offEB);
            bitmap = reinterpret_cast<unsigned long long*>(ws + pl.offBitmap);
            counter = reinterpret_cast<unsigned int*>(ws + pl.offCounter);
            EnergyArgs ea{pa.energyPartial, pl.mPad, pb.energyPartial, pl.nPad, pl.kTiles, eA, eB,
                          bitmap, pl.bitmapWords, counter};
            if (e == hipSuccess) {
                StageTimer t(th, ST_ENERGY);
                e = launch_energy(ea, s);
            }
        }
        // Correction screen: recompute outputs whose condition |alpha| |op(A)_i|_2 |op(B)_j|_2 / |c|
        // exceeds 2^8 / unit (or that are not finite). unit = 2^-24 (FP32 rounding) for FP32,
        // bf16x6 and FP16; bf16x3 drops the lo*lo term (relative size up to 2^-18 per term), so
        // its screen is four times wider (2^-22), which keeps its worst per-element relative
        // error near the FP32 path's. FP16 block-scaled products bound the normwise error only.
        const double amag = std::sqrt((double)p.alpha.real * p.alpha.real + (double)p.alpha.imag * p.alpha.imag);
        const double unit = pl.precision == ACLBLAS_CGEMM3M_BF16X3 ? 0x1p-22 : 0x1p-24;
        GemmArgs g{};
        g.Ap = ws + pl.offA;
        g.aPlaneBytes = pl.aPlaneBytes;
        g.mPad = pl.mPad;
        g.Bp = ws + pl.offB;
        g.bPlaneBytes = pl.bPlaneBytes;
        g.nPad = pl.nPad;
        g.m = p.m;
        g.n = p.n;
        g.kSteps = pl.kSteps;
        g.parts = pl.parts;
        g.tilesM = (p.m + pl.bm - 1) / pl.bm;
        g.tilesN = (p.n + pl.bn - 1) / pl.bn;
        g.identity = pl.identity;
        g.alphaR = p.alpha.real;
        g.alphaI = p.alpha.imag;
        g.betaR = p.beta.real;
        g.betaI = p.beta.imag;
        g.betaZero = is_zero(p.beta) ? 1 : 0;
        g.C = p.C;
        g.ldc = p.ldc;
        g.partial = (pl.parts > 1 || pl.path == 4) ? reinterpret_cast<float*>(ws + pl.offPartial) : nullptr;
        if (pl.path == 4) {
            g.tilesM = pl.mPad / pl.bm;
            g.tilesN = pl.nPad / pl.bn;
        }
        g.partStride = pl.partStride;
        g.scaleA = pa.scaleExp;
        g.scaleB = pb.scaleExp;
        g.kBlocks128 = (pl.kPad + 127) / 128;
        if (pl.format == FMT_F16) {
            g.sumA = reinterpret_cast<const int*>(ws + pl.offSumA);
            g.sumB = reinterpret_cast<const int*>(ws + pl.offSumB);
        }
        g.correction = pl.correction;
        g.eA = eA;
        g.eB = eB;
        g.corrScale = amag * unit;
        g.corrTau = 0x1p-8;
        g.bitmap = bitmap;
        g.src = SourceOperands{p.A, p.lda, (int)p.ta, p.B, p.ldb, (int)p.tb, p.k};
        if (e == hipSuccess) {
            StageTimer t(th, ST_GEMM);
            if (pl.path == 4 && pl.raw)
                e = launch_gemm_bf16_raw_products(g, pl.pieces, s);
            else if (pl.path == 4)
                e = launch_gemm_lowp_products(g, pl.pieces, pl.variant == 5, s);
            else if (pl.format == FMT_F32)
                e = launch_gemm_f32(g, pl.bm, pl.bn, pl.variant, s);
            else if (pl.format == FMT_BF16)
                e = launch_gemm_bf16(g, pl.bm, pl.bn, pl.pieces, pl.variant, s);
            else
                e = launch_gemm_f16(g, pl.bm, pl.bn, pl.variant, s);
        }
        if ((pl.parts > 1 || pl.path == 4) && e == hipSuccess) {
            MergeArgs mg{};
            mg.partial = g.partial;
            mg.partStride = pl.partStride;
            mg.parts = pl.path == 4 ? 3 : pl.parts;
            mg.products = pl.path == 4 ? 1 : 0;
            mg.splits = pl.path == 4 ? pl.parts : 1;
            mg.identity = pl.identity;
            mg.ldp = pl.mPad;
            mg.m = p.m;
            mg.n = p.n;
            mg.alphaR = g.alphaR;
            mg.alphaI = g.alphaI;
            mg.betaR = g.betaR;
            mg.betaI = g.betaI;
            mg.betaZero = g.betaZero;
            mg.C = p.C;
            mg.ldc = p.ldc;
            mg.correction = pl.correction;
            mg.eA = eA;
            mg.eB = eB;
            mg.corrScale = g.corrScale;
            mg.corrTau = g.corrTau;
            mg.bitmap = bitmap;
            mg.src = g.src;
            StageTimer t(th, ST_MERGE);
            e = launch_merge(mg, s);
        }
        if (pl.correction && e == hipSuccess) {
            RecomputeArgs r{};
            r.A = p.A;
            r.lda = p.lda;
            r.opA = (int)p.ta;
            r.B = p.B;
            r.ldb = p.ldb;
            r.opB = (int)p.tb;
            r.m = p.m;
            r.n = p.n;
            r.k = p.k;
            r.alphaR = p.alpha.real;
            r.alphaI = p.alpha.imag;
            r.betaR = p.beta.real;
            r.betaI = p.beta.imag;
            r.betaZero = g.betaZero;
            r.bitmap = bitmap;
            r.words = pl.bitmapWords;
            r.C = p.C;
            r.ldc = p.ldc;
            r.counter = counter;
            r.netlib = pl.correction == 2 ? 1 : 0;
            r.alphaRf = p.alpha.real;
            r.alphaIf = p.alpha.imag;
            r.betaRf = p.beta.real;
            r.betaIf = p.beta.imag;
            StageTimer t(th, ST_RECOMPUTE);
            e = launch_recompute(r, s);
            h->lastCounter = counter;
            h->lastCounterSlots = 1;
        }
    }
    if (e != hipSuccess) {
        log_error(op, "kernel launch failed: %s", hipGetErrorString(e));
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    if (h->timing) h->timedCalls++;
    return ACLBLAS_STATUS_SUCCESS;
}

static void record_plan(aclblasHandle_t h, const Plan& pl)
{
    h->lastPlan[0] = pl.path;
    h->lastPlan[1] = pl.bm;
    h->lastPlan[2] = pl.parts;
    h->lastPlan[3] = pl.correction;
    h->lastPlan[4] = pl.precision;
    h->lastPlan[5] = pl.identity;
    h->lastPlan[6] = -1;
    h->lastPlan[7] = (long long)pl.total;
}

aclblasStatus_t run_problem(aclblasHandle_t handle, aclblasOperation_t ta, aclblasOperation_t tb, int m, int n,
                            int k, const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                            const aclblasComplex* B, int ldb, const aclblasComplex* beta, aclblasComplex* C,
                            int ldc, const aclblasCgemm3mConfig* cfg, const char* op, ScratchSource scratch)
{
    aclblasStatus_t st = validate(handle, ta, tb, m, n, k, alpha, lda, ldb, beta, ldc, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    if (m == 0 || n == 0) {
        Plan none;
        record_plan(handle, none);
        return ACLBLAS_STATUS_SUCCESS;
    }
    const Problem p{ta, tb, m, n, k, *alpha, *beta, A, lda, B, ldb, C, ldc};
    st = validate_pointers(is_zero(p.alpha) ? 0 : k, A, B, C, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    Plan pl;
    if (!make_plan(p, cfg, handle->cus, pl, op)) return ACLBLAS_STATUS_INVALID_VALUE;
    record_plan(handle, pl);
    if (pl.path == 0) return ACLBLAS_STATUS_SUCCESS;
    st = ensure_device(handle);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        log_error(op, "no HIP device available");
        return st;
    }
    // The plan depends on the CU count; re-plan now that it is known.
    if (!make_plan(p, cfg, handle->cus, pl, op)) return ACLBLAS_STATUS_INVALID_VALUE;
    record_plan(handle, pl);
    char* ws = nullptr;
    if (pl.total > 0 && scratch != nullptr) {
        // the caller's fixed device array instead of the handle's workspace: nothing allocated
        size_t cap = 0;
        st = scratch(&ws, &cap, op);
        if (st != ACLBLAS_STATUS_SUCCESS) return st;
        if (pl.total > cap) {
            log_error(op, "this call needs %zu bytes of temporary storage and the fixed array holds %zu", pl.total,
                      cap);
            return ACLBLAS_STATUS_NOT_SUPPORTED;
        }
    } else if (pl.total > 0) {
        st = acquire_workspace(handle, pl.total, &ws, op);
        if (st != ACLBLAS_STATUS_SUCCESS) return st;
    }
    return enqueue(handle, p, pl, ws, op, scratch != nullptr);
}

}  // namespace c3m

using namespace c3m;

extern "C" aclblasStatus_t aclblasCgemm3m(aclblasHandle_t handle, aclblasOperation_t transa,
                                          aclblasOperation_t transb, int m, int n, int k,
                                          const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                                          const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                          aclblasComplex* C, int ldc)
{
    return run_problem(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, kOp);
}

// C <- alpha*op(A)*op(B) + beta*C with four real products (re = Ar*Br - Ai*Bi,
// im = Ar*Bi + Ai*Br), accumulated directly into (re, im) by the LDS-free FP32 kernel.
// Reference: ops-blas/blas/gemm/arch35/cgemm_host.cpp:732-738 over LaunchComplexGemm
// (:701-730). Differences from aclblasCgemm3m kept from there: a bad op returns
// INVALID_ENUM (no remap, :713), and A and B must be non-null whenever k > 0, alpha == 0
// included (skipProduct applies to the 3m entry only, :721-722).
extern "C" aclblasStatus_t aclblasCgemm(aclblasHandle_t handle, aclblasOperation_t transa,
                                        aclblasOperation_t transb, int m, int n, int k, const aclblasComplex* alpha,
                                        const aclblasComplex* A, int lda, const aclblasComplex* B, int ldb,
                                        const aclblasComplex* beta, aclblasComplex* C, int ldc)
{
    const char* op = "aclblasCgemm";
    if (handle == nullptr) {
        log_error(op, "no handle was given (create one with aclblasCreate)");
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    const auto valid = [](aclblasOperation_t t) {
        return t == ACLBLAS_OP_N || t == ACLBLAS_OP_T || t == ACLBLAS_OP_C;
    };
    if (!valid(transa) || !valid(transb)) {
        log_error(op, "invalid transa/transb: %d/%d", (int)transa, (int)transb);
        return ACLBLAS_STATUS_INVALID_ENUM;
    }
    aclblasStatus_t st = validate(handle, transa, transb, m, n, k, alpha, lda, ldb, beta, ldc, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    if (m == 0 || n == 0) {
        Plan none;
        record_plan(handle, none);
        return ACLBLAS_STATUS_SUCCESS;
    }
    st = validate_pointers(k, A, B, C, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    const aclblasCgemm3mConfig cfg{ACLBLAS_CGEMM3M_FP32, 0, 0, 0, ID_FOUR, 0};
    return run_problem(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &cfg, op);
}

extern "C" aclblasStatus_t aclblasCgemm3mEx(aclblasHandle_t handle, aclblasOperation_t transa,
                                            aclblasOperation_t transb, int m, int n, int k,
                                            const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                                            const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                            aclblasComplex* C, int ldc, const aclblasCgemm3mConfig* config)
{
    return run_problem(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, config,
                       "aclblasCgemm3mEx");
}

extern "C" aclblasStatus_t aclblasCgemm3mGetWorkspaceSize(aclblasHandle_t handle, aclblasOperation_t transa,
                                                          aclblasOperation_t transb, int m, int n, int k,
                                                          const aclblasComplex* alpha, const aclblasComplex* beta,
                                                          const aclblasCgemm3mConfig* config, size_t* bytes)
{
    const char* op = "aclblasCgemm3mGetWorkspaceSize";
    if (bytes == nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    const int lda = std::max(1, transa != ACLBLAS_OP_N ? k : m);
    const int ldb = std::max(1, transb != ACLBLAS_OP_N ? n : k);
    aclblasStatus_t st = validate(handle, transa, transb, m, n, k, alpha, lda, ldb, beta, std::max(1, m), op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    const Problem p{transa, transb, m, n, k, *alpha, *beta, nullptr, lda, nullptr, ldb, nullptr, std::max(1, m)};
    if (ensure_device(handle) != ACLBLAS_STATUS_SUCCESS) handle->cus = 0;
    Plan pl;
    if (!make_plan(p, config, handle->cus, pl, op)) return ACLBLAS_STATUS_INVALID_VALUE;
    *bytes = pl.total;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasCgemm3mGetLastPlan(aclblasHandle_t handle, int sync, long long* info)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (info == nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    for (int t = 0; t < 8; ++t) info[t] = handle->lastPlan[t];
    info[6] = -1;
    if (sync && handle->lastCounter != nullptr && handle->lastCounterSlots > 0) {
        std::vector<unsigned int> v((size_t)handle->lastCounterSlots, 0u);
        if (hipStreamSynchronize(handle->stream) != hipSuccess ||
            hipMemcpy(v.data(), handle->lastCounter, v.size() * sizeof(unsigned int), hipMemcpyDeviceToHost) !=
                hipSuccess) {
            return ACLBLAS_STATUS_EXECUTION_FAILED;
        }
        long long total = 0;
        for (unsigned int x : v) total += x;
        info[6] = total;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasCgemm3mSetTiming(aclblasHandle_t handle, int enable)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    handle->timing = enable ? 1 : 0;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasCgemm3mGetTiming(aclblasHandle_t handle, double* ms, long long* calls)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (ms == nullptr || calls == nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    for (int t = 0; t < ST_COUNT; ++t) ms[t] = 0.0;
    aclblasStatus_t st = ACLBLAS_STATUS_SUCCESS;
    for (auto& ev : handle->timed) {
        float v = 0.0f;
        if (hipEventSynchronize(ev.stop) == hipSuccess && hipEventElapsedTime(&v, ev.start, ev.stop) == hipSuccess)
            ms[ev.stage] += v;
        else
            st = ACLBLAS_STATUS_EXECUTION_FAILED;
        (void)hipEventDestroy(ev.start);
        (void)hipEventDestroy(ev.stop);
    }
    handle->timed.clear();
    *calls = handle->timedCalls;
    handle->timedCalls = 0;
    return st;
}

extern "C" aclblasStatus_t aclblasCgemm3mResidual(aclblasHandle_t handle, aclblasOperation_t transa,
                                                  aclblasOperation_t transb, int m, int n, int k,
                                                  const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                                                  const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                                  const aclblasComplex* Cinit, const aclblasComplex* C, int ldc,
                                                  double* out)
{
    const char* op = "aclblasCgemm3mResidual";
    aclblasStatus_t st = validate(handle, transa, transb, m, n, k, alpha, lda, ldb, beta, ldc, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    if (out == nullptr || C == nullptr || (k > 0 && (A == nullptr || B == nullptr)) ||
        (!is_zero(*beta) && Cinit == nullptr))
        return ACLBLAS_STATUS_INVALID_VALUE;
    for (int t = 0; t < 4; ++t) out[t] = 0.0;
    if (m == 0 || n == 0) return ACLBLAS_STATUS_SUCCESS;
    st = ensure_device(handle);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    const int bm = (m + 63) / 64, bn = (n + 63) / 64;
    const size_t nblk = (size_t)bm * bn;
    double* dOut = nullptr;
    if (hipMalloc(&dOut, nblk * 5 * sizeof(double)) != hipSuccess) return ACLBLAS_STATUS_ALLOC_FAILED;
    ResidualArgs r{};
    r.A = A;
    r.lda = lda;
    r.opA = (int)transa;
    r.B = B;
    r.ldb = ldb;
    r.opB = (int)transb;
    r.m = m;
    r.n = n;
    r.k = k;
    r.alphaR = alpha->real;
    r.alphaI = alpha->imag;
    r.betaR = beta->real;
    r.betaI = beta->imag;
    r.betaZero = is_zero(*beta) ? 1 : 0;
    r.Cinit = Cinit;
    r.C = C;
    r.ldc = ldc;
    r.out = dOut;
    hipError_t e = launch_residual(r, bm, bn, handle->stream);
    std::vector<double> host(nblk * 5);
    if (e == hipSuccess) e = hipStreamSynchronize(handle->stream);
    if (e == hipSuccess) e = hipMemcpy(host.data(), dOut, host.size() * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dOut);
    if (e != hipSuccess) {
        log_error(op, "residual failed: %s", hipGetErrorString(e));
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    double mx = 0.0, sum = 0.0, cnt = 0.0, mxAbs = 0.0, mxRef = 0.0;
    for (size_t b = 0; b < nblk; ++b) {
        mx = std::max(mx, host[b * 5 + 0]);
        sum += host[b * 5 + 1];
        cnt += host[b * 5 + 2];
        mxAbs = std::max(mxAbs, host[b * 5 + 3]);
        mxRef = std::max(mxRef, host[b * 5 + 4]);
    }
    out[0] = mx;
    out[1] = cnt > 0 ? sum / cnt : 0.0;
    out[2] = mxRef > 0 ? mxAbs / mxRef : mxAbs;
    out[3] = mxRef;
    return ACLBLAS_STATUS_SUCCESS;
}
