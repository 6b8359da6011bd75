// grouped_host.cpp -- aclblasCgemm3mGroupedBatched: many independent complex GEMMs in one
// persistent product launch (operands read as stored, 64 x 64 wave blocks pulled from one
// queue), with the one-launch small-problem kernel beside it for members the table gives to
// it, and a merge and a recompute launch when some member needs them. No reference
// counterpart (the reference's batched entry is the uniform strided one,
// ops-blas/blas/gemm/arch35/cgemm_host.cpp:740-782).
//
// Plans per problem come from the autotuned table (tuned_table.h) with grouped = 1, else a
// fixed rule: 128 tiles for problems at least 256 x 256, 64 tiles otherwise, no K split (the
// batch itself fills the CUs). The descriptor block (problem arguments, pack items, tile
// list, merge items) is built on the host and kept on the device per content, so a repeated
// call (or a hipGraph replay of one) launches kernels only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "host_internal.h"
#include "launch.h"
#include "tuned_table.h"

namespace c3m {

static const TunedEntry* tuned_bin(int mnl, int kl, bool grouped)
{
    for (const TunedEntry& e : kTunedTable)
        if (e.tile != 0 && e.mnLog2 == mnl && e.kLog2 == kl && e.grouped == (grouped ? 1 : 0)) return &e;
    return nullptr;
}

bool tuned_lookup(long long m, long long n, long long k, bool grouped, int* tile, int* parts)
{
    if (m <= 0 || n <= 0 || k <= 0) return false;
    int mnl = 0, kl = 0;
    for (unsigned long long v = (unsigned long long)(m * n); v > 1; v >>= 1) ++mnl;
    for (unsigned long long v = (unsigned long long)k; v > 1; v >>= 1) ++kl;
    if (const TunedEntry* e = tuned_bin(mnl, kl, grouped)) {
        *tile = e->tile;
        *parts = e->parts;
        return true;
    }
    // A one-problem class the tuner did not sample, between two k classes of the same output
    // size that both take the one-launch kernel without k slices, takes it too (single-wave
    // form when both do, else four waves). On 47 seeded random shapes in the 11 such classes
    // (outputs up to 2^21, k to 2^8) it ran 1.35x faster (geometric mean of kernel-time sums)
    // than the occupancy rule's packed plan, faster on 39, at most 8% slower on the others.
    if (!grouped) {
        const TunedEntry* lo = tuned_bin(mnl, kl - 1, false);
        const TunedEntry* hi = tuned_bin(mnl, kl + 1, false);
        const auto oneLaunch = [](const TunedEntry* e) { return e != nullptr && (e->tile == 16 || e->tile == 32) && e->parts == 1; };
        if (oneLaunch(lo) && oneLaunch(hi)) {
            *tile = (lo->tile == 16 && hi->tile == 16) ? 16 : 32;
            *parts = 1;
            return true;
        }
    }
    return false;
}

// Content key of a descriptor block, 8 bytes at a time: each word is folded in with a
// multiply and a shift (the key only selects a cached block; memcmp confirms the match).
static uint64_t block_hash(const void* data, size_t bytes)
{
    const unsigned char* p = static_cast<const unsigned char*>(data);
    uint64_t h = 0x9e3779b97f4a7c15ull ^ bytes;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 0xff51afd7ed558ccdull;
        h ^= h >> 32;
    }
    for (; i < bytes; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}

// Buffer size of a descriptor slot: a power of two of at least 4 KiB, so that a reused slot
// rarely needs larger buffers.
static size_t slot_capacity(size_t bytes)
{
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

static bool slot_buffers(_aclblas_handle::Staged& st, size_t cap)
{
    RelaxedCapture relaxed;
    st.host = st.device = nullptr;
    st.capacity = 0;
    if (hipHostMalloc(&st.host, cap, hipHostMallocDefault) != hipSuccess) {
        st.host = nullptr;
        return false;
    }
    if (hipMalloc(&st.device, cap) != hipSuccess) {
        (void)hipHostFree(st.host);
        st.host = nullptr;
        st.device = nullptr;
        return false;
    }
    st.capacity = cap;
    return true;
}

// Device copy of a descriptor block. The first call with a given content copies it from the
// slot's pinned host buffer on the stream; later calls with the same content reuse the device
// copy without any copy, so a hipGraph captured after one eager call holds kernels only. If
// the first call is itself being captured, the copy becomes a graph node and a later eager
// call uploads again. A slot used while capturing is marked and never reused. A new content
// takes a new slot until kMaxStaged slots exist, then the least recently used one, after the
// host has waited for the event recorded after that slot's last kernels (one earlier call of
// this handle, long finished in a steady loop): no device-wide synchronisation and no
// allocation per call. *slot receives the slot index (for grouped_run to record its event).
static const char* stage_descriptors(aclblasHandle_t h, const std::vector<char>& blob, hipStream_t s, size_t* slot,
                                     hipError_t* err)
{
    const bool capturing = stream_capturing(s);
    const uint64_t key = block_hash(blob.data(), blob.size());
    const size_t none = h->staged.size();
    size_t idx = none;
    for (size_t i = 0; i < h->staged.size(); ++i) {
        const auto& st = h->staged[i];
        if (st.hash == key && st.bytes == blob.size() && std::memcmp(st.host, blob.data(), blob.size()) == 0) idx = i;
    }
    if (idx == none) {
        size_t loose = 0, victim = none;
        for (size_t i = 0; i < h->staged.size(); ++i) {
            if (h->staged[i].captured) continue;
            ++loose;
            if (victim == none || h->staged[i].lastUse < h->staged[victim].lastUse) victim = i;
        }
        if (!capturing && loose >= kMaxStaged) {
            _aclblas_handle::Staged& v = h->staged[victim];
            if (v.recorded) {
                *err = hipEventSynchronize(v.done);
                if (*err != hipSuccess) return nullptr;
            }
            if (v.capacity < blob.size()) {  // its buffers are idle now; replace them with larger ones
                (void)hipHostFree(v.host);
                (void)hipFree(v.device);
                if (!slot_buffers(v, slot_capacity(blob.size()))) {
                    *err = hipErrorOutOfMemory;
                    (void)hipEventDestroy(v.done);
                    h->staged.erase(h->staged.begin() + (std::ptrdiff_t)victim);
                    return nullptr;
                }
            }
            idx = victim;
        } else {
            _aclblas_handle::Staged st{};
            if (!slot_buffers(st, slot_capacity(blob.size()))) {
                *err = hipErrorOutOfMemory;
                return nullptr;
            }
            RelaxedCapture relaxed;
            if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess) {
                *err = hipGetLastError();
                (void)hipHostFree(st.host);
                (void)hipFree(st.device);
                return nullptr;
            }
            h->staged.push_back(st);
            idx = h->staged.size() - 1;
        }
        _aclblas_handle::Staged& fresh = h->staged[idx];
        fresh.hash = key;
        fresh.bytes = blob.size();
        fresh.uploaded = false;
        fresh.recorded = false;
        fresh.doneStream = nullptr;
        std::memcpy(fresh.host, blob.data(), blob.size());
    }
    _aclblas_handle::Staged& hit = h->staged[idx];
    hit.lastUse = ++h->stageClock;
    if (capturing) hit.captured = true;
    // A slot last read on another stream: that read (and the upload before it) finishes
    // before this stream goes on. Not while capturing: a capture may not wait on an event
    // recorded outside it, and the slot, now marked captured, is never rewritten.
    if (hit.recorded && hit.doneStream != s && !capturing) {
        *err = hipStreamWaitEvent(s, hit.done, 0);
        if (*err != hipSuccess) return nullptr;
    }
    if (!hit.uploaded) {
        *err = hipMemcpyAsync(hit.device, hit.host, hit.bytes, hipMemcpyHostToDevice, s);
        if (*err != hipSuccess) return nullptr;
        if (!capturing) hit.uploaded = true;
    }
    *slot = idx;
    return static_cast<const char*>(hit.device);
}

struct GroupedLayout {
    std::vector<Problem> probs;     // GEMM problems (alpha != 0, k > 0, m, n > 0)
    std::vector<Plan> plans;
    std::vector<size_t> base;       // workspace offset of each problem's region
    std::vector<Problem> scaleOnly; // alpha == 0 or k == 0 with beta != 1
    std::vector<Problem> small;     // members on the one-launch kernel (no pack, no workspace)
    std::vector<Plan> smallPlans;
    size_t descBytes = 0, offQueue = 0, total = 0;
    size_t nProb = 0, nPack = 0, nQuad = 0, nMerge = 0;
    size_t nCorr = 0;               // quad-kernel members with the rounding correction on
    size_t nCorrSplit = 0;          // those of them with a K split (energies before the merge)
    size_t energyBytes = 0;         // descriptor bytes of their energy items
    long long packBlocks = 0;
    int maxMergeOutputs = 0;
    bool raw = true;                // the quad kernel reads the members' operands as stored
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// How the quad kernel covers one member's output with 64 x 64 blocks: bm x bn blocks; fm (fn)
// 32-row (32-column) fragments in the last block row (column), 1 when at most 32 outputs are
// left there; whole 128 x 128 regions of full blocks (the first bmQ x bnQ blocks) go to quads
// of one region each, the other blocks to quads of four blocks of like cost.
struct QuadCover {
    int bm, bn, fm, fn, bmQ, bnQ;
};
static QuadCover quad_cover(int m, int n)
{
    QuadCover q;
    q.bm = (m + 63) / 64;
    q.bn = (n + 63) / 64;
    q.fm = m - 64 * (q.bm - 1) > 32 ? 2 : 1;
    q.fn = n - 64 * (q.bn - 1) > 32 ? 2 : 1;
    q.bmQ = (q.fm == 2 ? q.bm : q.bm - 1) & ~1;
    q.bnQ = (q.fn == 2 ? q.bn : q.bn - 1) & ~1;
    return q;
}

// Workspace of a member the quad kernel reads raw: no planes, only the K-split partial planes
// and, with the correction on, the bitmap and counter; a K-split member also keeps its energy
// partials (from the energy-only pass, two halves per k tile) and energies, which the merge
// reads. A fused member's energies are formed by the quad kernel itself, in the same order.
static void raw_member_workspace(Plan& pl)
{
    size_t total = 0;
    auto region = [&](size_t bytes) {
        const size_t off = align256(total);
        total = off + bytes;
        return off;
    };
    pl.offA = pl.offB = pl.offEPA = pl.offEPB = 0;
    pl.aPlaneBytes = pl.bPlaneBytes = 0;  // no planes: pack items of this member write none
    if (pl.parts > 1) pl.offPartial = region(pl.partStride * 4 * (size_t)pl.parts);
    if (pl.correction) {
        if (pl.parts > 1) {
            pl.offEPA = region(2 * (size_t)pl.kTiles * pl.mPad * 8);
            pl.offEPB = region(2 * (size_t)pl.kTiles * pl.nPad * 8);
            pl.offEA = region((size_t)pl.mPad * 8);
            pl.offEB = region((size_t)pl.nPad * 8);
        }
        pl.offBitmap = region((size_t)pl.bitmapWords * 8);
        pl.offCounter = region(256);
    }
    pl.total = total;
}

// Plan of one grouped member: the tuned table's plan for grouped members (tile 64 / 128 with
// K partitions, or the one-launch kernel, tile 32 / 16, unsliced), else 128 tiles for members
// of at least 256 x 256 and 64 tiles below, no K split; the rounding correction of a single
// aclblasCgemm3m call (on when k >= 1024); the caller's
// tile/kSplit/identity/correction override when given (variant 5: the one-launch kernel). The one-problem
// form of the same plan (variant: LDS-free kernel for 128 tiles, as the grouped kernel uses)
// gives byte-identical results, which the strided-batched entry relies on.
bool grouped_batch_fills(long long blocks64, int cus) { return cus > 0 && blocks64 >= 8LL * cus; }

aclblasCgemm3mConfig grouped_member_config(int m, int n, int k, const aclblasCgemm3mConfig* over, bool batchFills)
{
    aclblasCgemm3mConfig cfg{ACLBLAS_CGEMM3M_FP32, 0, 1, 0, 0, 1};
    int tt = 0, tp = 0;
    bool small = false;
    if (over != nullptr && over->variant == 5) {
        small = true;
        tt = over->tile == 16 ? 16 : 32;
    } else if (over != nullptr && over->tile != 0) {
        cfg.tile = over->tile;
        cfg.kSplit = over->kSplit != 0 ? over->kSplit : 1;
    } else if (tuned_lookup(m, n, k, true, &tt, &tp)) {
        small = tt == 32 || tt == 16;
        cfg.tile = small ? 0 : tt;
        cfg.kSplit = small ? 1 : tp;
    } else {
        cfg.tile = (m >= 256 && n >= 256) ? 128 : 64;
    }
    if (over != nullptr && over->identity != 0) cfg.identity = over->identity;
    if (over != nullptr && over->correction != 0) cfg.correction = over->correction;
    if (small) {  // four waves per tile (table tile 32) or one (tile 16); no k slices
        cfg.variant = 5;
        cfg.tile = tt == 16 ? 16 : 0;
        cfg.kSplit = 1;
    } else {
        cfg.variant = cfg.tile == 128 ? 3 : 1;
        // In a batch that fills the GPU without K splits, a member with the rounding
        // correction on takes none unless the call pins one: the persistent kernel then forms
        // its energies from its own loads, where a split member needs an energy pass over its
        // operands beside it (whose workgroups hold CUs the persistent kernel waits for), a
        // partial-plane merge and a longer recompute list, and the persistent queue balances
        // such a batch without the split. A small batch keeps the table's split, which it
        // needs to fill the CUs. (The same rule for every form of a call, so a member's bytes
        // do not depend on the form.)
        const bool corrected = cfg.correction != 0 ? cfg.correction > 0 : k >= 1024;
        const bool pinned = over != nullptr && (over->tile != 0 || over->kSplit != 0);
        if (batchFills && corrected && !pinned) cfg.kSplit = 1;
    }
    return cfg;
}

static aclblasStatus_t build_layout(aclblasHandle_t h, int count, const aclblasOperation_t* ta,
                                    const aclblasOperation_t* tb, const int* m, const int* n, const int* k,
                                    const aclblasComplex* alpha, const aclblasComplex* const* A, const int* lda,
                                    const aclblasComplex* const* B, const int* ldb, const aclblasComplex* beta,
                                    aclblasComplex* const* C, const int* ldc, bool withPointers,
                                    const aclblasCgemm3mConfig* over, GroupedLayout& L, const char* op,
                                    int fills = -1)
{
    if (fills < 0) {
        long long blocks64 = 0;
        for (int i = 0; i < count; ++i)
            if (m[i] > 0 && n[i] > 0 && k[i] > 0) blocks64 += (long long)((m[i] + 63) / 64) * ((n[i] + 63) / 64);
        fills = grouped_batch_fills(blocks64, h->cus) ? 1 : 0;
    }
    // variant 7 keeps the earlier form: members packed into planes by one launch first; so
    // does a batch on the standard identity (the raw kernel forms the fused identity's planes)
    L.raw = !(over != nullptr && (over->variant == 7 || over->identity == ID_STANDARD));
    // The raw quad kernel addresses a tile's elements by 32-bit byte offsets (at most 68 leading
    // dimensions of 8 bytes from the tile's base); wider leading dimensions take the packed form.
    for (int i = 0; i < count && withPointers; ++i)
        if ((long long)std::max(lda[i], ldb[i]) * 8 * 68 >= (1LL << 31)) L.raw = false;
    // The tuned table gives a class to the one-launch kernel from batches of that class alone.
    // In a batch whose quad-kernel members carry at least as much work, its persistent
    // workgroups hold every CU until their queue empties, so the one-launch launches would in
    // effect run after it: such members join the quad kernel instead (64 x 64 tiles, no K
    // split), where they fill its queue. Unless tiles, K splits or the one-launch kernel are
    // pinned; the same in both forms (variant 7 included), so both run the same plans.
    bool absorb = false;
    if (over == nullptr || ((over->variant == 0 || over->variant == 7) && over->tile == 0 && over->kSplit == 0)) {
        double quadWork = 0.0, smallWork = 0.0;
        for (int i = 0; i < count; ++i) {
            if (m[i] <= 0 || n[i] <= 0 || k[i] <= 0) continue;
            const double w = (double)m[i] * n[i] * k[i];
            (grouped_member_config(m[i], n[i], k[i], over, fills != 0).variant == 5 ? smallWork : quadWork) += w;
        }
        absorb = smallWork > 0.0 && quadWork >= smallWork;
    }
    for (int i = 0; i < count; ++i) {
        const aclblasStatus_t st =
            validate(h, ta[i], tb[i], m[i], n[i], k[i], &alpha[i], withPointers ? lda[i] : std::max(1, ta[i] != ACLBLAS_OP_N ? k[i] : m[i]),
                     withPointers ? ldb[i] : std::max(1, tb[i] != ACLBLAS_OP_N ? n[i] : k[i]), &beta[i],
                     withPointers ? ldc[i] : std::max(1, m[i]), op);
        if (st != ACLBLAS_STATUS_SUCCESS) return st;
        if (m[i] == 0 || n[i] == 0) continue;
        Problem p{ta[i], tb[i], m[i], n[i], k[i], alpha[i], beta[i],
                  withPointers ? A[i] : nullptr, withPointers ? lda[i] : 1,
                  withPointers ? B[i] : nullptr, withPointers ? ldb[i] : 1,
                  withPointers ? C[i] : nullptr, withPointers ? ldc[i] : std::max(1, m[i])};
        if (withPointers) {
            const aclblasStatus_t sp = validate_pointers(is_zero(p.alpha) ? 0 : p.k, p.A, p.B, p.C, op);
            if (sp != ACLBLAS_STATUS_SUCCESS) return sp;
        }
        if (p.k == 0 || is_zero(p.alpha)) {
            if (!(p.beta.real == 1.0f && p.beta.imag == 0.0f)) L.scaleOnly.push_back(p);
            continue;
        }
        aclblasCgemm3mConfig cfg = grouped_member_config(p.m, p.n, p.k, over, fills != 0);
        if (absorb && cfg.variant == 5) {
            cfg.tile = 64;
            cfg.kSplit = 1;
            cfg.variant = 1;
        }
        Plan pl;
        if (!make_plan(p, &cfg, h->cus, pl, op)) return ACLBLAS_STATUS_INVALID_VALUE;
        if (pl.path == 5) {  // the one-launch kernel: no pack, no workspace
            L.small.push_back(p);
            L.smallPlans.push_back(pl);
            continue;
        }
        if (L.raw) raw_member_workspace(pl);
        L.probs.push_back(p);
        L.plans.push_back(pl);
    }
    size_t total = 0;
    size_t loose = 0;  // 64 x 64 blocks outside whole 128 x 128 regions
    for (size_t i = 0; i < L.probs.size(); ++i) {
        L.base.push_back(total);
        total = align256(total + L.plans[i].total);
        const Plan& pl = L.plans[i];
        const QuadCover qc = quad_cover(L.probs[i].m, L.probs[i].n);
        const size_t bm = (size_t)qc.bm, bn = (size_t)qc.bn;
        const size_t whole = (size_t)(qc.bmQ / 2) * (size_t)(qc.bnQ / 2);
        L.nQuad += whole * (size_t)pl.parts;
        loose += (bm * bn - 4 * whole) * (size_t)pl.parts;
        if (pl.parts > 1) {
            ++L.nMerge;
            L.maxMergeOutputs = std::max<long long>(L.maxMergeOutputs, (long long)L.probs[i].m * L.probs[i].n);
        }
        if (pl.correction) ++L.nCorr;
        if (pl.correction && pl.parts > 1) ++L.nCorrSplit;
    }
    L.nQuad += (loose + 3) / 4;
    L.nProb = L.probs.size();
    // raw: energy-only items of the corrected K-split members; energy items: one per corrected
    // member (raw: bitmap clears, plus the K-split members' energy sums)
    L.nPack = 2 * (L.raw ? L.nCorrSplit : L.nProb);
    L.energyBytes = (L.nCorr + (L.raw ? L.nCorrSplit : 0)) * sizeof(EnergyArgs);
    L.descBytes = align256(L.nProb * sizeof(GroupProblem)) + align256(L.nPack * sizeof(GroupPack)) +
                  align256(4 * L.nQuad * sizeof(GroupTile)) + align256(L.nMerge * sizeof(GroupMerge)) +
                  align256(L.small.size() * sizeof(GroupSmall)) + align256(L.energyBytes) +
                  align256(L.nCorr * sizeof(RecomputeArgs));
    L.offQueue = total;
    total += 256;
    L.total = L.nProb ? total : 0;
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace c3m

using namespace c3m;

extern "C" aclblasStatus_t aclblasCgemm3mGroupedBatchedGetWorkspaceSize(aclblasHandle_t handle, int count,
                                                                        const aclblasOperation_t* transa,
                                                                        const aclblasOperation_t* transb,
                                                                        const int* m, const int* n, const int* k,
                                                                        int precision, size_t* bytes)
{
    const char* op = "aclblasCgemm3mGroupedBatchedGetWorkspaceSize";
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (bytes == nullptr || count < 0) return ACLBLAS_STATUS_INVALID_VALUE;
    if (count == 0) {
        *bytes = 0;
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (transa == nullptr || transb == nullptr || m == nullptr || n == nullptr || k == nullptr)
        return ACLBLAS_STATUS_INVALID_VALUE;
    if (precision != ACLBLAS_CGEMM3M_FP32) return ACLBLAS_STATUS_NOT_SUPPORTED;
    if (ensure_device(handle) != ACLBLAS_STATUS_SUCCESS) handle->cus = 0;
    std::vector<aclblasComplex> one(count, aclblasComplex{1.0f, 0.0f}), zero(count, aclblasComplex{0.0f, 0.0f});
    // The larger of the two forms' needs: the call reads the operands as stored unless a
    // leading dimension is too wide for that kernel's offsets, which only the call sees.
    GroupedLayout L, Lp;
    aclblasCgemm3mConfig packed{};
    packed.variant = 7;
    aclblasStatus_t st = build_layout(handle, count, transa, transb, m, n, k, one.data(), nullptr, nullptr, nullptr,
                                      nullptr, zero.data(), nullptr, nullptr, false, nullptr, L, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    st = build_layout(handle, count, transa, transb, m, n, k, one.data(), nullptr, nullptr, nullptr, nullptr,
                      zero.data(), nullptr, nullptr, false, &packed, Lp, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    *bytes = std::max(L.total, Lp.total);
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t c3m::grouped_run(aclblasHandle_t handle, int count, const aclblasOperation_t* transa,
                                   const aclblasOperation_t* transb, const int* m, const int* n, const int* k,
                                   const aclblasComplex* alpha, const aclblasComplex* const* A, const int* lda,
                                   const aclblasComplex* const* B, const int* ldb, const aclblasComplex* beta,
                                   aclblasComplex* const* C, const int* ldc, int precision,
                                   const aclblasCgemm3mConfig* over, const char* op, int fills)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (count < 0) return ACLBLAS_STATUS_INVALID_VALUE;
    if (count == 0) return ACLBLAS_STATUS_SUCCESS;
    if (transa == nullptr || transb == nullptr || m == nullptr || n == nullptr || k == nullptr || alpha == nullptr ||
        A == nullptr || lda == nullptr || B == nullptr || ldb == nullptr || beta == nullptr || C == nullptr ||
        ldc == nullptr) {
        log_error(op, "argument arrays must not be nullptr");
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (precision != ACLBLAS_CGEMM3M_FP32) {
        log_error(op, "precision %d is not supported for grouped batches", precision);
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    aclblasStatus_t st = ensure_device(handle);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        // validation first, as the one-problem entry does, then report the missing device
        GroupedLayout L0;
        handle->cus = 0;
        const aclblasStatus_t v = build_layout(handle, count, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta,
                                               C, ldc, true, over, L0, op, fills);
        return v != ACLBLAS_STATUS_SUCCESS ? v : st;
    }
    GroupedLayout L;
    st = build_layout(handle, count, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, true, over, L, op,
                      fills);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    hipStream_t s = handle->stream;
    hipError_t e = hipSuccess;
    if (L.nProb > 0 || !L.small.empty()) {
        char* ws = nullptr;
        if (L.nProb > 0) {
            st = acquire_workspace(handle, L.total, &ws, op);
            if (st != ACLBLAS_STATUS_SUCCESS) return st;
        }
        // Descriptor block: [problems][pack items][quads: four 64 x 64 blocks each][merge items]
        // [one-launch members: four-wave form, then single-wave form]
        std::vector<char> blob(L.descBytes, 0);
        GroupProblem* gp = reinterpret_cast<GroupProblem*>(blob.data());
        const size_t offPack = align256(L.nProb * sizeof(GroupProblem));
        GroupPack* pk = reinterpret_cast<GroupPack*>(blob.data() + offPack);
        const size_t offTiles = offPack + align256(L.nPack * sizeof(GroupPack));
        GroupTile* tl = reinterpret_cast<GroupTile*>(blob.data() + offTiles);
        const size_t offMerge = offTiles + align256(4 * L.nQuad * sizeof(GroupTile));
        GroupMerge* mg = reinterpret_cast<GroupMerge*>(blob.data() + offMerge);
        struct Cost {
            long long cost;
            GroupTile t;
        };
        struct Quad {
            long long cost;
            GroupTile t[4];
        };
        std::vector<Quad> quads;
        std::vector<Cost> loose;
        unsigned int* queue0 = L.nProb > 0 ? reinterpret_cast<unsigned int*>(ws + L.offQueue) : nullptr;
        quads.reserve(L.nQuad);
        // members with the rounding correction: energies and recompute (last). Packed form:
        // one energy item per member (sums of the pack's partials, bitmap clear). Raw form:
        // bitmap clears for every such member before the quad kernel (the first item also
        // clears the tile queue), and energy sums of the K-split ones, whose merge screens.
        std::vector<EnergyArgs> energies, sums;
        std::vector<RecomputeArgs> recomputes;
        long long maxEnergyWork = 0, maxSumWork = 0, maxWords = 0;
        long long packBlocks = 0;
        size_t nm = 0, npk = 0;
        for (size_t i = 0; i < L.nProb; ++i) {
            const Problem& p = L.probs[i];
            const Plan& pl = L.plans[i];
            char* pws = ws + L.base[i];
            // the same argument blocks the one-problem path builds
            PackArgs pa{}, pb{};
            for (int role = 0; role < 2; ++role) {
                PackArgs& a = role == 0 ? pa : pb;
                const bool isA = role == 0;
                const aclblasOperation_t t = isA ? p.ta : p.tb;
                a.src = isA ? (const void*)p.A : (const void*)p.B;
                a.ld = isA ? p.lda : p.ldb;
                a.mn = isA ? p.m : p.n;
                a.kk = p.k;
                a.mnPad = isA ? pl.mPad : pl.nPad;
                a.kPad = pl.kPad;
                a.mnContig = isA ? (t == ACLBLAS_OP_N) : (t != ACLBLAS_OP_N);
                a.conj = t == ACLBLAS_OP_C;
                a.role = role;
                a.identity = pl.identity;
                a.format = pl.format;
                a.pieces = pl.pieces;
                a.nplanes = 3;  // grouped members use a 3M identity
                a.planes = L.raw ? nullptr : pws + (isA ? pl.offA : pl.offB);
                a.planeBytes = isA ? pl.aPlaneBytes : pl.bPlaneBytes;
                a.energyPartial =
                    pl.correction ? reinterpret_cast<double*>(pws + (isA ? pl.offEPA : pl.offEPB)) : nullptr;
                a.scaleExp = nullptr;
                if (L.raw && !(pl.correction && pl.parts > 1)) continue;  // read as stored by the quad kernel
                pk[npk].a = a;
                pk[npk].firstBlock = packBlocks;
                ++npk;
                packBlocks += (long long)(a.mnPad / 64) * ((a.kPad + 63) / 64);
            }
            GemmArgs& g = gp[i].g;
            g = GemmArgs{};
            g.Ap = L.raw ? nullptr : pa.planes;
            g.aPlaneBytes = pl.aPlaneBytes;
            g.mPad = pl.mPad;
            g.Bp = L.raw ? nullptr : pb.planes;
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
            g.partial = pl.parts > 1 ? reinterpret_cast<float*>(pws + pl.offPartial) : nullptr;
            g.partStride = pl.partStride;
            g.src = SourceOperands{p.A, p.lda, (int)p.ta, p.B, p.ldb, (int)p.tb, p.k};
            // the one-problem path's rounding correction (FP32 products: unit 2^-24)
            g.correction = pl.correction;
            if (pl.correction) {
                const double amag =
                    std::sqrt((double)p.alpha.real * p.alpha.real + (double)p.alpha.imag * p.alpha.imag);
                g.energyInline = L.raw && pl.parts == 1 ? 1 : 0;
                if (!g.energyInline) {
                    g.eA = reinterpret_cast<const double*>(pws + pl.offEA);
                    g.eB = reinterpret_cast<const double*>(pws + pl.offEB);
                }
                g.corrScale = amag * 0x1p-24;
                g.corrTau = 0x1p-8;
                g.bitmap = reinterpret_cast<unsigned long long*>(pws + pl.offBitmap);
                if (!L.raw) {
                    energies.push_back(EnergyArgs{pa.energyPartial, pl.mPad, pb.energyPartial, pl.nPad, pl.kTiles,
                                                  const_cast<double*>(g.eA), const_cast<double*>(g.eB), g.bitmap,
                                                  pl.bitmapWords, nullptr});
                    maxEnergyWork =
                        std::max<long long>(maxEnergyWork, energy_work(pl.mPad, pl.nPad, pl.kTiles, pl.bitmapWords));
                } else {
                    energies.push_back(EnergyArgs{nullptr, 0, nullptr, 0, 0, nullptr, nullptr, g.bitmap,
                                                  pl.bitmapWords, energies.empty() ? queue0 : nullptr});
                    maxEnergyWork = std::max<long long>(maxEnergyWork, pl.bitmapWords);
                    if (pl.parts > 1) {
                        sums.push_back(EnergyArgs{pa.energyPartial, pl.mPad, pb.energyPartial, pl.nPad, pl.kTiles,
                                                  const_cast<double*>(g.eA), const_cast<double*>(g.eB), nullptr, 0,
                                                  nullptr});
                        maxSumWork = std::max<long long>(maxSumWork, energy_work(pl.mPad, pl.nPad, pl.kTiles, 0));
                    }
                }
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
                r.bitmap = g.bitmap;
                r.words = pl.bitmapWords;
                r.C = p.C;
                r.ldc = p.ldc;
                r.counter = nullptr;
                r.netlib = pl.correction == 2 ? 1 : 0;
                r.alphaRf = p.alpha.real;
                r.alphaIf = p.alpha.imag;
                r.betaRf = p.beta.real;
                r.betaIf = p.beta.imag;
                recomputes.push_back(r);
                maxWords = std::max<long long>(maxWords, pl.bitmapWords);
            }
            // 64 x 64 blocks: whole 128 x 128 regions of full blocks as quads in the 128 x 128
            // kernel's wave order (wave w: block row 2r + w % 2, block column 2c + w / 2); the
            // other blocks (odd last block row or column, and a last block row or column with
            // at most 32 outputs, which the kernel runs on one 32-wide fragment) go to the loose
            // list, costed by their MFMA work: k steps x 32 x 32 fragments
            const long long stepsPerPart = (pl.kSteps + pl.parts - 1) / pl.parts;
            const QuadCover qc = quad_cover(p.m, p.n);
            const int bm = qc.bm, bn = qc.bn;
            for (int part = 0; part < pl.parts; ++part) {
                for (int tn = 0; tn < qc.bnQ; tn += 2)
                    for (int tm = 0; tm < qc.bmQ; tm += 2) {
                        Quad qd;
                        qd.cost = 4 * stepsPerPart;
                        for (int w = 0; w < 4; ++w) qd.t[w] = GroupTile{(int)i, part, tm + w % 2, tn + w / 2};
                        quads.push_back(qd);
                    }
                for (int tn = 0; tn < bn; ++tn)
                    for (int tm = 0; tm < bm; ++tm)
                        if (tm >= qc.bmQ || tn >= qc.bnQ) {
                            const int frags = (tm == bm - 1 ? qc.fm : 2) * (tn == bn - 1 ? qc.fn : 2);
                            loose.push_back(Cost{frags * stepsPerPart, GroupTile{(int)i, part, tm, tn}});
                        }
            }
            if (pl.parts > 1) {
                MergeArgs& a = mg[nm++].a;
                a = MergeArgs{};
                a.partial = g.partial;
                a.partStride = pl.partStride;
                a.parts = pl.parts;
                a.ldp = pl.mPad;
                a.m = p.m;
                a.n = p.n;
                a.alphaR = g.alphaR;
                a.alphaI = g.alphaI;
                a.betaR = g.betaR;
                a.betaI = g.betaI;
                a.betaZero = g.betaZero;
                a.C = p.C;
                a.ldc = p.ldc;
                a.correction = g.correction;
                a.eA = g.eA;
                a.eB = g.eB;
                a.corrScale = g.corrScale;
                a.corrTau = g.corrTau;
                a.bitmap = g.bitmap;
                a.src = g.src;
            }
        }
        // Longest first; ties in (problem, part, tn, tm) order, so the list is a pure function
        // of the arguments. Loose blocks are grouped four at a time in that order (blocks of
        // like cost share a workgroup); the last quad is filled with empty slots.
        auto before = [](long long ca, const GroupTile& a, long long cb, const GroupTile& b) {
            if (ca != cb) return ca > cb;
            if (a.problem != b.problem) return a.problem < b.problem;
            if (a.part != b.part) return a.part < b.part;
            if (a.tn != b.tn) return a.tn < b.tn;
            return a.tm < b.tm;
        };
        std::sort(loose.begin(), loose.end(),
                  [&](const Cost& a, const Cost& b) { return before(a.cost, a.t, b.cost, b.t); });
        for (size_t j = 0; j < loose.size(); j += 4) {
            Quad qd;
            qd.cost = loose[j].cost;
            for (size_t w = 0; w < 4; ++w)
                qd.t[w] = j + w < loose.size() ? loose[j + w].t : GroupTile{-1, 0, 0, 0};
            quads.push_back(qd);
        }
        std::sort(quads.begin(), quads.end(),
                  [&](const Quad& a, const Quad& b) { return before(a.cost, a.t[0], b.cost, b.t[0]); });
        if (quads.size() != L.nQuad || L.nQuad > (size_t)std::numeric_limits<int>::max()) {
            log_error(op, "internal error: %zu quads laid out, %zu counted", quads.size(), L.nQuad);
            return ACLBLAS_STATUS_INTERNAL_ERROR;
        }
        size_t ti = 0;
        for (const Quad& qd : quads)
            for (int w = 0; w < 4; ++w) tl[ti++] = qd.t[w];
        // one-launch members, grouped by waves per tile (one launch per form)
        const size_t offSmall = offMerge + align256(L.nMerge * sizeof(GroupMerge));
        GroupSmall* sm = reinterpret_cast<GroupSmall*>(blob.data() + offSmall);
        size_t nSmall[2] = {0, 0};     // four-wave, single-wave members
        long long smallBlocks[2] = {0, 0};
        for (int form = 0; form < 2; ++form)
            for (size_t i = 0; i < L.small.size(); ++i) {
                const Problem& p = L.small[i];
                const Plan& pl = L.smallPlans[i];
                if ((pl.smallWaves == 1) != (form == 1)) continue;
                GroupSmall& it = sm[nSmall[0] + nSmall[1]];
                const double amag =
                    std::sqrt((double)p.alpha.real * p.alpha.real + (double)p.alpha.imag * p.alpha.imag);
                it.a = SmallArgs{p.A, p.lda, (int)p.ta, p.B, p.ldb, (int)p.tb, p.m, p.n, p.k, pl.identity,
                                 p.alpha.real, p.alpha.imag, p.beta.real, p.beta.imag, is_zero(p.beta) ? 1 : 0,
                                 p.C, p.ldc, pl.correction, amag * 0x1p-24, 0x1p-8, nullptr, 1, pl.smallWaves,
                                 nullptr, nullptr, nullptr};
                it.tilesM = (p.m + 15) / 16;
                it.tiles = (int)pl.smallTiles;
                it.firstBlock = smallBlocks[form];
                smallBlocks[form] += pl.smallTiles;
                ++nSmall[form];
            }
        const size_t offEnergy = offSmall + align256(L.small.size() * sizeof(GroupSmall));
        const size_t offRecompute = offEnergy + align256(L.energyBytes);
        unsigned int* queue = queue0;
        if (energies.size() != L.nCorr || recomputes.size() != L.nCorr || npk != L.nPack ||
            sums.size() != (L.raw ? L.nCorrSplit : 0)) {
            log_error(op, "internal error: %zu corrected members and %zu pack items laid out, %zu and %zu counted",
                      energies.size(), npk, L.nCorr, L.nPack);
            return ACLBLAS_STATUS_INTERNAL_ERROR;
        }
        if (L.nCorr > 0) {
            std::memcpy(blob.data() + offEnergy, energies.data(), energies.size() * sizeof(EnergyArgs));
            if (!sums.empty())
                std::memcpy(blob.data() + offEnergy + energies.size() * sizeof(EnergyArgs), sums.data(),
                            sums.size() * sizeof(EnergyArgs));
            std::memcpy(blob.data() + offRecompute, recomputes.data(), L.nCorr * sizeof(RecomputeArgs));
        }

        size_t slot = 0;
        const char* dDesc = stage_descriptors(handle, blob, s, &slot, &e);
        if (dDesc == nullptr) {
            log_error(op, "descriptor staging failed: %s", hipGetErrorString(e));
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        // Work that the quad kernel does not wait for runs on the handle's second stream, forked
        // after the descriptor upload and joined before the merge (under hipGraph capture this
        // becomes two parallel branches): the one-launch members, which share nothing with
        // the quad-kernel members, and, in the raw form, the energies of the K-split corrected
        // members, which only their merge reads. Side-stream kernel times are not in the
        // stage timings.
        const GroupSmall* dSmall = reinterpret_cast<const GroupSmall*>(dDesc + offSmall);
        const GroupProblem* dProb = reinterpret_cast<const GroupProblem*>(dDesc);
        const GroupPack* dPack = reinterpret_cast<const GroupPack*>(dDesc + offPack);
        const GroupTile* dTiles = reinterpret_cast<const GroupTile*>(dDesc + offTiles);
        const GroupMerge* dMerge = reinterpret_cast<const GroupMerge*>(dDesc + offMerge);
        const EnergyArgs* dEnergy = reinterpret_cast<const EnergyArgs*>(dDesc + offEnergy);
        const bool fork = L.nProb > 0 && (!L.small.empty() || !sums.empty());
        hipStream_t ss = s;
        // Raw form: the tile-queue reset and bitmap clears go first, before the fork, so that
        // the quad kernel right behind them does not wait for CUs the second stream's kernels
        // took first.
        if (e == hipSuccess && L.raw && L.nProb > 0) {
            if (L.nCorr > 0) {
                StageTimer t(handle, ST_ENERGY);
                e = launch_energy_grouped(dEnergy, (int)L.nCorr, maxEnergyWork, s);  // also the queue
            } else {
                e = launch_fill_u64(reinterpret_cast<unsigned long long*>(queue), 1, s);  // the tile queue
            }
        }
        if (fork) {
            RelaxedCapture relaxed;
            if (handle->side == nullptr && hipStreamCreateWithFlags(&handle->side, hipStreamNonBlocking) != hipSuccess)
                handle->side = nullptr;
            if (handle->forkEv == nullptr &&
                hipEventCreateWithFlags(&handle->forkEv, hipEventDisableTiming) != hipSuccess)
                handle->forkEv = nullptr;
            if (handle->joinEv == nullptr &&
                hipEventCreateWithFlags(&handle->joinEv, hipEventDisableTiming) != hipSuccess)
                handle->joinEv = nullptr;
            if (handle->side == nullptr || handle->forkEv == nullptr || handle->joinEv == nullptr) {
                log_error(op, "second stream or its events could not be created");
                return ACLBLAS_STATUS_EXECUTION_FAILED;
            }
            ss = handle->side;
            e = hipEventRecord(handle->forkEv, s);
            if (e == hipSuccess) e = hipStreamWaitEvent(ss, handle->forkEv, 0);
        }
        if (e == hipSuccess && L.raw && !sums.empty()) {
            e = launch_energy_partial_grouped(dPack, (int)L.nPack, packBlocks, ss);
            if (e == hipSuccess) e = launch_energy_grouped(dEnergy + energies.size(), (int)sums.size(), maxSumWork, ss);
        }
        // single-wave form first: it is short, and a launch that follows the quad kernel's
        // start finds every CU held by it until the queue empties
        for (int f = 0; f < 2 && e == hipSuccess; ++f) {
            const int form = 1 - f;
            if (nSmall[form] == 0) continue;
            if (fork) {
                e = launch_gemm_small_grouped(dSmall + (form == 0 ? 0 : nSmall[0]), (int)nSmall[form],
                                              smallBlocks[form], form == 0 ? 4 : 1, ss);
            } else {
                StageTimer t(handle, ST_GEMM);
                e = launch_gemm_small_grouped(dSmall + (form == 0 ? 0 : nSmall[0]), (int)nSmall[form],
                                              smallBlocks[form], form == 0 ? 4 : 1, ss);
            }
        }
        if (fork && e == hipSuccess) e = hipEventRecord(handle->joinEv, ss);
        if (e == hipSuccess && !L.raw && L.nPack > 0) {
            StageTimer t(handle, ST_PACK);
            e = launch_pack_grouped(dPack, (int)L.nPack, packBlocks, queue, s);  // also clears the queue
        }
        if (e == hipSuccess && !L.raw && L.nCorr > 0) {  // packed form: energy sums and bitmap clears
            StageTimer t(handle, ST_ENERGY);
            e = launch_energy_grouped(dEnergy, (int)L.nCorr, maxEnergyWork, s);
        }
        if (e == hipSuccess && L.nProb > 0) {
            StageTimer t(handle, ST_GEMM);
            const int cus = handle->cus > 0 ? handle->cus : 256;
            if (!quads.empty())
                e = launch_gemm_f32_grouped_quads(dProb, dTiles, (int)quads.size(), queue,
                                                  std::min<int>((int)quads.size(), cus), L.raw, s);
        }
        if (fork && e == hipSuccess) e = hipStreamWaitEvent(s, handle->joinEv, 0);
        if (e == hipSuccess && nm > 0) {
            StageTimer t(handle, ST_MERGE);
            e = launch_merge_grouped(dMerge, (int)nm, L.maxMergeOutputs, s);
        }
        if (e == hipSuccess && L.nCorr > 0) {
            StageTimer t(handle, ST_RECOMPUTE);
            e = launch_recompute_grouped(reinterpret_cast<const RecomputeArgs*>(dDesc + offRecompute), (int)L.nCorr,
                                         maxWords, s);
        }
        // the slot is free for reuse once this call's kernels are done (a captured slot is
        // never reused)
        _aclblas_handle::Staged& used = handle->staged[slot];
        if (e == hipSuccess && !used.captured) {
            e = hipEventRecord(used.done, s);
            used.doneStream = s;
            used.recorded = e == hipSuccess;
        }
    }
    for (const Problem& p : L.scaleOnly) {
        if (e != hipSuccess) break;
        ScaleCArgs a{p.m, p.n, p.beta.real, p.beta.imag, is_zero(p.beta) ? 1 : 0, p.C, p.ldc};
        StageTimer t(handle, ST_SCALE);
        e = launch_scale_c(a, s);
    }
    if (e != hipSuccess) {
        log_error(op, "kernel launch failed: %s", hipGetErrorString(e));
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    if (handle->timing) handle->timedCalls++;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasCgemm3mGroupedBatched(aclblasHandle_t handle, int count,
                                                        const aclblasOperation_t* transa,
                                                        const aclblasOperation_t* transb, const int* m, const int* n,
                                                        const int* k, const aclblasComplex* alpha,
                                                        const aclblasComplex* const* A, const int* lda,
                                                        const aclblasComplex* const* B, const int* ldb,
                                                        const aclblasComplex* beta, aclblasComplex* const* C,
                                                        const int* ldc, int precision)
{
    return grouped_run(handle, count, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, precision,
                       nullptr, "aclblasCgemm3mGroupedBatched");
}

extern "C" aclblasStatus_t aclblasCgemm3mGroupedBatchedEx(aclblasHandle_t handle, int count,
                                                          const aclblasOperation_t* transa,
                                                          const aclblasOperation_t* transb, const int* m,
                                                          const int* n, const int* k, const aclblasComplex* alpha,
                                                          const aclblasComplex* const* A, const int* lda,
                                                          const aclblasComplex* const* B, const int* ldb,
                                                          const aclblasComplex* beta, aclblasComplex* const* C,
                                                          const int* ldc, const aclblasCgemm3mConfig* config)
{
    const char* op = "aclblasCgemm3mGroupedBatchedEx";
    // variant 5 pins the one-launch kernel for every member (tile 16: its single-wave form;
    // members are not k-sliced); otherwise tile 64 / 128 and kSplit pin the quad kernel's
    // plan; variant 7 packs the quad members into planes first instead of reading them raw
    const bool small = config != nullptr && config->variant == 5;
    if (config != nullptr &&
        ((config->variant != 0 && config->variant != 5 && config->variant != 7) ||
         (small ? (config->tile != 0 && config->tile != 16) || config->kSplit > 1
                : config->tile != 0 && config->tile != 64 && config->tile != 128) ||
         config->kSplit < 0 || config->kSplit > 16 || config->identity < 0 || config->identity > ID_STANDARD ||
         config->correction < -1 || config->correction > 2)) {
        log_error(op, "invalid grouped config (tile %d, kSplit %d, identity %d, variant %d, correction %d)",
                  config->tile, config->kSplit, config->identity, config->variant, config->correction);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return grouped_run(handle, count, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                       config != nullptr ? config->precision : ACLBLAS_CGEMM3M_FP32, config, op);
}
