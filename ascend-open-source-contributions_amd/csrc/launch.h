// launch.h -- host-callable launchers of the gfx950 kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>

#include "kernel_args.h"

namespace c3m {

hipError_t launch_pack(const PackArgs& a, hipStream_t s);
hipError_t launch_pack_pair(const PackArgs& a, const PackArgs& b, hipStream_t s);
hipError_t launch_scale_exponents(const ScaleArgs& a, const ScaleArgs& b, hipStream_t s);  // A and B, one launch
hipError_t launch_scale_summary(const ScaleSummaryArgs& a, hipStream_t s);
hipError_t launch_energy(const EnergyArgs& a, hipStream_t s);
hipError_t launch_gemm_f32(const GemmArgs& g, int bm, int bn, int variant, hipStream_t s);
hipError_t launch_gemm_bf16(const GemmArgs& g, int bm, int bn, int pieces, int variant, hipStream_t s);
hipError_t launch_gemm_lowp_products(const GemmArgs& g, int pieces, bool lds, hipStream_t s);
int lowp_products_lds_unroll(int pieces);
hipError_t launch_gemm_f16(const GemmArgs& g, int bm, int bn, int variant, hipStream_t s);
hipError_t launch_gemm_small(const SmallArgs& g, hipStream_t s);
hipError_t launch_gemm_small_grouped(const GroupSmall* items, int nitems, long long blocks, int waves, hipStream_t s);
hipError_t launch_merge(const MergeArgs& a, hipStream_t s);
hipError_t launch_scale_c(const ScaleCArgs& a, hipStream_t s);
hipError_t launch_recompute(const RecomputeArgs& a, hipStream_t s);
hipError_t launch_residual(const ResidualArgs& a, int blocksM, int blocksN, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, long long words, hipStream_t s);
hipError_t launch_pack_grouped(const GroupPack* items, int nitems, long long blocks, unsigned int* queue,
                               hipStream_t s);
hipError_t launch_energy_partial_grouped(const GroupPack* items, int nitems, long long blocks, hipStream_t s);
hipError_t launch_energy_partial_pair(const PackArgs& a, const PackArgs& b, hipStream_t s);
hipError_t launch_gemm_bf16_raw_products(const GemmArgs& g, int pieces, hipStream_t s);
hipError_t launch_gemm_f32_grouped_quads(const GroupProblem* probs, const GroupTile* tiles, int nquads,
                                         unsigned int* queue, int blocks, bool raw, hipStream_t s);
hipError_t launch_merge_grouped(const GroupMerge* items, int nitems, int maxOutputs, hipStream_t s);
hipError_t launch_energy_grouped(const EnergyArgs* items, int nitems, long long maxWork, hipStream_t s);
// threads an energy item wants (maxWork of launch_energy_grouped is the maximum over its items)
long long energy_work(int mnPadA, int mnPadB, int kTiles, long long bitmapWords);
hipError_t launch_recompute_grouped(const RecomputeArgs* items, int nitems, long long maxWords, hipStream_t s);

}  // namespace c3m
