// partition.hip — stable K-way partition and stream compaction.
//
// Replaces libWR/kernels/compact.cu:6-17 (one global atomic per survivor,
// nondeterministic order) with an ordered, atomic-free reduce-then-scan:
//   1. k_part_count  — each 1024-item tile counts its items per category
//                      with wave ballots (one popcount per wave per category);
//   2. k_part_scan   — one block scans the K x tiles count matrix
//                      (category-major), giving every (category, tile) its
//                      output offset;
//   3. k_part_scatter— each tile ranks its items stably (ballot prefix inside
//                      the wave, LDS prefix across the 4 waves, running total
//                      across the 4 rounds) and writes them.
// The output equals std::stable_partition / std::copy_if on the host.
#include <hip/hip_runtime.h>

#include "wrk_common.h"

namespace wrk {

constexpr uint32_t kTile = 1024;       // items per tile (256 lanes x 4 rounds)
constexpr uint32_t kRounds = kTile / 256;
constexpr uint32_t kMaxCat = 8;

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <uint32_t K>
__global__ __launch_bounds__(256) void k_part_count(const uint8_t* __restrict__ cat, uint32_t n,
                                                    uint32_t numTiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t waveCnt[4][K];
    const uint32_t tile = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t mine[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) mine[k] = 0;
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        uint32_t j = tile * kTile + r * 256 + threadIdx.x;
        uint32_t c = j < n ? cat[j] : 0xFFu;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) mine[k] += __popcll(__ballot(c == k));
    }
    if (lane == 0) {
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) waveCnt[wave][k] = mine[k];
    }
    __syncthreads();
    if (threadIdx.x < K) {
        uint32_t k = threadIdx.x;
        counts[k * numTiles + tile] = waveCnt[0][k] + waveCnt[1][k] + waveCnt[2][k] + waveCnt[3][k];
    }
}

// Exclusive scan over m = K * numTiles counts by one 1024-lane block;
// offsets[k] = first output slot of category k, offsets[K] = total kept.
__global__ __launch_bounds__(1024) void k_part_scan(uint32_t* __restrict__ counts, uint32_t m, uint32_t numTiles,
                                                    uint32_t K, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t carry;
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t begin = threadIdx.x * per;
    const uint32_t end = min(begin + per, m);
    uint32_t local = 0;
    for (uint32_t i = begin; i < end; ++i) local += counts[i];
    // block exclusive scan of `local`
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = local;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    if (lane == 63) waveSum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < 16; ++w) { uint32_t t = waveSum[w]; waveSum[w] = acc; acc += t; }
        carry = acc;
    }
    __syncthreads();
    uint32_t run = waveSum[wave] + incl - local;
    for (uint32_t i = begin; i < end; ++i) { uint32_t c = counts[i]; counts[i] = run; run += c; }
    __syncthreads();
    if (threadIdx.x < K) offsets[threadIdx.x] = (numTiles > 0) ? counts[threadIdx.x * numTiles] : 0u;
    if (threadIdx.x == 0) offsets[K] = carry;
}

template <uint32_t K>
__global__ __launch_bounds__(256) void k_part_scatter(const uint8_t* __restrict__ cat, const uint32_t* __restrict__ in,
                                                      uint32_t n, uint32_t numTiles,
                                                      const uint32_t* __restrict__ scanned, uint32_t* __restrict__ out) {
    __shared__ uint32_t waveCnt[4][K];
    const uint32_t tile = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    uint32_t base[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) base[k] = scanned[k * numTiles + tile];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        uint32_t j = tile * kTile + r * 256 + threadIdx.x;
        uint32_t c = j < n ? cat[j] : 0xFFu;
        uint32_t rank = 0;
        uint32_t myWaveCnt[K];
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            uint64_t m = __ballot(c == k);
            myWaveCnt[k] = __popcll(m);
            if (c == k) rank = __popcll(m & lt);
        }
        if (lane == 0) {
#pragma unroll
            for (uint32_t k = 0; k < K; ++k) waveCnt[wave][k] = myWaveCnt[k];
        }
        __syncthreads();
        if (c < K) {
            uint32_t before = 0;
            for (uint32_t w = 0; w < wave; ++w) before += waveCnt[w][c];
            uint32_t item = in ? in[j] : j;
            out[base[c] + before + rank] = item;
        }
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) base[k] += waveCnt[0][k] + waveCnt[1][k] + waveCnt[2][k] + waveCnt[3][k];
        __syncthreads();
    }
}

template <uint32_t K>
static int partition_impl(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t* out,
                          uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes, hipStream_t s) {
    const uint32_t numTiles = ceil_div(n, kTile);
    WRK_REQUIRE((uint64_t)K * numTiles * 4u <= scratchBytes, "wrk_partition: scratch too small");
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    if (numTiles > 0) {
        hipLaunchKernelGGL(k_part_count<K>, dim3(numTiles), dim3(256), 0, s, category, n, numTiles, counts);
    }
    hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, counts, K * numTiles, numTiles, K, offsetsOut);
    if (numTiles > 0) {
        hipLaunchKernelGGL(k_part_scatter<K>, dim3(numTiles), dim3(256), 0, s, category, in, n, numTiles, counts, out);
    }
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

}  // namespace wrk

using namespace wrk;

extern "C" uint64_t wrk_partition_scratch_bytes(uint32_t n, uint32_t numCategories) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint32_t k = numCategories < 1 ? 1 : numCategories;
    return (tiles * k + 64) * 4;
}

extern "C" int wrk_partition(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t numCategories,
                             uint32_t* out, uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes,
                             void* stream) {
    WRK_REQUIRE(offsetsOut && scratch, "wrk_partition: null argument");
    WRK_REQUIRE(n == 0 || (category && out), "wrk_partition: null argument");
    hipStream_t s = as_stream(stream);
    switch (numCategories) {
        case 1: return partition_impl<1>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 2: return partition_impl<2>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 3: return partition_impl<3>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 4: return partition_impl<4>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 5: return partition_impl<5>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 6: return partition_impl<6>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 7: return partition_impl<7>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 8: return partition_impl<8>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        default: set_error("wrk_partition: numCategories must be in [1, 8]"); return WRK_ERR_INVALID;
    }
}

// keep[i] != 0 -> category 0 (kept), else category 1 (dropped: >= K).
namespace wrk {
__global__ __launch_bounds__(256) void k_keep_to_cat(const uint8_t* __restrict__ keep, uint32_t n,
                                                     uint8_t* __restrict__ cat) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) cat[i] = keep[i] ? 0 : 1;
}
}  // namespace wrk

extern "C" int wrk_compact(const uint8_t* keep, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* countOut,
                           void* scratch, uint64_t scratchBytes, void* stream) {
    WRK_REQUIRE(countOut && scratch, "wrk_compact: null argument");
    WRK_REQUIRE(n == 0 || (keep && out), "wrk_compact: null argument");
    // keep bytes are used directly as categories when they are 0/1: map
    // nonzero -> 0 into the tail of the scratch buffer.
    const uint32_t numTiles = ceil_div(n, kTile);
    uint64_t countsBytes = ((uint64_t)numTiles + 64) * 4;
    uint64_t catOff = (countsBytes + 255) & ~uint64_t(255);
    WRK_REQUIRE(scratchBytes >= catOff + n, "wrk_compact: scratch too small (need wrk_compact_scratch_bytes)");
    hipStream_t s = as_stream(stream);
    uint8_t* cat = static_cast<uint8_t*>(scratch) + catOff;
    if (n > 0) hipLaunchKernelGGL(k_keep_to_cat, dim3(ceil_div(n, 256)), dim3(256), 0, s, keep, n, cat);
    // offsets for K=1: [0, total] -> total lands in countOut via a 2-word tail
    uint32_t* offs = static_cast<uint32_t*>(scratch) + numTiles + 8;
    int rc = partition_impl<1>(cat, in, n, out, offs, scratch, countsBytes, s);
    if (rc != WRK_OK) return rc;
    WRK_HIP_TRY(hipMemcpyAsync(countOut, offs + 1, 4, hipMemcpyDeviceToDevice, s));
    return WRK_OK;
}

extern "C" uint64_t wrk_compact_scratch_bytes(uint32_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t countsBytes = (tiles + 64) * 4;
    return ((countsBytes + 255) & ~uint64_t(255)) + n;
}
