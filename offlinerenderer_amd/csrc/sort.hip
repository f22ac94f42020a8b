// sort.hip — stable LSD radix sort of (uint64 key, uint32 value) pairs for
// the BVH build's Morton codes (inside optixAccelBuild in the reference,
// libWR/src/internal/scene.cpp:525-550, which has no source counterpart).
//
// One pass per 8-bit digit, three steps:
//   1. k_sort_count   — each 2048-key tile counts its digits in an LDS
//                       histogram; counts are stored digit-major
//                       (counts[digit * tiles + tile]);
//   2. a device-wide exclusive scan of the 256 x tiles counts gives every
//      (digit, tile) its first output position;
//   3. k_sort_scatter — each tile ranks its keys stably in 8 rounds of 256
//                       (a key's peers with the same digit found with 8 wave
//                       ballots, one bit each; waves ordered through LDS),
//                       stages the tile in LDS in digit order, and writes
//                       each digit's run to its output range, so neighbouring
//                       lanes write neighbouring addresses.
// Every pass is stable, so the result equals std::stable_sort by key.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wrk_common.h"

namespace wrk {

namespace {
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortRounds = 8;                         // keys per thread per tile
constexpr uint32_t kSortTile = kSortThreads * kSortRounds;  // 2048 keys
constexpr uint32_t kRadix = 256;                            // 8-bit digits

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

__global__ __launch_bounds__(256) void k_sort_count(const uint64_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                    uint32_t mask, uint32_t numTiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kRadix];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(uint32_t)(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    counts[threadIdx.x * numTiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_sort_scatter(const uint64_t* __restrict__ keysIn,
                                                      const uint32_t* __restrict__ valsIn, uint32_t n, uint32_t shift,
                                                      uint32_t mask, uint32_t numTiles,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ offsets,
                                                      uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    __shared__ uint64_t sKey[kSortTile];
    __shared__ uint32_t sVal[kSortTile];
    __shared__ uint32_t sStart[kRadix];    // tile-local first position of each digit's run
    __shared__ uint32_t sNext[kRadix];     // next free tile-local position of each digit
    __shared__ uint32_t sWave[4][kRadix];  // this round: keys of each digit per wave
    __shared__ uint32_t sPre[4][kRadix];   // this round: first position of each digit per wave
    __shared__ uint32_t sWaveSum[4];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t tile = blockIdx.x;
    const uint64_t below = lanes_below();

    // digit runs of this tile: exclusive scan of its histogram (thread t
    // holds digit t)
    const uint32_t h = counts[t * numTiles + tile];
    uint32_t incl = h;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) sWaveSum[wave] = incl;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) sWave[w][t] = 0;
    __syncthreads();
    uint32_t start = incl - h;
    for (uint32_t w = 0; w < wave; ++w) start += sWaveSum[w];
    sStart[t] = start;
    sNext[t] = start;
    __syncthreads();

    const uint32_t base = tile * kSortTile;
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + t;
        const bool valid = i < n;
        const uint64_t key = valid ? keysIn[i] : 0ull;
        const uint32_t val = valid ? valsIn[i] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        // lanes of this wave holding the same digit: agree on all 8 bits
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t set = __ballot(bit);
            peers &= bit ? set : ~set;
        }
        const uint32_t rank = __popcll(peers & below);
        if (valid && rank == 0) sWave[wave][d] = __popcll(peers);
        __syncthreads();
        {
            // digit t: this round's keys of wave 0, then 1, 2, 3, after the
            // keys of earlier rounds
            const uint32_t c0 = sWave[0][t], c1 = sWave[1][t], c2 = sWave[2][t], c3 = sWave[3][t];
            const uint32_t b0 = sNext[t];
            sPre[0][t] = b0;
            sPre[1][t] = b0 + c0;
            sPre[2][t] = b0 + c0 + c1;
            sPre[3][t] = b0 + c0 + c1 + c2;
            sNext[t] = b0 + c0 + c1 + c2 + c3;
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w) sWave[w][t] = 0;
        }
        __syncthreads();
        if (valid) {
            const uint32_t p = sPre[wave][d] + rank;
            sKey[p] = key;
            sVal[p] = val;
        }
    }
    __syncthreads();
    // the tile in digit order: each digit's run goes to its output range
    const uint32_t cnt = min(kSortTile, n - base);
    for (uint32_t p = t; p < cnt; p += kSortThreads) {
        const uint64_t key = sKey[p];
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        const uint32_t dst = offsets[d * numTiles + tile] + (p - sStart[d]);
        keysOut[dst] = key;
        valsOut[dst] = sVal[p];
    }
}

uint64_t radix_sort_scratch_bytes(uint32_t n) {
    const uint64_t m = (uint64_t)ceil_div(n, kSortTile) * kRadix;
    return 2 * align256(m * 4) + align256(exclusive_scan_scratch_bytes((uint32_t)m));
}

int radix_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n,
                         uint32_t beginBit, uint32_t endBit, void* scratch, uint64_t scratchBytes, hipStream_t s) {
    WRK_REQUIRE(beginBit <= endBit && endBit <= 64, "radix sort: bad bit range");
    if (n <= 1 || beginBit == endBit) return WRK_OK;
    WRK_REQUIRE(keys && vals && keysTmp && valsTmp && scratch, "radix sort: null argument");
    WRK_REQUIRE(scratchBytes >= radix_sort_scratch_bytes(n), "radix sort: scratch too small");
    const uint32_t numTiles = ceil_div(n, kSortTile);
    const uint32_t m = numTiles * kRadix;
    char* base = static_cast<char*>(scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(base + align256((size_t)m * 4));
    uint32_t* scanTmp = reinterpret_cast<uint32_t*>(base + 2 * align256((size_t)m * 4));
    const uint64_t scanBytes = scratchBytes - 2 * align256((size_t)m * 4);
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = keysTmp;
    uint32_t* vout = valsTmp;
    for (uint32_t shift = beginBit; shift < endBit; shift += 8) {
        const uint32_t bits = std::min(8u, endBit - shift);
        const uint32_t mask = (1u << bits) - 1u;
        hipLaunchKernelGGL(k_sort_count, dim3(numTiles), dim3(kSortThreads), 0, s, kin, n, shift, mask, numTiles,
                           counts);
        WRK_CHECK_LAUNCH();
        const int rc = exclusive_scan_u32(counts, offsets, m, scanTmp, scanBytes, nullptr, s);
        if (rc != WRK_OK) return rc;
        hipLaunchKernelGGL(k_sort_scatter, dim3(numTiles), dim3(kSortThreads), 0, s, kin, vin, n, shift, mask,
                           numTiles, counts, offsets, kout, vout);
        WRK_CHECK_LAUNCH();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        WRK_HIP_TRY(hipMemcpyAsync(keys, kin, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
        WRK_HIP_TRY(hipMemcpyAsync(vals, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    }
    return WRK_OK;
}

}  // namespace wrk

extern "C" uint64_t wrk_radix_sort_scratch_bytes(uint32_t n) { return wrk::radix_sort_scratch_bytes(n); }

extern "C" int wrk_radix_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp,
                                        uint32_t n, uint32_t beginBit, uint32_t endBit, void* scratch,
                                        uint64_t scratchBytes, void* stream) {
    return wrk::radix_sort_pairs_u64(keys, vals, keysTmp, valsTmp, n, beginBit, endBit, scratch, scratchBytes,
                                     wrk::as_stream(stream));
}
