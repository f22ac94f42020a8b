// partition.hip — stable K-way partition and stream compaction.
//
// Replaces libWR/kernels/compact.cu:6-17 (one global atomic per survivor,
// nondeterministic order) with an ordered, atomic-free reduce-then-scan:
//   1. k_part_count  — each 1024-item tile counts its items per category
//                      with wave ballots (one popcount per wave per category;
//                      one 4-byte category load per lane when aligned);
//   2. k_scan_*      — a three-launch scan of the K x tiles count matrix
//                      (category-major; one launch when it fits one
//                      segment) gives every (category, tile) its output
//                      offset;
//   3. k_part_scatter— each tile ranks its items stably (ballot prefix inside
//                      the wave, LDS prefix across the 4 waves, running total
//                      across the 4 rounds) and writes them.
// The output equals std::stable_partition / std::copy_if on the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "wrk_common.h"

namespace wrk {

constexpr uint32_t kTile = 1024;       // items per tile (256 lanes x 4 rounds)
constexpr uint32_t kRounds = kTile / 256;
constexpr uint32_t kMaxCat = 16;

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

// The same counts with one 4-byte load per lane: lane t of a tile holds its
// items 4t .. 4t+3 (counts do not depend on which lane holds an item). For a
// 4-byte-aligned category array; the last partial word is read bytewise so
// nothing past n is touched.
template <uint32_t K>
__global__ __launch_bounds__(256) void k_part_count_vec(const uint8_t* __restrict__ cat, uint32_t n,
                                                        uint32_t numTiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t waveCnt[4][K];
    const uint32_t tile = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t j0 = tile * kTile + threadIdx.x * 4u;
    uint32_t word = 0xFFFFFFFFu;
    if (j0 + 3u < n) {
        word = reinterpret_cast<const uint32_t*>(cat)[j0 >> 2];
    } else {
        for (uint32_t b = 0; b < 4u; ++b)
            if (j0 + b < n) word = (word & ~(0xFFu << (8u * b))) | ((uint32_t)cat[j0 + b] << (8u * b));
    }
    uint32_t mine[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) mine[k] = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) {
        const uint32_t c = (word >> (8u * b)) & 0xFFu;
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

// Exclusive scan over m = K * numTiles counts (category-major), in three
// launches: per-segment totals, one block scanning the (few) segment totals,
// then each segment rescanned with its carry. offsets[k] = first output slot
// of category k, offsets[K] = total kept.
constexpr uint32_t kScanSeg = 1024 * 8;  // counts per segment (1024 lanes x 8)

__device__ __forceinline__ uint32_t block_exclusive_scan_1024(uint32_t v, uint32_t* waveSum, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) waveSum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < 16; ++w) { uint32_t t = waveSum[w]; waveSum[w] = acc; acc += t; }
        *total = acc;
    }
    __syncthreads();
    uint32_t r = waveSum[wave] + incl - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(1024) void k_scan_segments(const uint32_t* __restrict__ counts, uint32_t m,
                                                        uint32_t* __restrict__ segTotals) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t total;
    const uint32_t base = blockIdx.x * kScanSeg + threadIdx.x * 8;
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i)
        if (base + i < m) local += counts[base + i];
    block_exclusive_scan_1024(local, waveSum, &total);
    if (threadIdx.x == 0) segTotals[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_totals(uint32_t* __restrict__ segTotals, uint32_t numSeg,
                                                      uint32_t K, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t total;
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < numSeg; b += 1024) {
        uint32_t i = b + threadIdx.x;
        uint32_t v = i < numSeg ? segTotals[i] : 0u;
        uint32_t ex = block_exclusive_scan_1024(v, waveSum, &total);
        if (i < numSeg) segTotals[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[K] = carry;
}

__global__ __launch_bounds__(1024) void k_scan_apply(uint32_t* __restrict__ counts, uint32_t m,
                                                     const uint32_t* __restrict__ segOffsets, uint32_t numTiles,
                                                     uint32_t K, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t total;
    const uint32_t base = blockIdx.x * kScanSeg + threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        v[i] = base + i < m ? counts[base + i] : 0u;
        local += v[i];
    }
    uint32_t run = segOffsets[blockIdx.x] + block_exclusive_scan_1024(local, waveSum, &total);
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        uint32_t e = base + i;
        if (e < m) {
            counts[e] = run;
            if (e % numTiles == 0) offsets[e / numTiles] = run;  // first tile of category e/numTiles
        }
        run += v[i];
    }
}

// The whole scan in one block when the count matrix fits one segment
// (m <= kScanSeg: up to 8192 / K tiles), which saves two launches on the
// small queues of short requests.
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* __restrict__ counts, uint32_t m, uint32_t numTiles,
                                                      uint32_t K, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t total;
    const uint32_t base = threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        v[i] = base + i < m ? counts[base + i] : 0u;
        local += v[i];
    }
    uint32_t run = block_exclusive_scan_1024(local, waveSum, &total);
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        uint32_t e = base + i;
        if (e < m) {
            counts[e] = run;
            if (e % numTiles == 0) offsets[e / numTiles] = run;
        }
        run += v[i];
    }
    if (threadIdx.x == 0) offsets[K] = total;
}

// ----------------------------------------------- general exclusive scan
// out[i] = sum of in[0 .. i) over m values (in == out allowed), with the
// same segment structure as the count-matrix scan: one block per 8,192
// values, one block over the segment totals, one pass applying them; a
// single launch when m fits one segment.
__global__ __launch_bounds__(1024) void k_gscan_single(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint32_t m, uint32_t* __restrict__ total) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t tot;
    const uint32_t base = threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        v[i] = base + i < m ? in[base + i] : 0u;
        local += v[i];
    }
    uint32_t run = block_exclusive_scan_1024(local, waveSum, &tot);
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        if (base + i < m) out[base + i] = run;
        run += v[i];
    }
    if (total && threadIdx.x == 0) *total = tot;
}

__global__ __launch_bounds__(1024) void k_gscan_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint32_t m, const uint32_t* __restrict__ segOffsets) {
    __shared__ uint32_t waveSum[16];
    __shared__ uint32_t total;
    const uint32_t base = blockIdx.x * kScanSeg + threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        v[i] = base + i < m ? in[base + i] : 0u;
        local += v[i];
    }
    uint32_t run = segOffsets[blockIdx.x] + block_exclusive_scan_1024(local, waveSum, &total);
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        if (base + i < m) out[base + i] = run;
        run += v[i];
    }
}

uint64_t exclusive_scan_scratch_bytes(uint32_t m) { return ((uint64_t)ceil_div(m, kScanSeg) + 16) * 4; }

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t m, uint32_t* scratch, uint64_t scratchBytes,
                       uint32_t* total, hipStream_t s) {
    if (m == 0) return WRK_OK;
    const uint32_t numSeg = ceil_div(m, kScanSeg);
    WRK_REQUIRE(scratchBytes >= exclusive_scan_scratch_bytes(m), "exclusive scan: scratch too small");
    if (numSeg == 1) {
        hipLaunchKernelGGL(k_gscan_single, dim3(1), dim3(1024), 0, s, in, out, m, total);
    } else {
        // segment totals, their scan (the grand total lands in scratch[numSeg]),
        // then each segment with its carry
        uint32_t* seg = scratch;
        hipLaunchKernelGGL(k_scan_segments, dim3(numSeg), dim3(1024), 0, s, in, m, seg);
        hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, s, seg, numSeg, 0u, seg + numSeg);
        hipLaunchKernelGGL(k_gscan_apply, dim3(numSeg), dim3(1024), 0, s, in, out, m, seg);
        if (total) WRK_HIP_TRY(hipMemcpyAsync(total, seg + numSeg, 4, hipMemcpyDeviceToDevice, s));
    }
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

__global__ void k_zero_offsets(uint32_t* offsets, uint32_t n) {
    if (threadIdx.x < n) offsets[threadIdx.x] = 0;
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

// Scatter with the layout of k_part_count_vec: lane t holds the tile's items
// 4t .. 4t+3, so item order within the tile is (lane, byte). An item of
// category k lands after the tile's earlier items of k: those of earlier
// waves (LDS totals), of earlier lanes of its wave (ballot prefixes over all
// four bytes) and of earlier bytes of its own lane. One barrier per tile.
template <uint32_t K>
__global__ __launch_bounds__(256) void k_part_scatter_vec(const uint8_t* __restrict__ cat,
                                                          const uint32_t* __restrict__ in, uint32_t n,
                                                          uint32_t numTiles, const uint32_t* __restrict__ scanned,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t waveTot[4][K];
    const uint32_t tile = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    const uint32_t j0 = tile * kTile + threadIdx.x * 4u;
    uint32_t word = 0xFFFFFFFFu;
    if (j0 + 3u < n) {
        word = reinterpret_cast<const uint32_t*>(cat)[j0 >> 2];
    } else {
        for (uint32_t b = 0; b < 4u; ++b)
            if (j0 + b < n) word = (word & ~(0xFFu << (8u * b))) | ((uint32_t)cat[j0 + b] << (8u * b));
    }
    uint32_t c[4];
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) c[b] = (word >> (8u * b)) & 0xFFu;
    uint32_t before[K];  // items of category k in earlier lanes of this wave
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        uint32_t tot = 0, bef = 0;
#pragma unroll
        for (uint32_t b = 0; b < 4u; ++b) {
            const uint64_t m = __ballot(c[b] == k);
            tot += __popcll(m);
            bef += __popcll(m & lt);
        }
        before[k] = bef;
        if (lane == 0) waveTot[wave][k] = tot;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) {
        const uint32_t k = c[b];
        if (k >= K) continue;
        uint32_t bef = 0;  // select, not a dynamically indexed register array
#pragma unroll
        for (uint32_t kk = 0; kk < K; ++kk) bef = kk == k ? before[kk] : bef;
        uint32_t pos = scanned[k * numTiles + tile] + bef;
        for (uint32_t w = 0; w < wave; ++w) pos += waveTot[w][k];
        for (uint32_t b2 = 0; b2 < b; ++b2) pos += c[b2] == k ? 1u : 0u;
        out[pos] = in ? in[j0 + b] : j0 + b;
    }
}

template <uint32_t K>
static int partition_impl(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t* out,
                          uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes, hipStream_t s) {
    const uint32_t numTiles = ceil_div(n, kTile);
    const uint32_t m = K * numTiles;
    const uint32_t numSeg = ceil_div(m, kScanSeg);
    WRK_REQUIRE(((uint64_t)m + numSeg) * 4u <= scratchBytes, "wrk_partition: scratch too small");
    if (numTiles == 0) {
        hipLaunchKernelGGL(k_zero_offsets, dim3(1), dim3(64), 0, s, offsetsOut, K + 1);
        WRK_CHECK_LAUNCH();
        return WRK_OK;
    }
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    uint32_t* segTotals = counts + m;
    if ((reinterpret_cast<uintptr_t>(category) & 3u) == 0)
        hipLaunchKernelGGL(k_part_count_vec<K>, dim3(numTiles), dim3(256), 0, s, category, n, numTiles, counts);
    else
        hipLaunchKernelGGL(k_part_count<K>, dim3(numTiles), dim3(256), 0, s, category, n, numTiles, counts);
    if (numSeg == 1) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, counts, m, numTiles, K, offsetsOut);
    } else {
        hipLaunchKernelGGL(k_scan_segments, dim3(numSeg), dim3(1024), 0, s, counts, m, segTotals);
        hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, s, segTotals, numSeg, K, offsetsOut);
        hipLaunchKernelGGL(k_scan_apply, dim3(numSeg), dim3(1024), 0, s, counts, m, segTotals, numTiles, K,
                           offsetsOut);
    }
    if ((reinterpret_cast<uintptr_t>(category) & 3u) == 0)
        hipLaunchKernelGGL(k_part_scatter_vec<K>, dim3(numTiles), dim3(256), 0, s, category, in, n, numTiles, counts,
                           out);
    else
        hipLaunchKernelGGL(k_part_scatter<K>, dim3(numTiles), dim3(256), 0, s, category, in, n, numTiles, counts,
                           out);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

// Scratch bytes of a partition of n items into k categories: the k x tiles
// count matrix, one total per scan segment, and a small tail.
constexpr uint64_t partition_scratch_bytes(uint64_t n, uint32_t k) {
    const uint64_t m = (n + kTile - 1) / kTile * (k < 1 ? 1 : k);
    return (m + (m + kScanSeg - 1) / kScanSeg + 64) * 4;
}

static int partition_dispatch(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t numCategories,
                              uint32_t* out, uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes,
                              hipStream_t s) {
    switch (numCategories) {
        case 1: return partition_impl<1>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 2: return partition_impl<2>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 3: return partition_impl<3>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 4: return partition_impl<4>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 5: return partition_impl<5>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 6: return partition_impl<6>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 7: return partition_impl<7>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 8: return partition_impl<8>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 9: return partition_impl<9>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 10: return partition_impl<10>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 11: return partition_impl<11>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 12: return partition_impl<12>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 13: return partition_impl<13>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 14: return partition_impl<14>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 15: return partition_impl<15>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        case 16: return partition_impl<16>(category, in, n, out, offsetsOut, scratch, scratchBytes, s);
        default: set_error("wrk_partition: numCategories must be in [1, 16]"); return WRK_ERR_INVALID;
    }
}

// ------------------------------------------------ allocation-free variant
// The count matrix of wrk_partition_noalloc lives in this array, sized for
// the largest queue that entry point accepts (2^26 items, 16 categories:
// 4,194,816 bytes). Each GPU holds its own copy for the life of the process.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): it uses the per-device address from
// hipGetSymbolAddress.
constexpr uint32_t kNoAllocMaxItems = 1u << 26;
constexpr uint32_t kNoAllocMaxCategories = 16;
constexpr uint64_t kNoAllocScratchBytes =
    (partition_scratch_bytes(kNoAllocMaxItems, kNoAllocMaxCategories) + 255) & ~uint64_t(255);
namespace noalloc {
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
}

constexpr int kMaxDevices = 64;
static std::mutex g_noallocMutex;
static void* g_noallocScratch[kMaxDevices];  // device address of noalloc::scratch per device, once resolved

// Device address of the array on the current device, resolved on the first
// call made with that device current.
static void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_noallocMutex);
    if (!g_noallocScratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc::scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("wrk_partition_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_noallocScratch[device] = p;
    }
    return g_noallocScratch[device];
}

}  // namespace wrk

using namespace wrk;

extern "C" uint64_t wrk_partition_scratch_bytes(uint32_t n, uint32_t numCategories) {
    return partition_scratch_bytes(n, numCategories);
}

extern "C" int wrk_partition(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t numCategories,
                             uint32_t* out, uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes,
                             void* stream) {
    WRK_REQUIRE(offsetsOut && scratch, "wrk_partition: null argument");
    WRK_REQUIRE(n == 0 || (category && out), "wrk_partition: null argument");
    return partition_dispatch(category, in, n, numCategories, out, offsetsOut, scratch, scratchBytes,
                              as_stream(stream));
}

extern "C" uint64_t wrk_exclusive_scan_scratch_bytes(uint32_t n) { return exclusive_scan_scratch_bytes(n); }

extern "C" int wrk_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* scratch,
                                      uint64_t scratchBytes, void* stream) {
    WRK_REQUIRE(n == 0 || (in && out && scratch), "wrk_exclusive_scan_u32: null argument");
    return exclusive_scan_u32(in, out, n, static_cast<uint32_t*>(scratch), scratchBytes, total, as_stream(stream));
}

extern "C" uint32_t wrk_partition_noalloc_max_items(void) { return kNoAllocMaxItems; }

extern "C" int wrk_partition_noalloc(const uint8_t* category, const uint32_t* in, uint32_t n,
                                     uint32_t numCategories, uint32_t* out, uint32_t* offsetsOut, void* stream) {
    WRK_REQUIRE(offsetsOut, "wrk_partition_noalloc: null argument");
    WRK_REQUIRE(n == 0 || (category && out), "wrk_partition_noalloc: null argument");
    if (n > kNoAllocMaxItems) {
        set_error("wrk_partition_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                  std::to_string(kNoAllocMaxItems) + " items (use wrk_partition with a scratch buffer)");
        return WRK_ERR_INVALID;
    }
    WRK_REQUIRE(numCategories >= 1 && numCategories <= kNoAllocMaxCategories,
                "wrk_partition_noalloc: numCategories must be in [1, 16]");
    int device = -1;
    WRK_HIP_TRY(hipGetDevice(&device));
    WRK_REQUIRE(device >= 0 && device < kMaxDevices, "wrk_partition_noalloc: device index out of range");
    void* scratch = noalloc_scratch(device);
    if (!scratch) return WRK_ERR_HIP;
    return partition_dispatch(category, in, n, numCategories, out, offsetsOut, scratch, kNoAllocScratchBytes,
                              as_stream(stream));
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
    uint64_t countsBytes = ((uint64_t)numTiles + ceil_div(numTiles, kScanSeg) + 64) * 4;
    uint64_t catOff = (countsBytes + 255) & ~uint64_t(255);
    WRK_REQUIRE(scratchBytes >= catOff + n, "wrk_compact: scratch too small (need wrk_compact_scratch_bytes)");
    hipStream_t s = as_stream(stream);
    uint8_t* cat = static_cast<uint8_t*>(scratch) + catOff;
    if (n > 0) hipLaunchKernelGGL(k_keep_to_cat, dim3(ceil_div(n, 256)), dim3(256), 0, s, keep, n, cat);
    // offsets for K=1: [0, total] -> total lands in countOut via a 2-word tail
    uint32_t* offs = static_cast<uint32_t*>(scratch) + numTiles + ceil_div(numTiles, kScanSeg) + 8;
    int rc = partition_impl<1>(cat, in, n, out, offs, scratch, countsBytes, s);
    if (rc != WRK_OK) return rc;
    WRK_HIP_TRY(hipMemcpyAsync(countOut, offs + 1, 4, hipMemcpyDeviceToDevice, s));
    return WRK_OK;
}

extern "C" uint64_t wrk_compact_scratch_bytes(uint32_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t countsBytes = (tiles + (tiles + kScanSeg - 1) / kScanSeg + 64) * 4;
    return ((countsBytes + 255) & ~uint64_t(255)) + n;
}
