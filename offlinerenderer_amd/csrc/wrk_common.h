// wrk_common.h — error plumbing shared by the C-ABI entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/wrk.h"

namespace wrk {

void set_error(const std::string& msg);

#define WRK_HIP_TRY(expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::wrk::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return WRK_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

#define WRK_CHECK_LAUNCH()                                                                  \
    do {                                                                                    \
        hipError_t _e = hipGetLastError();                                                  \
        if (_e != hipSuccess) {                                                             \
            ::wrk::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));       \
            return WRK_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

#define WRK_REQUIRE(cond, msg)                                                              \
    do {                                                                                    \
        if (!(cond)) {                                                                      \
            ::wrk::set_error(msg);                                                          \
            return WRK_ERR_INVALID;                                                         \
        }                                                                                   \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Device-wide exclusive scan of m uint32 values (partition.hip): out[i] =
// sum of in[0 .. i); in == out is allowed; *total (device, optional) gets
// the sum of all m values.
uint64_t exclusive_scan_scratch_bytes(uint32_t m);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t m, uint32_t* scratch, uint64_t scratchBytes,
                       uint32_t* total, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs by key bits [beginBit,
// endBit) (sort.hip). The sorted pairs end in keys/vals; keysTmp/valsTmp
// are ping-pong buffers of n entries.
uint64_t radix_sort_scratch_bytes(uint32_t n);
int radix_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n,
                         uint32_t beginBit, uint32_t endBit, void* scratch, uint64_t scratchBytes, hipStream_t s);

// Compressed 8-wide node (96 bytes): child boxes quantised to 8 bits per
// plane on a per-axis power-of-two grid anchored at `origin` (decoded as
// origin + q * 2^e, rounded outward at build time so the decoded box always
// contains the exact one). origin.w holds the three biased float exponents
// (bits 0-7 x, 8-15 y, 16-23 z).
struct alignas(16) Node8 {
    float4 origin;
    uint32_t child[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
};
static_assert(sizeof(Node8) == 96, "Node8 must be 96 bytes");

constexpr uint32_t kBvhWidth = 8;

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kPrimMask = 0x0FFFFFFFu;  // primitive id bits of a hit (ids < 2^27)

// Shade category of a triangle (the queue its hits go to): the material kind
// 0..4, an unknown kind joins the emissive queue (whose shader flushes it as
// the default case of shade.cu:79-81 does), and an out-of-range material id
// ends the path (shade.cu:35-38).
constexpr uint32_t kCatDone = 5;
__host__ __device__ inline uint32_t shade_category(uint32_t materialId, uint32_t numMaterials,
                                                   const wrk_material* materials) {
    if (materialId >= numMaterials) return kCatDone;
    const uint32_t type = materials[materialId].type;
    return type <= WRK_MAT_GGX_TRANSMISSION ? type : (uint32_t)WRK_MAT_EMISSIVE;
}
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
#ifndef WRK_LEAF_MAX
#define WRK_LEAF_MAX 2
#endif
// triangles per leaf (at most 16: 4-bit count). With 8-wide nodes and
// batched leaf tests, 2 measured best: leaves of 1, 2, 3, 4, 6, 8 gave
// material 1436, 1486, 1457, 1434, 1355, 1285 Mpaths/s (leaf-batch
// threshold 5/8 for all six).
constexpr uint32_t kLeafMax = WRK_LEAF_MAX;
constexpr float kBoxPadRel = 9.5367431640625e-07f;  // 2^-20 outward box padding
#ifndef WRK_STACK_LDS
#define WRK_STACK_LDS 8
#endif
// With lanes refilled in batches of 16, caps of 12, 16 and 24 rays per lane
// gave material 2059-2063, 2067-2074 and 2073-2078 Mpaths/s, large within noise.
#ifndef WRK_RAYS_PER_LANE
#define WRK_RAYS_PER_LANE 24
#endif
constexpr uint32_t kStackLds = WRK_STACK_LDS;        // traversal stack entries per lane in LDS (power of 2)
constexpr uint32_t kStackSpill = 192 - WRK_STACK_LDS; // further entries per lane in scratch
constexpr uint32_t kRaysPerLane = WRK_RAYS_PER_LANE;  // upper bound of rays per lane in a wave's fetch chunk
constexpr uint32_t kStackTotal = kStackLds + kStackSpill;

}  // namespace wrk
