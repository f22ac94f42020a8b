// bvh.hip — GPU LBVH build (63-bit Morton codes, LDS-staged radix sort,
// agglomerative binary tree, bottom-up boxes and SAH cost rows, SAH-optimal
// collapse to compressed 8-wide nodes) and watertight FP32 closest-hit /
// any-hit traversal for gfx950.
//
// Replaces the OptiX GAS build (libWR/src/internal/scene.cpp:493-550) and
// optixTrace (libWR/kernels/trace.cu:27-51,127-151).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "traverse.h"
#include "wrk_common.h"
#include "wrk_device.h"

namespace wrk {

// ----------------------------------------------------------------- build
struct alignas(16) BinNode {
    float4 lo;   // xyz box, w = left child ref bits
    float4 hi;   // xyz box, w = right child ref bits
    uint32_t first, last, pad0, pad1;
};

enum Counter : uint32_t {
    C_NODES = 0,      // 4-wide nodes allocated
    C_TASKS_NEXT = 1, // tasks produced for the next level
    C_MAXDEPTH = 2,
    C_BIN_NODES = 3,  // binary nodes created by the agglomerative build
    C_CMIN = 4,       // ordered-uint centroid min (3)
    C_CMAX = 8,       // ordered-uint centroid max (3)
    C_BMIN = 12,      // ordered-uint scene min (3)
    C_BMAX = 16,      // ordered-uint scene max (3)
    C_COUNT = 20
};

__device__ __forceinline__ uint32_t ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __host__ __forceinline__ float unordered(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__global__ void k_init_counters(uint32_t* c) {
    uint32_t i = threadIdx.x;
    if (i < C_COUNT) {
        uint32_t v = 0;
        if ((i >= C_CMIN && i < C_CMIN + 3) || (i >= C_BMIN && i < C_BMIN + 3)) v = 0xFFFFFFFFu;
        c[i] = v;
    }
}

// Per-triangle boxes; each block writes its partial scene and centroid
// bounds (12 floats), reduced by k_reduce_bounds.
__global__ __launch_bounds__(256) void k_tri_bounds(const float* __restrict__ verts, const uint32_t* __restrict__ idx,
                                                    uint32_t n, float4* triLo, float4* triHi, float* partials) {
    __shared__ float red[4][12];
    uint32_t t = blockIdx.x * 256 + threadIdx.x;
    float v[12];
#pragma unroll
    for (int a = 0; a < 3; ++a) { v[a] = INFINITY; v[3 + a] = -INFINITY; v[6 + a] = INFINITY; v[9 + a] = -INFINITY; }
    if (t < n) {
        uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
        float lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float x0 = verts[i0 * 3 + a], x1 = verts[i1 * 3 + a], x2 = verts[i2 * 3 + a];
            lo[a] = fminf(x0, fminf(x1, x2));
            hi[a] = fmaxf(x0, fmaxf(x1, x2));
            float c = (lo[a] + hi[a]) * 0.5f;
            v[a] = c; v[3 + a] = c; v[6 + a] = lo[a]; v[9 + a] = hi[a];
        }
        triLo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        triHi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], __shfl_xor(v[a], off));
            v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], off));
            v[6 + a] = fminf(v[6 + a], __shfl_xor(v[6 + a], off));
            v[9 + a] = fmaxf(v[9 + a], __shfl_xor(v[9 + a], off));
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) red[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        uint32_t k = threadIdx.x;
        bool isMin = (k % 6) < 3;
        float r = red[0][k];
        for (int w = 1; w < 4; ++w) r = isMin ? fminf(r, red[w][k]) : fmaxf(r, red[w][k]);
        partials[(size_t)blockIdx.x * 12 + k] = r;
    }
}

// One block folds the per-block partials into the counters record
// (ordered-uint min/max of centroids and of the scene box).
__global__ __launch_bounds__(1024) void k_reduce_bounds(const float* __restrict__ partials, uint32_t numBlocks,
                                                        uint32_t* counters) {
    __shared__ float red[16][12];
    float v[12];
#pragma unroll
    for (int a = 0; a < 3; ++a) { v[a] = INFINITY; v[3 + a] = -INFINITY; v[6 + a] = INFINITY; v[9 + a] = -INFINITY; }
    for (uint32_t b = threadIdx.x; b < numBlocks; b += 1024) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            float x = partials[(size_t)b * 12 + k];
            v[k] = ((k % 6) < 3) ? fminf(v[k], x) : fmaxf(v[k], x);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            float x = __shfl_xor(v[k], off);
            v[k] = ((k % 6) < 3) ? fminf(v[k], x) : fmaxf(v[k], x);
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) red[threadIdx.x >> 6][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        uint32_t k = threadIdx.x;
        bool isMin = (k % 6) < 3;
        float r = red[0][k];
        for (int w = 1; w < 16; ++w) r = isMin ? fminf(r, red[w][k]) : fmaxf(r, red[w][k]);
        // k: 0-2 centroid min, 3-5 centroid max, 6-8 box min, 9-11 box max
        uint32_t slot = (k < 3) ? C_CMIN + k : (k < 6) ? C_CMAX + (k - 3) : (k < 9) ? C_BMIN + (k - 6) : C_BMAX + (k - 9);
        counters[slot] = ordered(r);
    }
}

typedef uint64_t MortonKey;
constexpr unsigned kMortonBits = 63;

// 21 bits per axis -> every third bit of a 63-bit code
__device__ __forceinline__ uint64_t expand_bits21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x001F00000000FFFFull;
    v = (v | (v << 16)) & 0x001F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Morton code of the triangle centroid in the centroid bounds: 63 bits (21
// per axis) so that neighbouring small triangles of a dense mesh get distinct
// codes and the hierarchy splits them by position rather than by index.
__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ triLo, const float4* __restrict__ triHi,
                                                uint32_t n, const uint32_t* __restrict__ counters,
                                                MortonKey* keys, uint32_t* vals) {
    uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    float4 lo = triLo[t], hi = triHi[t];
    float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float mn = unordered(counters[C_CMIN + a]);
        float mx = unordered(counters[C_CMAX + a]);
        float ext = mx - mn;
        float u = ext > 0.0f ? (c[a] - mn) / ext : 0.0f;
        float s = fminf(fmaxf(u * 2097152.0f, 0.0f), 2097151.0f);
        q[a] = (uint32_t)s;
    }
    keys[t] = (expand_bits21(q[0]) << 2) | (expand_bits21(q[1]) << 1) | expand_bits21(q[2]);
    vals[t] = t;
}

__device__ __forceinline__ void child_box(const BinNode* bnodes, const float4* triLo, const float4* triHi,
                                          const uint32_t* vals, uint32_t ref, float lo[3], float hi[3]) {
    if (ref & kLeafBit) {
        uint32_t t = vals[ref & ~kLeafBit];
        float4 a = triLo[t], b = triHi[t];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    } else {
        float4 a = bnodes[ref].lo, b = bnodes[ref].hi;
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    }
}

// ---- agglomerative binary build (PLOC: Meister and Bittner 2018, "Parallel
// locally-ordered clustering for bounding volume hierarchy construction").
// Clusters start as the triangles in Morton order; each round every cluster
// finds the neighbour within WRK_PLOC_RADIUS positions whose merged box has
// the least surface area, mutual nearest neighbours merge into a new binary
// node, and the survivors are compacted in order. The result is a binary
// tree of SAH quality close to a top-down sweep, unlike a Morton-split
// (Karras 2012) hierarchy whose splits ignore triangle sizes.
#ifndef WRK_PLOC_RADIUS
#define WRK_PLOC_RADIUS 8
#endif

__global__ __launch_bounds__(256) void k_ploc_init(const float4* __restrict__ triLo, const float4* __restrict__ triHi,
                                                   const uint32_t* __restrict__ vals, uint32_t n, uint32_t* cid,
                                                   float4* clo, float4* chi) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t t = vals[i];
    cid[i] = kLeafBit | i;  // leaf reference: position in Morton order
    clo[i] = triLo[t];
    chi[i] = triHi[t];
}

__device__ __forceinline__ float merged_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;
}

// Nearest neighbour of every cluster within the radius. Pairs are ordered
// strictly by (merged area, not an aligned pair {2k, 2k+1}, lower index,
// higher index), so the least pair is always mutual (at least one merge per
// round), and equal areas (a regular grid, duplicated triangles) resolve to
// aligned pairs that are all mutual, so such inputs still halve each round.
__global__ __launch_bounds__(256) void k_ploc_nn(const float4* __restrict__ clo, const float4* __restrict__ chi,
                                                 uint32_t c, uint32_t* nn) {
    constexpr int R = WRK_PLOC_RADIUS;
    __shared__ float4 sLo[256 + 2 * R], sHi[256 + 2 * R];
    const int base = (int)(blockIdx.x * 256) - R;
    for (int k = threadIdx.x; k < 256 + 2 * R; k += 256) {
        int g = base + k;
        if (g >= 0 && g < (int)c) { sLo[k] = clo[g]; sHi[k] = chi[g]; }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= (int)c) return;
    const int li = (int)threadIdx.x + R;
    const float4 alo = sLo[li], ahi = sHi[li];
    // (c >= 2 here; a NaN area, from invalid vertices, counts as +inf, so
    // every cluster still names a neighbour and the rounds make progress)
    float best = INFINITY;
    int bj = (i ^ 1) < (int)c ? (i ^ 1) : i - 1;
    const int j0 = max(i - R, 0), j1 = min(i + R, (int)c - 1);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        float a = merged_area(alo, ahi, sLo[j - base], sHi[j - base]);
        if (a != a) a = INFINITY;
        // ascending j: among equal areas the first candidate is the smallest
        // (min, max) pair, so only the aligned partner i ^ 1 can displace it
        if (a < best || (a == best && j == (i ^ 1))) { best = a; bj = j; }
    }
    nn[i] = (uint32_t)bj;
}

// Mutual nearest neighbours i < j merge into a new binary node kept at i;
// j leaves. keep[i] = 1 for survivors.
__global__ __launch_bounds__(256) void k_ploc_merge(uint32_t c, const uint32_t* __restrict__ nn, uint32_t* cid,
                                                    float4* clo, float4* chi, BinNode* bnodes, uint32_t* counters,
                                                    uint32_t* keep) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c) return;
    const uint32_t j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    if (mutual && i > j) { keep[i] = 0; return; }
    keep[i] = 1;
    if (!mutual) return;
    const float4 alo = clo[i], ahi = chi[i], blo = clo[j], bhi = chi[j];
    const uint32_t k = atomicAdd(&counters[C_BIN_NODES], 1u);
    const float4 lo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.0f);
    const float4 hi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.0f);
    BinNode& b = bnodes[k];
    b.lo = make_float4(lo.x, lo.y, lo.z, __uint_as_float(cid[i]));
    b.hi = make_float4(hi.x, hi.y, hi.z, __uint_as_float(cid[j]));
    cid[i] = k;
    clo[i] = lo;
    chi[i] = hi;
}

__global__ __launch_bounds__(256) void k_ploc_compact(uint32_t c, const uint32_t* __restrict__ keep,
                                                      const uint32_t* __restrict__ pos, const uint32_t* __restrict__ cidIn,
                                                      const float4* __restrict__ loIn, const float4* __restrict__ hiIn,
                                                      uint32_t* cidOut, float4* loOut, float4* hiOut,
                                                      uint32_t* counters) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c) return;
    if (i == c - 1) counters[C_TASKS_NEXT] = pos[i] + keep[i];
    if (!keep[i]) return;
    const uint32_t d = pos[i];
    cidOut[d] = cidIn[i];
    loOut[d] = loIn[i];
    hiOut[d] = hiIn[i];
}

// Subtree triangle counts, bottom-up by level (held in `last` until the
// ranges are assigned).
__global__ __launch_bounds__(256) void k_ploc_sizes(const uint32_t* __restrict__ level, uint32_t count,
                                                    BinNode* bnodes) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    BinNode& b = bnodes[level[i]];
    const uint32_t l = __float_as_uint(b.lo.w), r = __float_as_uint(b.hi.w);
    b.last = ((l & kLeafBit) ? 1u : bnodes[l].last) + ((r & kLeafBit) ? 1u : bnodes[r].last);
}

// Top-down by level: each node's subtree gets the contiguous triangle range
// [first, first + size); leaves are renumbered to their new positions and
// valsOut lists the triangles in that order, so every subtree of the binary
// tree is one contiguous range, as the collapse and the leaves require.
__global__ __launch_bounds__(256) void k_ploc_ranges(const uint32_t* __restrict__ level, uint32_t count,
                                                     bool rootLevel, BinNode* bnodes,
                                                     const uint32_t* __restrict__ valsIn, uint32_t* valsOut) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    BinNode& b = bnodes[level[i]];
    const uint32_t first = rootLevel ? 0u : b.first;
    const uint32_t size = b.last;
    uint32_t l = __float_as_uint(b.lo.w), r = __float_as_uint(b.hi.w);
    uint32_t sizeL;
    if (l & kLeafBit) {
        valsOut[first] = valsIn[l & ~kLeafBit];
        l = kLeafBit | first;
        sizeL = 1;
    } else {
        bnodes[l].first = first;
        sizeL = bnodes[l].last;
    }
    if (r & kLeafBit) {
        valsOut[first + sizeL] = valsIn[r & ~kLeafBit];
        r = kLeafBit | (first + sizeL);
    } else {
        bnodes[r].first = first + sizeL;
    }
    b.lo.w = __uint_as_float(l);
    b.hi.w = __uint_as_float(r);
    b.first = first;
    b.last = first + size - 1u;
}

// Top-down: the internal children of one level's nodes form the next level
// (one append per wave).
__global__ __launch_bounds__(256) void k_expand_level(const uint32_t* __restrict__ level, uint32_t count,
                                                      const BinNode* __restrict__ bnodes, uint32_t* next,
                                                      uint32_t* nextCount) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t k0 = kEmptyChild, k1 = kEmptyChild;
    uint32_t num = 0;
    if (i < count) {
        const BinNode& b = bnodes[level[i]];
        uint32_t l = __float_as_uint(b.lo.w), r = __float_as_uint(b.hi.w);
        if (!(l & kLeafBit)) { k0 = l; num = 1; }
        if (!(r & kLeafBit)) { if (num) k1 = r; else k0 = r; ++num; }
    }
    uint32_t incl = num;
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    uint32_t total = __shfl(incl, 63);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(nextCount, total);
    base = __shfl(base, 0);
    uint32_t pos = base + incl - num;
    if (num > 0) next[pos] = k0;
    if (num > 1) next[pos + 1] = k1;
}

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// SAH costs for the collapse to W-wide nodes (Ylitie, Karras, Laine 2017,
// "Efficient incoherent ray traversal on GPUs through compressed wide
// BVHs", section 4.1). For a binary node n and i = 1..W, C(n, i) is the
// least cost of representing n's subtree as at most i children of a wide
// node: C(n, 1) = min(leaf cost, A(n) * c_node + best split of W slots
// between n's two children), C(n, i > 1) = min(C(n, i - 1), best split of
// i slots). A is the surface area. The collapse follows the argmins top-down,
// recomputing the same float sums, so its choices match the table exactly.
#ifndef WRK_SAH_CPRIM
#define WRK_SAH_CPRIM 0.3f
#endif
constexpr float kSahNode = 1.0f;

__device__ __forceinline__ float sah_leaf_cost(float area, uint32_t count) {
    return count <= kLeafMax ? area * (WRK_SAH_CPRIM * (float)count) : INFINITY;
}

// cost row of a child reference: a single triangle costs its leaf cost for
// every i; an inner binary node has its row in `cost`
template <int W>
__device__ __forceinline__ void sah_row(const float* cost, const float4* triLo, const float4* triHi,
                                        const uint32_t* vals, uint32_t ref, float row[W]) {
    if (ref & kLeafBit) {
        uint32_t t = vals[ref & ~kLeafBit];
        float4 a = triLo[t], b = triHi[t];
        float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
        float c = sah_leaf_cost(half_area(lo, hi), 1u);
#pragma unroll
        for (int i = 0; i < W; ++i) row[i] = c;
    } else {
#pragma unroll
        for (int i = 0; i < W; ++i) row[i] = cost[(size_t)ref * W + i];
    }
}

// best split of i slots (i >= 2) between two children: min over k of
// cl(k) + cr(i - k); *kBest gets the smallest minimising k
template <int W>
__device__ __forceinline__ float sah_split(const float cl[W], const float cr[W], int i, int* kBest) {
    float best = INFINITY;
    int kb = 1;
    for (int k = 1; k < i; ++k) {
        float c = cl[k - 1] + cr[i - k - 1];
        if (c < best) { best = c; kb = k; }
    }
    *kBest = kb;
    return best;
}

// Bottom-up: boxes (and SAH cost rows) of one level from its children, whose
// boxes (deeper levels) were written by earlier launches.
template <int W>
__global__ __launch_bounds__(256) void k_refit_level(const uint32_t* __restrict__ level, uint32_t count,
                                                     BinNode* bnodes, const float4* __restrict__ triLo,
                                                     const float4* __restrict__ triHi,
                                                     const uint32_t* __restrict__ vals, float* cost) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t node = level[i];
    BinNode* b = &bnodes[node];
    uint32_t l = __float_as_uint(b->lo.w), r = __float_as_uint(b->hi.w);
    float llo[3], lhi[3], rlo[3], rhi[3];
    child_box(bnodes, triLo, triHi, vals, l, llo, lhi);
    child_box(bnodes, triLo, triHi, vals, r, rlo, rhi);
    float lo[3] = {fminf(llo[0], rlo[0]), fminf(llo[1], rlo[1]), fminf(llo[2], rlo[2])};
    float hi[3] = {fmaxf(lhi[0], rhi[0]), fmaxf(lhi[1], rhi[1]), fmaxf(lhi[2], rhi[2])};
    b->lo = make_float4(lo[0], lo[1], lo[2], __uint_as_float(l));
    b->hi = make_float4(hi[0], hi[1], hi[2], __uint_as_float(r));
    float cl[W], cr[W], row[W];
    sah_row<W>(cost, triLo, triHi, vals, l, cl);
    sah_row<W>(cost, triLo, triHi, vals, r, cr);
    const float area = half_area(lo, hi);
    int k;
    const float inner = sah_split<W>(cl, cr, W, &k) + area * kSahNode;
    row[0] = fminf(sah_leaf_cost(area, b->last - b->first + 1u), inner);
    for (int j = 2; j <= W; ++j) row[j - 1] = fminf(row[j - 2], sah_split<W>(cl, cr, j, &k));
#pragma unroll
    for (int j = 0; j < W; ++j) cost[(size_t)node * W + j] = row[j];
}

// Collapse one binary subtree root into a W-wide node following the SAH
// cost rows, and write it as a compressed Node8.
template <int W>
__global__ __launch_bounds__(256) void k_collapse(const uint4* __restrict__ tasksIn, uint32_t numTasks,
                                                  const BinNode* __restrict__ bnodes, const float4* __restrict__ triLo,
                                                  const float4* __restrict__ triHi, const uint32_t* __restrict__ vals,
                                                  const float* __restrict__ cost, void* nodesOut, uint32_t* counters,
                                                  uint4* tasksOut) {
    uint32_t ti = blockIdx.x * 256 + threadIdx.x;
    if (ti >= numTasks) return;
    uint4 task = tasksIn[ti];
    float sceneExtent = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        sceneExtent = fmaxf(sceneExtent, unordered(counters[C_BMAX + a]) - unordered(counters[C_BMIN + a]));
    static_assert(W == 8, "the node format is the compressed 8-wide Node8");
    uint32_t ref[W];
    uint32_t m = 0;
    if (task.x & kLeafBit) {  // single-triangle scene
        ref[m++] = task.x;
    } else {
        // follow the SAH table: split W slots between the root's children,
        // then each (subtree, slots) pair either into one child (a leaf or
        // an inner node, as its cost row decided) or on into its children
        uint32_t sr[2 * W], sj[2 * W];
        int top = 0;
        {
            const uint32_t l = __float_as_uint(bnodes[task.x].lo.w), r = __float_as_uint(bnodes[task.x].hi.w);
            float cl[W], cr[W];
            sah_row<W>(cost, triLo, triHi, vals, l, cl);
            sah_row<W>(cost, triLo, triHi, vals, r, cr);
            int k;
            (void)sah_split<W>(cl, cr, W, &k);
            sr[top] = r; sj[top++] = (uint32_t)(W - k);
            sr[top] = l; sj[top++] = (uint32_t)k;
        }
        while (top > 0) {
            --top;
            const uint32_t r = sr[top], j = sj[top];
            if (r & kLeafBit) { ref[m++] = r; continue; }
            float row[W];
            sah_row<W>(cost, triLo, triHi, vals, r, row);
            uint32_t jj = 1;  // fewest slots that reach C(r, j)
            while (jj < j && row[jj - 1] != row[j - 1]) ++jj;
            if (jj == 1) { ref[m++] = r; continue; }
            const uint32_t cl_ = __float_as_uint(bnodes[r].lo.w), cr_ = __float_as_uint(bnodes[r].hi.w);
            float cl[W], cr[W];
            sah_row<W>(cost, triLo, triHi, vals, cl_, cl);
            sah_row<W>(cost, triLo, triHi, vals, cr_, cr);
            int k;
            (void)sah_split<W>(cl, cr, (int)jj, &k);
            sr[top] = cr_; sj[top++] = jj - (uint32_t)k;
            sr[top] = cl_; sj[top++] = (uint32_t)k;
        }
    }
    float blo[W][3], bhi[W][3];
    uint32_t child[W];
    for (uint32_t c = 0; c < (uint32_t)W; ++c) {
        if (c >= m) {
            blo[c][0] = blo[c][1] = blo[c][2] = INFINITY;
            bhi[c][0] = bhi[c][1] = bhi[c][2] = -INFINITY;
            child[c] = kEmptyChild;
            continue;
        }
        uint32_t r = ref[c];
        float lo[3], hi[3];
        if (r & kLeafBit) {
            uint32_t t = vals[r & ~kLeafBit];
            float4 a = triLo[t], b = triHi[t];
            lo[0] = a.x; lo[1] = a.y; lo[2] = a.z; hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
            child[c] = kLeafBit | (r & ~kLeafBit);  // count-1 = 0
        } else {
            const BinNode& b = bnodes[r];
            lo[0] = b.lo.x; lo[1] = b.lo.y; lo[2] = b.lo.z; hi[0] = b.hi.x; hi[1] = b.hi.y; hi[2] = b.hi.z;
            uint32_t cnt = b.last - b.first + 1u;
            bool asLeaf = cnt <= kLeafMax;
            if (asLeaf) {
                // a leaf where its cost row chose the leaf cost
                float row[W];
                sah_row<W>(cost, triLo, triHi, vals, r, row);
                const float blo[3] = {b.lo.x, b.lo.y, b.lo.z}, bhi[3] = {b.hi.x, b.hi.y, b.hi.z};
                asLeaf = row[0] == sah_leaf_cost(half_area(blo, bhi), cnt);
            }
            if (asLeaf) {
                child[c] = kLeafBit | ((cnt - 1u) << 27) | b.first;
            } else {
                uint32_t slot = atomicAdd(&counters[C_NODES], 1u);
                uint32_t qi = atomicAdd(&counters[C_TASKS_NEXT], 1u);
                tasksOut[qi] = make_uint4(r, slot, task.z + 1u, 0u);
                child[c] = slot;
            }
        }
        // Pad outward so the float slab test stays conservative with
        // respect to the watertight triangle test, whose rounding can
        // report hits on an edge or vertex an ulp outside the exact bounds.
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pad = (fmaxf(fabsf(lo[a]), fabsf(hi[a])) + sceneExtent) * kBoxPadRel;
            blo[c][a] = lo[a] - pad;
            bhi[c][a] = hi[a] + pad;
        }
    }
    {
        // quantise: per axis the smallest power-of-two step with which 255
        // steps cover the union, then each plane rounded outward and checked
        // with the decoding expression the traversal uses, fma(q, s, origin)
        // (q * s is exact, so the fused and unfused forms agree)
        Node8 out;
        float org[3];
        uint32_t ebits[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float ulo = INFINITY, uhi = -INFINITY;
            for (uint32_t c = 0; c < m; ++c) {
                ulo = fminf(ulo, blo[c][a]);
                uhi = fmaxf(uhi, bhi[c][a]);
            }
            org[a] = ulo;
            int e = -126;
            const float ext = uhi - ulo;
            if (ext > 0.0f) {
                int ex;
                (void)frexpf(ext / 255.0f, &ex);  // ext / 255 < 2^ex
                e = ex < -126 ? -126 : ex;
            }
            for (int attempt = 0; attempt < 8; ++attempt) {
                const float sc = __uint_as_float((uint32_t)(e + 127) << 23);
                bool ok = true;
                for (uint32_t c = 0; c < (uint32_t)W; ++c) {
                    if (c >= m) {
                        out.qlo[a][c] = 255;
                        out.qhi[a][c] = 0;
                        continue;
                    }
                    float ql = floorf((blo[c][a] - ulo) / sc);
                    ql = fminf(fmaxf(ql, 0.0f), 255.0f);
                    while (ql > 0.0f && fmaf(ql, sc, ulo) > blo[c][a]) ql -= 1.0f;
                    float qh = ceilf((bhi[c][a] - ulo) / sc);
                    qh = fmaxf(qh, ql);
                    while (qh <= 255.0f && fmaf(qh, sc, ulo) < bhi[c][a]) qh += 1.0f;
                    if (qh > 255.0f) { ok = false; break; }
                    out.qlo[a][c] = (uint8_t)ql;
                    out.qhi[a][c] = (uint8_t)qh;
                }
                if (ok || e >= 127) break;
                ++e;
            }
            ebits[a] = (uint32_t)(e + 127);
        }
        out.origin = make_float4(org[0], org[1], org[2], __uint_as_float(ebits[0] | (ebits[1] << 8) | (ebits[2] << 16)));
#pragma unroll
        for (int c = 0; c < W; ++c) out.child[c] = child[c];
        static_cast<Node8*>(nodesOut)[task.y] = out;
    }
    atomicMax(&counters[C_MAXDEPTH], task.z);
}

__global__ __launch_bounds__(256) void k_reorder_tris(const uint32_t* __restrict__ vals, const float* __restrict__ verts,
                                                      const uint32_t* __restrict__ idx, uint32_t n, float4* tris) {
    uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    uint32_t t = vals[k];
    uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    tris[k * 3 + 0] = make_float4(verts[i0 * 3 + 0], verts[i0 * 3 + 1], verts[i0 * 3 + 2], __uint_as_float(t));
    tris[k * 3 + 1] = make_float4(verts[i1 * 3 + 0], verts[i1 * 3 + 1], verts[i1 * 3 + 2], 0.0f);
    tris[k * 3 + 2] = make_float4(verts[i2 * 3 + 0], verts[i2 * 3 + 1], verts[i2 * 3 + 2], 0.0f);
}

// ------------------------------------------------------------- traversal
// One lane per ray with per-wave dynamic ray fetch: each wave owns a chunk
// of raysPerLane*64 queue entries and a lane whose ray finishes takes the
// next entry of the chunk (ballot + popcount, no atomics), so lanes do not
// idle until the slowest ray of a static 64-ray group is done. raysPerLane
// is chosen per launch so the grid still fills every wave slot of the chip. The
// traversal stack keeps its top kStackLds entries in LDS (a per-lane ring,
// column-major so a wave touches distinct banks) and spills older entries
// to a per-lane scratch array. Closest-hit rays visit children near to far,
// shadow (any-hit) rays far to near. Each child box of a compressed node is
// decoded and tested in turn, the hit children (leaves included) are
// ordered by a branch-free sorting network and stacked, and leaves wait for
// a wave-wide leaf step, so the node decode and the triangle test never
// hold registers at the same time (64 VGPRs, 8 waves per SIMD, no spills).
// Box test: slabs (lo - o) * (1/d) on boxes padded outward by
// 2^-20 * (|coordinate| + scene extent) and rounded outward to the
// quantisation grid, far distance (clamped to the best hit) scaled
// by 1 + 2*gamma_3 (Ize 2013), so a box is never culled while it may hold
// a triangle the watertight test reports at t <= best (ties included). Ties
// in t resolve to the smaller primitive id, so the result does not depend
// on the traversal order.
#ifndef WRK_TRACE_WAVES
#define WRK_TRACE_WAVES 8
#endif
// Lanes a wave refills at once (see the refill in k_trace): 16 for scenes
// whose nodes and triangle records fit the 256 MB MALL, 32 for larger ones,
// whose node fetches go to HBM and make each refill's wait longer.
constexpr uint32_t kRefillMinResident = 16, kRefillMinLarge = 32;
constexpr uint64_t kMallBytes = 256ull << 20;
#ifndef WRK_LEAF_BATCH_NUM
#define WRK_LEAF_BATCH_NUM 5
#endif
#ifndef WRK_LEAF_BATCH_DEN
#define WRK_LEAF_BATCH_DEN 8
#endif
// What a finished ray does to its path when the traversal also resolves the
// result (the per-entry decisions of wrk_classify, taken where the ray ends
// instead of in a pass of their own). Pool arrays are indexed by the pool
// entry the queue names; categories by queue position.
struct TraceResolve {
    float4* rad;
    const float4* thr;
    const float4* contrib;
    const uint4* st;
    float4* result;
    const wrk_render_params* prm;  // indexed by the path's slot
    uint8_t* cat;
    const uint32_t* triMat;        // for hits on untagged triangles
    const wrk_material* materials;
    uint32_t numMaterials;
    uint32_t numTraceShadow;       // any hit: entries past this are shadow-only
    float env[3];
};

// Shadow ray done (trace.cu:27-51): an unoccluded ray adds the pending
// contribution; a shadow-only entry then resumes as a Lambertian path
// (reference mode: the continuation waited for the result) or is flushed.
__device__ __forceinline__ void resolve_shadow(const TraceResolve& R, uint32_t p, uint32_t qi, bool occluded) {
    if (!occluded) {
        const float4 L = R.rad[p];
        const float4 C = R.contrib[p];
        R.rad[p] = make_float4(L.x + C.x, L.y + C.y, L.z + C.z, L.w);
    }
    if (qi >= R.numTraceShadow) {
        const uint32_t z = R.st[p].z;
        if (z & FLAG_NEE_DONE) {
            R.cat[qi] = (uint8_t)WRK_MAT_LAMBERTIAN;
        } else {
            const float4 L = R.rad[p];
            R.result[p] = f4(clamp_fireflies(mk(L.x, L.y, L.z), R.prm[z >> FLAG_SLOT_SHIFT].fireflyClamp), 0.0f);
            R.cat[qi] = (uint8_t)kCatDone;
        }
    }
}

// Closest-hit ray done: a hit's shade category, or for a miss the
// environment term (trace.cu:143-145) and the flush (trace.cu:146-148).
__device__ __forceinline__ void resolve_closest(const TraceResolve& R, uint2* hits, uint32_t p, uint32_t qi,
                                                uint32_t bestPrim, float tbest) {
    // the path's result is written where it ends: a miss flushes it, and an
    // out-of-range material id ends it unflushed (shade.cu:35-38)
    f3 res = mk(0.0f, 0.0f, 0.0f);
    uint32_t c = kCatDone;
    if (bestPrim != 0xFFFFFFFFu) {
        const uint32_t prim = bestPrim & kPrimMask;
        hits[p] = make_uint2(__float_as_uint(tbest), prim);
        const uint32_t tag = bestPrim >> 28;
        c = tag != 0u ? tag - 1u : shade_category(R.triMat[prim], R.numMaterials, R.materials);
    } else {
        const uint32_t z = R.st[p].z;
        const float4 T = path_throughput(R.thr, p, z), L = R.rad[p];
        const uint32_t slot = z >> FLAG_SLOT_SHIFT;
        const f3 Lr = mk(L.x + T.x * R.env[0], L.y + T.y * R.env[1], L.z + T.z * R.env[2]);
        res = clamp_fireflies(Lr, R.prm[slot].fireflyClamp);
        R.rad[p] = f4(Lr, 0.0f);
    }
    if (c == kCatDone) R.result[p] = f4(res, 0.0f);
    R.cat[qi] = (uint8_t)c;
}

template <bool ANY, bool COUNT, int W, bool RESOLVE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(WRK_TRACE_WAVES, WRK_TRACE_WAVES)))
void k_trace(const void* __restrict__ nodesIn, const float4* __restrict__ tris,
                                               const float4* __restrict__ org, const float4* __restrict__ dir,
                                               const uint32_t* __restrict__ queue, uint32_t n,
                                               const uint32_t* __restrict__ countDev, uint2* __restrict__ hits,
                                               uint8_t* __restrict__ occl, unsigned long long* __restrict__ fetches,
                                               uint32_t raysPerLane, uint32_t refillMin, TraceResolve R) {
    __shared__ uint32_t sNode[kStackLds][256];
    __shared__ float sT[kStackLds][256];
    __shared__ uint32_t sQi[(ANY && !RESOLVE) ? 1 : 256];  // queue position of each lane's ray
    const uint32_t lane = threadIdx.x;
    uint32_t cnt = n;
    if (countDev) cnt = min(cnt, *countDev);
    // wave-uniform chunk bounds, kept in scalar registers (the launch
    // guarantees n + 256 * raysPerLane fits 32 bits)
    const uint32_t waveId = __builtin_amdgcn_readfirstlane((blockIdx.x * 256u + threadIdx.x) >> 6);
    const uint32_t chunkBegin = waveId * (64u * raysPerLane);
    const uint32_t chunkEnd = min(chunkBegin + 64u * raysPerLane, cnt);
    uint32_t cursor = chunkBegin;
    static_assert(W == 8, "traversal of compressed 8-wide nodes");

    uint32_t spillN[kStackSpill];
    float spillT[kStackSpill];
    bool active = false;
    uint32_t ray = 0, sp = 0, cur = kEmptyChild, leaf = kEmptyChild, bestPrim = 0xFFFFFFFFu;
    float4 o = make_float4(0, 0, 0, 0);
    float tmin = 0.0f, tbest = 0.0f, ix = 0.0f, iy = 0.0f, iz = 0.0f;
    RayPre rp;
    rp.kx = 0; rp.ky = 1; rp.kz = 2; rp.Sx = rp.Sy = rp.Sz = 0.0f; rp.o[0] = rp.o[1] = rp.o[2] = 0.0f;
    bool found = false;
    uint32_t nodeFetches = 0, triFetches = 0, stackSpills = 0;

    for (;;) {
        // refill lanes without a ray from the wave's chunk
        // once refillMin lanes wait for one (or all do): the wave then
        // waits for the new rays' loads (queue entry, then ray) once for
        // several lanes instead of at nearly every step (thresholds 1, 4,
        // 16, 32 gave material 2010-2014, 2043-2046, 2045-2047, 2032-2039
        // and large 1197-1207, 1238-1243, 1233-1250, 1265-1267 Mpaths/s;
        // handing each lane one of the chunk's next 64 queue entries ahead
        // of time, read across lanes at refill, measured 2% slower)
        uint64_t need = __ballot(!active);
        const uint32_t nNeed = (uint32_t)__popcll(need);
        if (nNeed >= refillMin && cursor < chunkEnd) {
            // this lane's rank among the lanes that need a ray
            const uint32_t qi = cursor + __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            if (!active && qi < chunkEnd) {
                ray = queue ? queue[qi] : (uint32_t)qi;
                if (RESOLVE || (!ANY && occl)) sQi[lane] = qi;
                o = org[ray];
                float4 d = dir[ray];
                tmin = o.w;
                tbest = d.w;
                bestPrim = 0xFFFFFFFFu;
                rp = ray_pre(o, d);
                ix = slab_inv(d.x);
                iy = slab_inv(d.y);
                iz = slab_inv(d.z);
                sp = 0;
                cur = 0;  // root
                leaf = kEmptyChild;
                found = false;
                active = true;
            }
            cursor += nNeed;
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;

        // one traversal step: a wave-wide leaf step or node step, then pop
        bool leafStep = false;
        {
            // Leaves found while descending wait in `leaf` (one per lane)
            // while the lane keeps visiting inner nodes; the wave tests its
            // pending leaves together once half of its active lanes hold
            // one, or when no lane has an inner node left to visit. This
            // keeps the node steps and the leaf steps each on most lanes of
            // the wave (measured with 4-triangle leaves: 1/4, 1/2, 5/8, 3/4,
            // all of the lanes gave material 1269, 1411, 1434, 1420, 1259
            // Mpaths/s; without the batching, a lane tests a leaf as soon as
            // it reaches one, 1140; with 2-triangle leaves 1/2, 5/8, 3/4 gave
            // material 1486, 1485, 1446 and large 1061, 1038, 979; after the
            // node test became one fma per plane, 3/8, 1/2, 5/8, 3/4 gave
            // material 1901, 1953-1964, 1985-1996, 1974 and large 1189,
            // 1200-1205, 1197-1204, 1159: the cheaper node steps leave the
            // leaf steps relatively dearer, so the wave waits for more of them;
            // with both triangle records of a leaf loaded together, 1/2, 5/8,
            // 3/4 gave material 1989-1995, 2013-2020, 1992-2000 and large
            // 1199-1224, 1204-1207, 1156-1166).
            const uint32_t nAct = __popcll(__ballot(true));
            const uint32_t nPend = __popcll(__ballot(leaf != kEmptyChild));
            const bool canStep = __ballot(cur != kEmptyChild && !(cur & kLeafBit)) != 0;
            leafStep = nPend != 0 && (!canStep || nPend * WRK_LEAF_BATCH_DEN >= nAct * WRK_LEAF_BATCH_NUM);
        }
        if (leafStep) {
            if (leaf != kEmptyChild) {
            // a pending leaf: its triangles
            const uint32_t first = leaf & 0x07FFFFFFu;
            const uint32_t count = ((leaf >> 27) & 0xFu) + 1u;
            // The first two triangles' records are loaded together, so a
            // two-triangle leaf waits for one memory round trip, not two;
            // they are then tested in leaf order.
            const float4* tp = tris + (size_t)first * 3;
            const float4 a0 = tp[0], a1 = tp[1], a2 = tp[2];
            float4 b0 = a0, b1 = a1, b2 = a2;
            if (count > 1) { b0 = tp[3]; b1 = tp[4]; b2 = tp[5]; }
            // bestPrim keeps the triangle's shade-category tag in bits 28..30
            auto test = [&](float4 v0, float4 v1, float4 v2) {
                if (COUNT) ++triFetches;
                float th;
                if (intersect_tri(rp, v0, v1, v2, tmin, tbest, &th)) {
                    const uint32_t prim = __float_as_uint(v0.w);
                    if (ANY) {
                        found = true;
                    } else if (th < tbest || prim < (bestPrim & kPrimMask)) {
                        tbest = th;
                        bestPrim = prim | ((__float_as_uint(v1.w) & 7u) << 28);
                    }
                }
            };
            test(a0, a1, a2);
            if (count > 1 && !(ANY && found)) test(b0, b1, b2);
            if constexpr (kLeafMax > 2) {
                for (uint32_t k = 2; k < count && !(ANY && found); ++k) {
                    const float4* tk = tris + (size_t)(first + k) * 3;
                    test(tk[0], tk[1], tk[2]);
                }
            }
            leaf = kEmptyChild;
            if (ANY && found) { cur = kEmptyChild; sp = 0; }
            }
        } else if (cur != kEmptyChild && !(cur & kLeafBit)) {
            if (COUNT) ++nodeFetches;
            float key[W];
            uint32_t code[W];
            node8_children(reinterpret_cast<const uint4*>(static_cast<const Node8*>(nodesIn) + cur), o, ix, iy, iz,
                           tmin, tbest, key, code);
            if (ANY && found) {
                cur = kEmptyChild;
                sp = 0;
            } else {
                // children in visiting order (the any-hit order costs 4%
                // more node visits if used for closest-hit rays; measured
                // 21.3 -> 20.5 ms per step for shadow rays against a full sort)
                order_children<ANY>(key, code);
                // push far children (farthest first), descend into the nearest
#pragma unroll
                for (int c = W - 1; c >= 1; --c) {
                    if (key[c] != INFINITY) {
                        if (sp >= kStackLds) {
                            if (COUNT) ++stackSpills;  // the LDS ring's oldest entry moves to scratch
                            uint32_t slot = sp & (kStackLds - 1);
                            spillN[sp - kStackLds] = sNode[slot][lane];
                            spillT[sp - kStackLds] = sT[slot][lane];
                        }
                        sNode[sp & (kStackLds - 1)][lane] = code[c];
                        sT[sp & (kStackLds - 1)][lane] = key[c];
                        ++sp;
                    }
                }
                cur = (key[0] != INFINITY) ? code[0] : kEmptyChild;
            }
        }
        if (cur != kEmptyChild && (cur & kLeafBit) && leaf == kEmptyChild) {
            leaf = cur;  // the nearest child is a leaf: it waits for a leaf step
            cur = kEmptyChild;
        }
        if (cur == kEmptyChild) {
            bool done = true;
            while (sp > 0) {
                --sp;
                uint32_t slot = sp & (kStackLds - 1);
                uint32_t nn = sNode[slot][lane];
                float tt = sT[slot][lane];
                if (sp >= kStackLds) {
                    sNode[slot][lane] = spillN[sp - kStackLds];
                    sT[slot][lane] = spillT[sp - kStackLds];
                }
                if (tt <= tbest * kRobust) {
                    if ((nn & kLeafBit) && leaf == kEmptyChild) {
                        leaf = nn;  // park it and look for an inner node
                        continue;
                    }
                    cur = nn;
                    done = false;
                    break;
                }
            }
            if (leaf != kEmptyChild) done = false;
            if (done) {
                if constexpr (RESOLVE) {
                    if constexpr (ANY) resolve_shadow(R, ray, sQi[lane], found);
                    else resolve_closest(R, hits, ray, sQi[lane], bestPrim, tbest);
                } else if (ANY) {
                    occl[ray] = found ? 1 : 0;
                } else {
                    const bool hit = bestPrim != 0xFFFFFFFFu;
                    hits[ray] = make_uint2(__float_as_uint(tbest), hit ? (bestPrim & kPrimMask) : 0xFFFFFFFFu);
                    if (occl) {
                        // closest hit with categories: the shade category of
                        // the hit triangle at this queue position (0xFF: a
                        // miss or an untagged triangle)
                        const uint32_t tag = bestPrim >> 28;
                        occl[sQi[lane]] = (hit && tag != 0u) ? (uint8_t)(tag - 1u) : (uint8_t)0xFF;
                    }
                }
                active = false;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&fetches[0], (unsigned long long)nodeFetches);
        atomicAdd(&fetches[1], (unsigned long long)triFetches);
        atomicAdd(&fetches[2], (unsigned long long)stackSpills);
    }
}

}  // namespace wrk

using namespace wrk;

namespace {
struct BuildLayout {
    size_t triLo, triHi, keysA, keysB, valsA, valsB, bnodes, levels, partials, tasksA, tasksB, counters, costs,
        sortTemp;
    size_t cidA, cidB, cloA, cloB, chiA, chiB, nn, keep, pos, scanTemp, scanTempBytes;
    size_t sortTempBytes, total;
};
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
BuildLayout layout_for(uint32_t n) {
    BuildLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    size_t nn = std::max<size_t>(n, 2);
    L.triLo = take(nn * 16);
    L.triHi = take(nn * 16);
    L.keysA = take(nn * sizeof(MortonKey));
    L.keysB = take(nn * sizeof(MortonKey));
    L.valsA = take(nn * 4);
    L.valsB = take(nn * 4);
    L.bnodes = take(nn * sizeof(BinNode));
    L.levels = take(nn * 4);
    L.partials = take(((nn + 255) / 256) * 12 * 4);
    L.tasksA = take(nn * 16);
    L.tasksB = take(nn * 16);
    L.counters = take(C_COUNT * 4);
    L.costs = take(nn * kBvhWidth * 4);
    L.sortTempBytes = radix_sort_scratch_bytes((uint32_t)nn);
    L.sortTemp = take(L.sortTempBytes);
    L.cidA = take(nn * 4);
    L.cidB = take(nn * 4);
    L.cloA = take(nn * 16);
    L.cloB = take(nn * 16);
    L.chiA = take(nn * 16);
    L.chiB = take(nn * 16);
    L.nn = take(nn * 4);
    L.keep = take(nn * 4);
    L.pos = take(nn * 4);
    L.scanTempBytes = exclusive_scan_scratch_bytes((uint32_t)nn);
    L.scanTemp = take(L.scanTempBytes);
    L.total = off;
    return L;
}
}  // namespace

extern "C" int wrk_bvh_query_sizes(uint32_t numTris, uint64_t* nodeBytes, uint64_t* triBytes,
                                   uint64_t* scratchBytes) {
    WRK_REQUIRE(nodeBytes && triBytes && scratchBytes, "wrk_bvh_query_sizes: null output");
    WRK_REQUIRE(numTris > 0 && numTris < (1u << 27), "wrk_bvh_query_sizes: numTris must be in [1, 2^27)");
    *nodeBytes = (uint64_t)std::max<uint32_t>(numTris, 1u) * sizeof(Node8);
    *triBytes = (uint64_t)numTris * 48u;
    *scratchBytes = layout_for(numTris).total;
    return WRK_OK;
}

extern "C" int wrk_bvh_build(const float* vertices, const uint32_t* indices, uint32_t numTris, void* nodes,
                             void* tris, void* scratch, uint64_t scratchBytes, wrk_bvh* out, void* stream) {
    WRK_REQUIRE(vertices && indices && nodes && tris && scratch && out, "wrk_bvh_build: null argument");
    WRK_REQUIRE(numTris > 0 && numTris < (1u << 27), "wrk_bvh_build: numTris must be in [1, 2^27)");
    BuildLayout L = layout_for(numTris);
    WRK_REQUIRE(scratchBytes >= L.total, "wrk_bvh_build: scratch too small");
    hipStream_t s = as_stream(stream);
    char* base = static_cast<char*>(scratch);
    float4* triLo = (float4*)(base + L.triLo);
    float4* triHi = (float4*)(base + L.triHi);
    MortonKey* keysA = (MortonKey*)(base + L.keysA);
    MortonKey* keysB = (MortonKey*)(base + L.keysB);
    uint32_t* valsA = (uint32_t*)(base + L.valsA);
    uint32_t* valsB = (uint32_t*)(base + L.valsB);
    BinNode* bnodes = (BinNode*)(base + L.bnodes);
    uint32_t* levels = (uint32_t*)(base + L.levels);
    float* partials = (float*)(base + L.partials);
    uint4* tasksA = (uint4*)(base + L.tasksA);
    uint4* tasksB = (uint4*)(base + L.tasksB);
    uint32_t* counters = (uint32_t*)(base + L.counters);
    float* costs = (float*)(base + L.costs);
    const uint32_t n = numTris;
    const uint32_t g = ceil_div(n, 256);

    hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL(k_tri_bounds, dim3(g), dim3(256), 0, s, vertices, indices, n, triLo, triHi, partials);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(1024), 0, s, partials, g, counters);
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, s, triLo, triHi, n, counters, keysA, valsA);
    WRK_CHECK_LAUNCH();
    // Morton order (stable: equal codes keep triangle order)
    const int sortRc = radix_sort_pairs_u64(keysA, valsA, keysB, valsB, n, 0u, kMortonBits, base + L.sortTemp,
                                            L.sortTempBytes, s);
    if (sortRc != WRK_OK) return sortRc;
    const uint32_t* sorted = valsA;  // triangles in Morton order
    uint32_t* leafOrder = valsB;     // triangles in leaf order (written below)
    uint32_t root = 0;
    const uint32_t* valsF = sorted;
    if (n > 1) {
        {
            uint32_t* cid[2] = {(uint32_t*)(base + L.cidA), (uint32_t*)(base + L.cidB)};
            float4* clo[2] = {(float4*)(base + L.cloA), (float4*)(base + L.cloB)};
            float4* chi[2] = {(float4*)(base + L.chiA), (float4*)(base + L.chiB)};
            uint32_t* nnb = (uint32_t*)(base + L.nn);
            uint32_t* keep = (uint32_t*)(base + L.keep);
            uint32_t* pos = (uint32_t*)(base + L.pos);
            hipLaunchKernelGGL(k_ploc_init, dim3(g), dim3(256), 0, s, triLo, triHi, sorted, n, cid[0], clo[0], chi[0]);
            WRK_CHECK_LAUNCH();
            uint32_t c = n;
            int cur = 0;
            for (uint32_t round = 0; c > 1; ++round) {
                if (round > 4u * n) {
                    set_error("wrk_bvh_build: clustering made no progress");
                    return WRK_ERR_INVALID;
                }
                const uint32_t gc = ceil_div(c, 256);
                hipLaunchKernelGGL(k_ploc_nn, dim3(gc), dim3(256), 0, s, clo[cur], chi[cur], c, nnb);
                hipLaunchKernelGGL(k_ploc_merge, dim3(gc), dim3(256), 0, s, c, nnb, cid[cur], clo[cur], chi[cur],
                                   bnodes, counters, keep);
                WRK_CHECK_LAUNCH();
                const int scanRc = exclusive_scan_u32(keep, pos, c, (uint32_t*)(base + L.scanTemp), L.scanTempBytes,
                                                      nullptr, s);
                if (scanRc != WRK_OK) return scanRc;
                hipLaunchKernelGGL(k_ploc_compact, dim3(gc), dim3(256), 0, s, c, keep, pos, cid[cur], clo[cur],
                                   chi[cur], cid[cur ^ 1], clo[cur ^ 1], chi[cur ^ 1], counters);
                WRK_CHECK_LAUNCH();
                uint32_t next = 0;
                WRK_HIP_TRY(hipMemcpyAsync(&next, counters + C_TASKS_NEXT, 4, hipMemcpyDeviceToHost, s));
                WRK_HIP_TRY(hipStreamSynchronize(s));
                if (next >= c || next == 0) {
                    set_error("wrk_bvh_build: clustering made no progress");
                    return WRK_ERR_INVALID;
                }
                c = next;
                cur ^= 1;
            }
            WRK_HIP_TRY(hipMemcpyAsync(&root, cid[cur], 4, hipMemcpyDeviceToHost, s));
            WRK_HIP_TRY(hipStreamSynchronize(s));
            if (root & kLeafBit) {
                set_error("wrk_bvh_build: clustering ended on a leaf");
                return WRK_ERR_INVALID;
            }
        }
        // binary levels top-down (each internal node lands in exactly one level) ...
        WRK_HIP_TRY(hipMemcpyAsync(levels, &root, 4, hipMemcpyHostToDevice, s));
        std::vector<uint32_t> off = {0u, 1u};
        uint32_t next = 0;
        for (;;) {
            size_t L0 = off.size() - 2;
            uint32_t cnt = off[L0 + 1] - off[L0];
            WRK_HIP_TRY(hipMemsetAsync(counters + C_TASKS_NEXT, 0, 4, s));
            hipLaunchKernelGGL(k_expand_level, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, levels + off[L0], cnt,
                               bnodes, levels + off[L0 + 1], counters + C_TASKS_NEXT);
            WRK_CHECK_LAUNCH();
            WRK_HIP_TRY(hipMemcpyAsync(&next, counters + C_TASKS_NEXT, 4, hipMemcpyDeviceToHost, s));
            WRK_HIP_TRY(hipStreamSynchronize(s));
            if (next == 0) break;
            if ((uint64_t)off.back() + next > (uint64_t)(n - 1)) {
                set_error("wrk_bvh_build: malformed binary hierarchy");
                return WRK_ERR_INVALID;
            }
            off.push_back(off.back() + next);
        }
        {
            // contiguous triangle ranges per subtree: sizes bottom-up, then
            // ranges top-down; leafOrder receives the triangles in that order
            for (size_t L0 = off.size() - 1; L0-- > 0;) {
                uint32_t cnt = off[L0 + 1] - off[L0];
                hipLaunchKernelGGL(k_ploc_sizes, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, levels + off[L0], cnt,
                                   bnodes);
            }
            for (size_t L0 = 0; L0 + 1 < off.size(); ++L0) {
                uint32_t cnt = off[L0 + 1] - off[L0];
                hipLaunchKernelGGL(k_ploc_ranges, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, levels + off[L0], cnt,
                                   L0 == 0, bnodes, sorted, leafOrder);
            }
            WRK_CHECK_LAUNCH();
            valsF = leafOrder;
        }
        // ... then boxes bottom-up, one launch per level
        for (size_t L0 = off.size() - 1; L0-- > 0;) {
            uint32_t cnt = off[L0 + 1] - off[L0];
            hipLaunchKernelGGL(k_refit_level<kBvhWidth>, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, levels + off[L0],
                               cnt, bnodes, triLo, triHi, valsF, costs);
        }
        WRK_CHECK_LAUNCH();
    }
    // level-by-level collapse; root 4-wide node is slot 0
    uint4 rootTask = make_uint4(n > 1 ? root : kLeafBit, 0u, 1u, 0u);
    WRK_HIP_TRY(hipMemcpyAsync(tasksA, &rootTask, sizeof(uint4), hipMemcpyHostToDevice, s));
    uint32_t one = 1;
    WRK_HIP_TRY(hipMemcpyAsync(counters + C_NODES, &one, 4, hipMemcpyHostToDevice, s));
    uint32_t numTasks = 1;
    uint4* tin = tasksA;
    uint4* tout = tasksB;
    while (numTasks > 0) {
        WRK_HIP_TRY(hipMemsetAsync(counters + C_TASKS_NEXT, 0, 4, s));
        hipLaunchKernelGGL(k_collapse<kBvhWidth>, dim3(ceil_div(numTasks, 256)), dim3(256), 0, s, tin, numTasks, bnodes,
                           triLo, triHi, valsF, (const float*)costs, nodes, counters, tout);
        WRK_CHECK_LAUNCH();
        WRK_HIP_TRY(hipMemcpyAsync(&numTasks, counters + C_TASKS_NEXT, 4, hipMemcpyDeviceToHost, s));
        WRK_HIP_TRY(hipStreamSynchronize(s));
        std::swap(tin, tout);
    }
    hipLaunchKernelGGL(k_reorder_tris, dim3(g), dim3(256), 0, s, valsF, vertices, indices, n, (float4*)tris);
    WRK_CHECK_LAUNCH();
    uint32_t hc[C_COUNT];
    WRK_HIP_TRY(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    WRK_HIP_TRY(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->nodes = (uint64_t)(uintptr_t)nodes;
    out->tris = (uint64_t)(uintptr_t)tris;
    out->numNodes = hc[C_NODES];
    out->numTris = n;
    out->maxDepth = hc[C_MAXDEPTH];
    out->nodeBytes = (uint32_t)sizeof(Node8);
    for (int a = 0; a < 3; ++a) {
        out->sceneLo[a] = unordered(hc[C_BMIN + a]);
        out->sceneHi[a] = unordered(hc[C_BMAX + a]);
    }
    if ((kBvhWidth - 1u) * out->maxDepth + 8u > kStackTotal) {
        set_error("wrk_bvh_build: BVH depth " + std::to_string(out->maxDepth) + " exceeds the traversal stack");
        return WRK_ERR_CAPACITY;
    }
    return WRK_OK;
}

#ifndef WRK_TRACE_ROUNDS
#define WRK_TRACE_ROUNDS 4
#endif
// Rays per lane for a launch of n rays: up to kRaysPerLane, while the grid
// still fills every wave slot (CUs x 4 SIMDs x 8 waves) WRK_TRACE_ROUNDS
// times over. Lanes that refill from their wave's chunk keep more lanes
// busy than one ray per lane (fewer than half of the lanes are active per
// VALU instruction in the closest-hit kernel). At 39M rays per launch 8 rays
// per lane gave +12% over one; 12 gave another 3.5% of closest-hit time,
// 16 no more, and at 30M rays the cap is within 1% of 8. Once lanes were
// refilled in batches, a cap of 24 gave another 1% (kRaysPerLane).
static uint32_t rays_per_lane(uint32_t n) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    uint64_t slots = (uint64_t)cus * 4u * 8u * 64u * WRK_TRACE_ROUNDS;
    uint64_t r = n / slots;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(r, kRaysPerLane));
}

template <bool ANY>
static int launch_trace(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue, uint32_t n,
                        const uint32_t* countDev, uint32_t* hits, uint8_t* occl, unsigned long long* fetches,
                        void* stream, const TraceResolve* resolve = nullptr) {
    WRK_REQUIRE(bvh && bvh->nodes && bvh->tris && org4 && dir4, "wrk_trace: null argument");
    WRK_REQUIRE(bvh->nodeBytes == sizeof(Node8), "wrk_trace: unknown node format");
    WRK_REQUIRE((kBvhWidth - 1u) * bvh->maxDepth + 8u <= kStackTotal, "wrk_trace: BVH deeper than the traversal stack");
    if (n == 0) return WRK_OK;
    const uint32_t rpl = rays_per_lane(n);
    WRK_REQUIRE((uint64_t)n + 256ull * rpl <= 0xFFFFFFFFull, "wrk_trace: too many rays for one launch");
    dim3 grid(ceil_div(n, 256u * rpl)), block(256);
    const uint64_t sceneBytes = (uint64_t)bvh->numNodes * bvh->nodeBytes + (uint64_t)bvh->numTris * 48u;
    const uint32_t refill = sceneBytes > kMallBytes ? kRefillMinLarge : kRefillMinResident;
    hipStream_t s = as_stream(stream);
    const void* nd = (const void*)bvh->nodes;
    const float4* tr = (const float4*)bvh->tris;
    const float4* o4 = (const float4*)org4;
    const float4* d4 = (const float4*)dir4;
    uint2* h2 = (uint2*)hits;
    if (resolve) {
        if (fetches)
            hipLaunchKernelGGL((k_trace<ANY, true, 8, true>), grid, block, 0, s, nd, tr, o4, d4, queue, n, countDev,
                               h2, occl, fetches, rpl, refill, *resolve);
        else
            hipLaunchKernelGGL((k_trace<ANY, false, 8, true>), grid, block, 0, s, nd, tr, o4, d4, queue, n, countDev,
                               h2, occl, nullptr, rpl, refill, *resolve);
    } else {
        const TraceResolve none{};
        if (fetches)
            hipLaunchKernelGGL((k_trace<ANY, true, 8, false>), grid, block, 0, s, nd, tr, o4, d4, queue, n, countDev,
                               h2, occl, fetches, rpl, refill, none);
        else
            hipLaunchKernelGGL((k_trace<ANY, false, 8, false>), grid, block, 0, s, nd, tr, o4, d4, queue, n, countDev,
                               h2, occl, nullptr, rpl, refill, none);
    }
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_trace_closest(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                                 uint32_t n, const uint32_t* countDev, uint32_t* hits, void* stream) {
    WRK_REQUIRE(hits, "wrk_trace_closest: null hits");
    return launch_trace<false>(bvh, org4, dir4, queue, n, countDev, hits, nullptr, nullptr, stream);
}

extern "C" int wrk_trace_closest_categorized(const wrk_bvh* bvh, const void* org4, const void* dir4,
                                             const uint32_t* queue, uint32_t n, const uint32_t* countDev,
                                             uint32_t* hits, uint8_t* categories, unsigned long long* fetches,
                                             void* stream) {
    WRK_REQUIRE(hits && categories, "wrk_trace_closest_categorized: null argument");
    return launch_trace<false>(bvh, org4, dir4, queue, n, countDev, hits, categories, fetches, stream);
}

__global__ __launch_bounds__(256) void k_tag_categories(float4* tris, uint32_t n, const uint32_t* __restrict__ triMat,
                                                        const wrk_material* __restrict__ materials,
                                                        uint32_t numMaterials) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t prim = __float_as_uint(tris[3 * (size_t)k].w);
    tris[3 * (size_t)k + 1].w = __uint_as_float(1u + shade_category(triMat[prim], numMaterials, materials));
}

extern "C" int wrk_bvh_tag_categories(const wrk_bvh* bvh, const wrk_scene_view* scene, void* stream) {
    WRK_REQUIRE(bvh && bvh->tris && scene && scene->triMaterialIds, "wrk_bvh_tag_categories: null argument");
    WRK_REQUIRE(scene->numMaterials == 0 || scene->materials, "wrk_bvh_tag_categories: null materials");
    WRK_REQUIRE(bvh->numTris == scene->numTris, "wrk_bvh_tag_categories: BVH and scene differ in triangles");
    if (bvh->numTris == 0) return WRK_OK;
    hipLaunchKernelGGL(k_tag_categories, dim3(ceil_div(bvh->numTris, 256)), dim3(256), 0, as_stream(stream),
                       (float4*)bvh->tris, bvh->numTris, (const uint32_t*)scene->triMaterialIds,
                       (const wrk_material*)scene->materials, scene->numMaterials);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_trace_any(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                             uint32_t n, const uint32_t* countDev, uint8_t* occluded, void* stream) {
    WRK_REQUIRE(occluded, "wrk_trace_any: null occluded");
    return launch_trace<true>(bvh, org4, dir4, queue, n, countDev, nullptr, occluded, nullptr, stream);
}

extern "C" int wrk_trace_closest_counted(const wrk_bvh* bvh, const void* org4, const void* dir4,
                                         const uint32_t* queue, uint32_t n, const uint32_t* countDev, uint32_t* hits,
                                         unsigned long long* fetches, void* stream) {
    WRK_REQUIRE(hits && fetches, "wrk_trace_closest_counted: null argument");
    return launch_trace<false>(bvh, org4, dir4, queue, n, countDev, hits, nullptr, fetches, stream);
}

static TraceResolve resolve_of(const wrk_scene_view* scene, const wrk_render_params* paramsTable,
                                const wrk_paths* paths, uint8_t* categories, uint32_t numTraceShadow) {
    TraceResolve R{};
    R.rad = (float4*)paths->rad;
    R.thr = (const float4*)paths->thr;
    R.contrib = (const float4*)paths->contrib;
    R.st = (const uint4*)paths->state;
    R.result = (float4*)paths->result;
    R.prm = paramsTable;
    R.cat = categories;
    R.numTraceShadow = numTraceShadow;
    if (scene) {
        R.triMat = (const uint32_t*)scene->triMaterialIds;
        R.materials = (const wrk_material*)scene->materials;
        R.numMaterials = scene->numMaterials;
        for (int a = 0; a < 3; ++a) R.env[a] = scene->envRadiance[a];
    }
    return R;
}

extern "C" int wrk_trace_closest_resolve(const wrk_bvh* bvh, const wrk_scene_view* scene,
                                         const wrk_render_params* paramsTable, const wrk_paths* paths,
                                         const uint32_t* queue, uint32_t n, uint8_t* categories,
                                         unsigned long long* fetches, void* stream) {
    WRK_REQUIRE(scene && paramsTable && paths && queue && categories, "wrk_trace_closest_resolve: null argument");
    WRK_REQUIRE(paths->org && paths->dir && paths->hit && paths->rad && paths->thr && paths->state && paths->result,
                "wrk_trace_closest_resolve: incomplete path pool");
    WRK_REQUIRE(scene->triMaterialIds && (scene->numMaterials == 0 || scene->materials),
                "wrk_trace_closest_resolve: incomplete scene view");
    const TraceResolve R = resolve_of(scene, paramsTable, paths, categories, 0);
    return launch_trace<false>(bvh, (const void*)paths->org, (const void*)paths->dir, queue, n, nullptr,
                               (uint32_t*)paths->hit, nullptr, fetches, stream, &R);
}

extern "C" int wrk_trace_any_resolve(const wrk_bvh* bvh, const wrk_render_params* paramsTable,
                                     const wrk_paths* paths, const uint32_t* queue, uint32_t numTraceShadow,
                                     uint32_t numShadow, uint8_t* categories, unsigned long long* fetches,
                                     void* stream) {
    WRK_REQUIRE(paramsTable && paths && queue && categories, "wrk_trace_any_resolve: null argument");
    WRK_REQUIRE(paths->savedOrg && paths->savedDir && paths->rad && paths->contrib && paths->state && paths->result,
                "wrk_trace_any_resolve: incomplete path pool");
    WRK_REQUIRE((uint64_t)numTraceShadow + numShadow <= 0xFFFFFFFFull, "wrk_trace_any_resolve: too many rays");
    const TraceResolve R = resolve_of(nullptr, paramsTable, paths, categories, numTraceShadow);
    return launch_trace<true>(bvh, (const void*)paths->savedOrg, (const void*)paths->savedDir, queue,
                              numTraceShadow + numShadow, nullptr, nullptr, nullptr, fetches, stream, &R);
}

extern "C" int wrk_trace_any_counted(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                                     uint32_t n, const uint32_t* countDev, uint8_t* occluded,
                                     unsigned long long* fetches, void* stream) {
    WRK_REQUIRE(occluded && fetches, "wrk_trace_any_counted: null argument");
    return launch_trace<true>(bvh, org4, dir4, queue, n, countDev, nullptr, occluded, fetches, stream);
}
