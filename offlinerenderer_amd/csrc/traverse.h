// traverse.h — the compressed 8-wide node test shared by the wave-cooperative
// traversal kernels (bvh.hip) and the one-path-per-thread traversal of the
// fused path kernel (wavefront.hip).
#pragma once

#include "wrk_common.h"
#include "wrk_device.h"

namespace wrk {

// Far distances are scaled by 1 + 2*gamma_3 (Ize 2013) so that a box is never
// culled while it may hold a triangle the watertight test reports at t <= best.
constexpr float kRobust = 1.0000005f;

// Reciprocal ray direction of the slab test, with components below 1e-20
// replaced by +-1e-20.
__device__ __forceinline__ float slab_inv(float d) { return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d)); }

// Child boxes of compressed node `raw` (six 16-byte words: origin | child
// 0-3 | child 4-7 | 48 quantised planes): code[c] is the child word and
// key[c] the near distance of the ray (origin o, reciprocal direction ix,
// iy, iz) into box c over [tmin, tbest], or +inf on a miss or an empty slot.
// A plane origin + q * 2^e sits at ray distance (origin + q * 2^e - o) * inv
// = q * (2^e * inv) + (origin - o) * inv, one fma per plane after a per-node
// setup (2^e * inv is exact). Per axis the ray's direction sign names the
// near and the far plane (lo then hi for inv >= 0), so no per-child min/max
// of the two slab distances is needed. The build pads every box outward by
// 2^-20 * (|coordinate| + scene extent) (kBoxPadRel), several times the
// rounding of either distance form for origins within a few scene extents,
// and far distances are scaled by kRobust, so a box that may hold the
// closest hit is never culled.
__device__ __forceinline__ void node8_children(const uint4* __restrict__ raw, float4 o, float ix, float iy, float iz,
                                               float tmin, float tbest, float key[8], uint32_t code[8]) {
    const uint4 r0 = raw[0], r1 = raw[1], r2 = raw[2], r3 = raw[3], r4 = raw[4], r5 = raw[5];
    const float ox = __uint_as_float(r0.x), oy = __uint_as_float(r0.y), oz = __uint_as_float(r0.z);
    const float sx = __uint_as_float((r0.w & 0xFFu) << 23);
    const float sy = __uint_as_float(((r0.w >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((r0.w >> 16) & 0xFFu) << 23);
    const float ax = sx * ix, ay = sy * iy, az = sz * iz;
    const float bx = (ox - o.x) * ix, by = (oy - o.y) * iy, bz = (oz - o.z) * iz;
    code[0] = r1.x; code[1] = r1.y; code[2] = r1.z; code[3] = r1.w;
    code[4] = r2.x; code[5] = r2.y; code[6] = r2.z; code[7] = r2.w;
    // plane words: qlo x (r3.x r3.y) y (r3.z r3.w) z (r4.x r4.y),
    //              qhi x (r4.z r4.w) y (r5.x r5.y) z (r5.z r5.w)
    const bool px = ix >= 0.0f, py = iy >= 0.0f, pz = iz >= 0.0f;
    const uint32_t nx[2] = {px ? r3.x : r4.z, px ? r3.y : r4.w}, fx[2] = {px ? r4.z : r3.x, px ? r4.w : r3.y};
    const uint32_t ny[2] = {py ? r3.z : r5.x, py ? r3.w : r5.y}, fy[2] = {py ? r5.x : r3.z, py ? r5.y : r3.w};
    const uint32_t nz[2] = {pz ? r4.x : r5.z, pz ? r4.y : r5.w}, fz[2] = {pz ? r5.z : r4.x, pz ? r5.w : r4.y};
    const float tfar = tbest;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int wi = c >> 2, sh = 8 * (c & 3);
        auto q = [&](uint32_t word) { return (float)((word >> sh) & 0xFFu); };
        const float tn = fmaxf(fmaxf(fmaf(q(nx[wi]), ax, bx), fmaf(q(ny[wi]), ay, by)),
                               fmaxf(fmaf(q(nz[wi]), az, bz), tmin));
        const float tf = fminf(fminf(fmaf(q(fx[wi]), ax, bx), fmaf(q(fy[wi]), ay, by)),
                               fminf(fmaf(q(fz[wi]), az, bz), tfar)) * kRobust;
        key[c] = (code[c] != kEmptyChild && tn <= tf) ? tn : INFINITY;
    }
}

// branch-free compare-exchange of (key, code) pairs (selects, no divergent blocks)
#define WRK_CSWAP(a, b)                                               \
    {                                                                 \
        const float ka = key[a], kb = key[b];                         \
        const uint32_t ca = code[a], cb = code[b];                    \
        const bool sw = kb < ka;                                      \
        key[a] = sw ? kb : ka; key[b] = sw ? ka : kb;                 \
        code[a] = sw ? cb : ca; code[b] = sw ? ca : cb;               \
    }

// Children in visiting order. Closest hit: near to far (Batcher odd-even
// merge network, 19 comparators; a 7-comparator tournament that only puts
// the nearest first measured 4.3% more node visits and 2.4% more
// closest-hit time). Any hit: the farthest child first (the
// keys are negated, and a tournament of 7 comparators puts the largest
// distance in slot 0; the order of the others barely matters to an any-hit
// query). Shadow rays leaving a surface meet their occluders toward their end
// more often than among the boxes around their origin (measured: 2-3x fewer
// node visits than near-first). Negated keys never cull on pop: tbest does
// not shrink for an any-hit query.
template <bool ANY>
__device__ __forceinline__ void order_children(float key[8], uint32_t code[8]) {
    if constexpr (ANY) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (key[c] != INFINITY) key[c] = -key[c];
        WRK_CSWAP(0, 1) WRK_CSWAP(2, 3) WRK_CSWAP(4, 5) WRK_CSWAP(6, 7)
        WRK_CSWAP(0, 2) WRK_CSWAP(4, 6) WRK_CSWAP(0, 4)
    } else {
        WRK_CSWAP(0, 1) WRK_CSWAP(2, 3) WRK_CSWAP(4, 5) WRK_CSWAP(6, 7)
        WRK_CSWAP(0, 2) WRK_CSWAP(1, 3) WRK_CSWAP(4, 6) WRK_CSWAP(5, 7)
        WRK_CSWAP(1, 2) WRK_CSWAP(5, 6)
        WRK_CSWAP(0, 4) WRK_CSWAP(1, 5) WRK_CSWAP(2, 6) WRK_CSWAP(3, 7)
        WRK_CSWAP(2, 4) WRK_CSWAP(3, 5)
        WRK_CSWAP(1, 2) WRK_CSWAP(3, 4) WRK_CSWAP(5, 6)
    }
}
#undef WRK_CSWAP

// One ray on one thread: the node test above and the watertight triangle
// test, children visited in order_children's order, a leaf's triangles
// tested as soon as it is reached. The stack keeps its top kStackLds
// entries in the caller's LDS column (sNode[k * 256], sT[k * 256]) and older
// entries in a per-thread scratch array. Closest hit: returns whether the ray
// hit; *tHit and *primHit get the nearest hit (ties in t to the smaller
// primitive id) with the triangle's shade-category tag in primHit bits
// 28..30, the same result as k_trace. Any hit: returns whether any triangle
// lies in [o.w, d.w].
template <bool ANY>
__device__ __forceinline__ bool trace_thread(const Node8* __restrict__ nodes, const float4* __restrict__ tris,
                                             float4 o, float4 d, uint32_t* sNode, float* sT, float* tHit,
                                             uint32_t* primHit) {
    uint32_t spillN[kStackSpill];
    float spillT[kStackSpill];
    const RayPre rp = ray_pre(o, d);
    const float ix = slab_inv(d.x), iy = slab_inv(d.y), iz = slab_inv(d.z);
    const float tmin = o.w;
    float tbest = d.w;
    uint32_t bestPrim = 0xFFFFFFFFu;
    uint32_t sp = 0, cur = 0;  // the root
    for (;;) {
        if (cur != kEmptyChild) {
            if (cur & kLeafBit) {
                const uint32_t first = cur & 0x07FFFFFFu;
                const uint32_t count = ((cur >> 27) & 0xFu) + 1u;
                // the first two triangles' records are loaded together (one
                // memory round trip for a two-triangle leaf), then tested in
                // leaf order
                const float4* tp = tris + (size_t)first * 3;
                const float4 a0 = tp[0], a1 = tp[1], a2 = tp[2];
                float4 b0 = a0, b1 = a1, b2 = a2;
                if (count > 1) { b0 = tp[3]; b1 = tp[4]; b2 = tp[5]; }
                bool found = false;
                auto test = [&](float4 v0, float4 v1, float4 v2) {
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
                if (ANY && found) return true;
                if (count > 1) test(b0, b1, b2);
                if (ANY && found) return true;
                if constexpr (kLeafMax > 2) {
                    for (uint32_t k = 2; k < count; ++k) {
                        const float4* tk = tris + (size_t)(first + k) * 3;
                        test(tk[0], tk[1], tk[2]);
                        if (ANY && found) return true;
                    }
                }
                cur = kEmptyChild;
            } else {
                float key[8];
                uint32_t code[8];
                node8_children(reinterpret_cast<const uint4*>(nodes + cur), o, ix, iy, iz, tmin, tbest, key, code);
                order_children<ANY>(key, code);
#pragma unroll
                for (int c = 7; c >= 1; --c) {
                    if (key[c] != INFINITY) {
                        const uint32_t slot = sp & (kStackLds - 1);
                        if (sp >= kStackLds) {
                            spillN[sp - kStackLds] = sNode[slot * 256];
                            spillT[sp - kStackLds] = sT[slot * 256];
                        }
                        sNode[slot * 256] = code[c];
                        sT[slot * 256] = key[c];
                        ++sp;
                    }
                }
                cur = key[0] != INFINITY ? code[0] : kEmptyChild;
                continue;
            }
        }
        // pop the nearest entry still in front of the best hit
        while (sp > 0) {
            --sp;
            const uint32_t slot = sp & (kStackLds - 1);
            const uint32_t nn = sNode[slot * 256];
            const float tt = sT[slot * 256];
            if (sp >= kStackLds) {
                sNode[slot * 256] = spillN[sp - kStackLds];
                sT[slot * 256] = spillT[sp - kStackLds];
            }
            if (tt <= tbest * kRobust) {
                cur = nn;
                break;
            }
        }
        if (cur == kEmptyChild) break;
    }
    *tHit = tbest;
    *primHit = bestPrim;
    return bestPrim != 0xFFFFFFFFu;
}

}  // namespace wrk
