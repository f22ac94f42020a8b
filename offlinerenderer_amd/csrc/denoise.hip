// denoise.hip — kernels of the learned denoiser (the OptiX AI denoiser of
// libWR/src/internal/denoiser.cpp:17-195 is closed; this U-Net replaces it).
//
// Tensors are NHWC bf16 with a channel pitch (so a convolution can write
// straight into one channel slice of a concatenation buffer). 3x3
// convolutions are implicit GEMMs on the gfx950 bf16 MFMA
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   forward  Y[p][co]    = sum_tap,ci W[tap][co][ci] X[p + off(tap)][ci]
//            (M = output channels, N = pixels, K = 9 taps x input channels);
//   data gradient: the forward kernel over dZ with the tap-flipped,
//            transposed weights W'[tap][ci][co] = W[8 - tap][co][ci];
//   weight gradient dW[tap][co][ci] = sum_p dZ[p][co] X[p + off(tap)][ci]
//            (M = co, N = ci, K = pixels), per-block partial sums then one
//            reduction (no float atomics: the result does not depend on
//            scheduling).
// MFMA operand layout (16x16x32): lane l holds row/column (l & 15) and
// K elements 8*(l >> 4) .. +7; the accumulator lane l holds rows
// 4*(l >> 4) + i, column l & 15.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "wrk_common.h"

namespace wrk {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// FP32 -> bf16, round to nearest even (NaN stays NaN), written out so the
// result does not depend on how the compiler lowers the conversion.
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    uint16_t h;
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        h = (uint16_t)((u >> 16) | 0x40u);
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
        h = (uint16_t)(u >> 16);
    }
    return __builtin_bit_cast(bf16, h);
}

// ------------------------------------------------------------ conv forward
constexpr int kTH = 8, kTW = 16;            // output pixels per block (4 waves x 2 rows x 16)
constexpr int kIH = kTH + 2, kIW = kTW + 2;  // input tile with the 1-pixel halo

// The MFMAs of one 32-channel input chunk over a block's 8x16 output tile:
// wave w takes output rows 2w, 2w + 1; sIn holds the input tile with its
// halo ([kIH * kIW][32]), sW the chunk's weights ([9][CT * 16][32]).
template <int CT>
__device__ __forceinline__ void conv_tile_mma(const bf16* sIn, const bf16* sW, int lane, int wave,
                                              f32x4 (&acc)[2][CT]) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3, dx = tap % 3;
        bf16x8 a[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c)
            a[c] = *(const bf16x8*)(sW + ((tap * CT + c) * 16 + (lane & 15)) * 32 + (lane >> 4) * 8);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int py = wave * 2 + r + dy, px = (lane & 15) + dx;
            const bf16x8 b = *(const bf16x8*)(sIn + (py * kIW + px) * 32 + (lane >> 4) * 8);
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[r][c] = mfma16(a[c], b, acc[r][c]);
        }
    }
}

// Bias, ReLU and the bf16 store of a block's 8x16 output tile at (x0, y0)
// of image base `img`, output channels co0 .. co0 + 16 * CT - 1 (those
// below cout); bv holds each lane's biases (conv_tile_bias).
template <int CT>
__device__ __forceinline__ void conv_tile_bias(const float* __restrict__ bias, uint32_t cout, int co0, int lane,
                                               f32x4 (&bv)[CT]) {
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int co = co0 + c * 16 + (lane >> 4) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[c][i] = (bias && co < (int)cout) ? bias[co + i] : 0.0f;
    }
}

// With rgb given, the tile's channels 0..2 are unpacked instead of stored:
// rgb[p][c] = max(expm1(y + base[p][c]), 0), as wrk_unet_unpack computes it.
template <int CT>
__device__ __forceinline__ void conv_tile_store(const f32x4 (&acc)[2][CT], const f32x4 (&bv)[CT],
                                                bf16* __restrict__ y, uint32_t yPitch, uint32_t cout, int co0,
                                                uint32_t relu, size_t img, int x0, int y0, uint32_t H, uint32_t W,
                                                int lane, int wave, bf16* __restrict__ yPool, uint32_t poolPitch,
                                                const bf16* __restrict__ base = nullptr, uint32_t basePitch = 0,
                                                float* __restrict__ rgb = nullptr) {
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int co = co0 + c * 16 + (lane >> 4) * 4;
        union { bf16 h[4]; uint2 u; } pk[2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[r][c][i] + bv[c][i];
                if (relu) v = fmaxf(v, 0.0f);
                pk[r].h[i] = f2bf(v);
            }
        const int ox = x0 + (lane & 15);
        if (rgb) {
            if (co == 0 && ox < (int)W) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int oy = y0 + wave * 2 + r;
                    if (oy >= (int)H) continue;
                    const size_t px = img + (size_t)oy * W + ox;
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const float q = (float)pk[r].h[c3] + (float)base[px * basePitch + c3];
                        rgb[px * 3 + c3] = fmaxf(expm1f(q), 0.0f);
                    }
                }
            }
        } else if (co < (int)cout && ox < (int)W) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int oy = y0 + wave * 2 + r;
                if (oy < (int)H) *(uint2*)(y + (img + (size_t)oy * W + ox) * yPitch + co) = pk[r].u;
            }
        }
        if (yPool) {
            // 2x2 average of the stored bf16 values in wrk_pool2_fwd's order
            // (row-major over the window): rows 2w, 2w + 1 are this wave's,
            // columns x, x + 1 are lanes l, l ^ 1
            uint2 nb[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                nb[r].x = (uint32_t)__shfl_xor((int)pk[r].u.x, 1);
                nb[r].y = (uint32_t)__shfl_xor((int)pk[r].u.y, 1);
            }
            const int oy = y0 + wave * 2;
            if ((lane & 1) == 0 && co < (int)cout && ox < (int)W && oy < (int)H) {
                union { bf16 h[4]; uint2 u; } q[2], o;
                q[0].u = nb[0];
                q[1].u = nb[1];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float s = 0.0f;
                    s += (float)pk[0].h[i];
                    s += (float)q[0].h[i];
                    s += (float)pk[1].h[i];
                    s += (float)q[1].h[i];
                    o.h[i] = f2bf(s * 0.25f);
                }
                *(uint2*)(yPool + (img / 4 + (size_t)(oy >> 1) * (W >> 1) + (ox >> 1)) * poolPitch + co) = o.u;
            }
        }
    }
}

// wCout is the weights' output-channel count (their tap stride); cout the
// channels computed (co0 + 16 * CT <= cout rounded up to 16). Input
// channels below cUp are the 2x nearest upsampling of xUp ((H/2, W/2),
// pitch upPitch), read at (y/2, x/2) while staging, so a decoder's
// [upsampled | skip] input needs no upsampled copy; the others come from x.
// At least 3 waves per SIMD: without the bound the 64-channel form takes
// 174 registers and runs 2 (its layers then take ~17 us longer each).
template <int CT>  // output-channel tiles of 16 per block
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
void k_conv_fwd(const bf16* __restrict__ x, uint32_t xPitch, uint32_t cin, const bf16* __restrict__ xUp,
                uint32_t upPitch, uint32_t cUp, const bf16* __restrict__ w, uint32_t wCout,
                const float* __restrict__ bias, bf16* __restrict__ y, uint32_t yPitch, uint32_t cout, uint32_t relu,
                uint32_t H, uint32_t W, uint32_t tilesX, bf16* __restrict__ yPool, uint32_t poolPitch) {
    __shared__ __attribute__((aligned(16))) bf16 sIn[kIH * kIW * 32];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * CT * 16 * 32];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int x0 = (int)(blockIdx.x % tilesX) * kTW, y0 = (int)(blockIdx.x / tilesX) * kTH;
    const int co0 = (int)blockIdx.y * CT * 16;
    const size_t img = (size_t)blockIdx.z * H * W;
    f32x4 acc[2][CT];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // Software pipeline: the input and weight tiles of the next 32 input
    // channels are loaded into registers while the MFMAs run on the current
    // ones in LDS, so global-load latency overlaps the math.
    constexpr int kInParts = kIH * kIW * 4, kWParts = 9 * CT * 16 * 4;
    constexpr int kInPer = (kInParts + 255) / 256, kWPer = (kWParts + 255) / 256;
    uint4 rIn[kInPer], rW[kWPer];
    auto fetch = [&](uint32_t ci0) {
        const bool up = ci0 < cUp;  // the same for the whole block
#pragma unroll
        for (int k = 0; k < kInPer; ++k) {
            const int e = t + 256 * k;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (e < kInParts) {
                const int pix = e >> 2, part = e & 3;
                const int gy = y0 + pix / kIW - 1, gx = x0 + pix % kIW - 1;
                if (gy >= 0 && gy < (int)H && gx >= 0 && gx < (int)W) {
                    if (up)
                        v = *(const uint4*)(xUp + (img / 4 + (size_t)(gy >> 1) * (W >> 1) + (gx >> 1)) * upPitch +
                                            ci0 + part * 8);
                    else
                        v = *(const uint4*)(x + (img + (size_t)gy * W + gx) * xPitch + ci0 + part * 8);
                }
            }
            rIn[k] = v;
        }
#pragma unroll
        for (int k = 0; k < kWPer; ++k) {
            const int e = t + 256 * k;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (e < kWParts) {
                const int row = e >> 2, part = e & 3;
                const int tap = row / (CT * 16), col = row % (CT * 16);
                if (co0 + col < (int)cout)
                    v = *(const uint4*)(w + ((size_t)tap * wCout + co0 + col) * cin + ci0 + part * 8);
            }
            rW[k] = v;
        }
    };
    fetch(0);
    for (uint32_t ci0 = 0; ci0 < cin; ci0 += 32) {
#pragma unroll
        for (int k = 0; k < kInPer; ++k) {
            const int e = t + 256 * k;
            if (e < kInParts) *(uint4*)(sIn + (e >> 2) * 32 + (e & 3) * 8) = rIn[k];
        }
#pragma unroll
        for (int k = 0; k < kWPer; ++k) {
            const int e = t + 256 * k;
            if (e < kWParts) *(uint4*)(sW + (e >> 2) * 32 + (e & 3) * 8) = rW[k];
        }
        __syncthreads();
        if (ci0 + 32 < cin) fetch(ci0 + 32);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 a[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c)
                a[c] = *(const bf16x8*)(sW + ((tap * CT + c) * 16 + (lane & 15)) * 32 + (lane >> 4) * 8);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int py = wave * 2 + r + dy, px = (lane & 15) + dx;
                const bf16x8 b = *(const bf16x8*)(sIn + (py * kIW + px) * 32 + (lane >> 4) * 8);
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[r][c] = mfma16(a[c], b, acc[r][c]);
            }
        }
        __syncthreads();
    }
    f32x4 bv[CT];
    conv_tile_bias<CT>(bias, cout, co0, lane, bv);
    conv_tile_store<CT>(acc, bv, y, yPitch, cout, co0, relu, img, x0, y0, H, W, lane, wave, yPool, poolPitch);
}

// The same convolution for 32 input channels (one chunk), over many tiles
// per block: the block stages its weight slice once, then walks the pixel
// tiles tile = blockIdx.x, blockIdx.x + gridDim.x, ... of all images, and
// loads the next tile's input into registers while the MFMAs and the store
// of the current one run. With one chunk, the kernel above has nothing to
// overlap its loads with and fetches the weights again for every tile.
// At least 4 waves per SIMD: at 3 it ran no faster than that kernel.
template <int CT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_conv_fwd_c32(const bf16* __restrict__ x, uint32_t xPitch, const bf16* __restrict__ w, uint32_t wCout,
                    const float* __restrict__ bias, bf16* __restrict__ y, uint32_t yPitch, uint32_t cout, uint32_t relu,
                    uint32_t H, uint32_t W, uint32_t tilesX, uint32_t tilesPerImage, uint32_t nTiles,
                    bf16* __restrict__ yPool, uint32_t poolPitch, const bf16* __restrict__ base, uint32_t basePitch,
                    float* __restrict__ rgb) {
    __shared__ __attribute__((aligned(16))) bf16 sIn[kIH * kIW * 32];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * CT * 16 * 32];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int co0 = (int)blockIdx.y * CT * 16;
    constexpr int kWParts = 9 * CT * 16 * 4;
    for (int e = t; e < kWParts; e += 256) {
        const int row = e >> 2, part = e & 3;
        const int tap = row / (CT * 16), col = row % (CT * 16);
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (co0 + col < (int)cout) v = *(const uint4*)(w + ((size_t)tap * wCout + co0 + col) * 32 + part * 8);
        *(uint4*)(sW + row * 32 + part * 8) = v;
    }
    // the biases are loaded once: a load in the tile loop would also wait
    // for the next tile's prefetch (loads complete in order)
    f32x4 bv[CT];
    conv_tile_bias<CT>(bias, cout, co0, lane, bv);
    constexpr int kInParts = kIH * kIW * 4, kInPer = (kInParts + 255) / 256;
    uint4 rIn[kInPer];
    auto fetch = [&](uint32_t tile) {
        const uint32_t tin = tile % tilesPerImage;
        const size_t img = (size_t)(tile / tilesPerImage) * H * W;
        const int x0 = (int)(tin % tilesX) * kTW, y0 = (int)(tin / tilesX) * kTH;
#pragma unroll
        for (int k = 0; k < kInPer; ++k) {
            const int e = t + 256 * k;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (e < kInParts) {
                const int pix = e >> 2, part = e & 3;
                const int gy = y0 + pix / kIW - 1, gx = x0 + pix % kIW - 1;
                if (gy >= 0 && gy < (int)H && gx >= 0 && gx < (int)W)
                    v = *(const uint4*)(x + (img + (size_t)gy * W + gx) * xPitch + part * 8);
            }
            rIn[k] = v;
        }
    };
    uint32_t tile = blockIdx.x;
    if (tile < nTiles) fetch(tile);
    for (; tile < nTiles; tile += gridDim.x) {
#pragma unroll
        for (int k = 0; k < kInPer; ++k) {
            const int e = t + 256 * k;
            if (e < kInParts) *(uint4*)(sIn + (e >> 2) * 32 + (e & 3) * 8) = rIn[k];
        }
        __syncthreads();  // the tile (and, the first time, the weights) are in LDS
        if (tile + gridDim.x < nTiles) fetch(tile + gridDim.x);
        f32x4 acc[2][CT];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        conv_tile_mma<CT>(sIn, sW, lane, wave, acc);
        const uint32_t tin = tile % tilesPerImage;
        conv_tile_store<CT>(acc, bv, y, yPitch, cout, co0, relu, (size_t)(tile / tilesPerImage) * H * W,
                            (int)(tin % tilesX) * kTW, (int)(tin / tilesX) * kTH, H, W, lane, wave, yPool,
                            poolPitch, base, basePitch, rgb);
        __syncthreads();  // every wave is done with sIn before the next tile overwrites it
    }
}

// ------------------------------------------------------- weight gradient
constexpr int kGH = 4, kGW = 32;  // pixels per K tile: 4 rows x 32 columns (4 MFMA K steps)

// Block: 32 output x 32 input channels, all 9 taps, a run of pixel tiles.
// Wave w: output channels 16*(w >> 1).., input channels 16*(w & 1)..
// X is staged three times, shifted by dx = 0, 1, 2, so every operand read is
// one aligned 16-byte LDS load.
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16* __restrict__ x, uint32_t xPitch, uint32_t cin,
                                                    const bf16* __restrict__ dz, uint32_t dzPitch, uint32_t cout,
                                                    uint32_t N, uint32_t H, uint32_t W, uint32_t tilesX,
                                                    uint32_t tilesPerImage, uint32_t tilesPerBlock,
                                                    float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) bf16 sZ[32 * kGH * kGW];
    __shared__ __attribute__((aligned(16))) bf16 sX[3 * 32 * (kGH + 2) * kGW];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t ciTiles = cin / 32;
    const int co0 = (int)(blockIdx.x / ciTiles) * 32, ci0 = (int)(blockIdx.x % ciTiles) * 32;
    const int coSub = (wave >> 1) * 16, ciSub = (wave & 1) * 16;
    f32x4 acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint64_t totalTiles = (uint64_t)N * tilesPerImage;
    const uint64_t tBegin = (uint64_t)blockIdx.y * tilesPerBlock;
    const uint64_t tEnd = tBegin + tilesPerBlock < totalTiles ? tBegin + tilesPerBlock : totalTiles;
    for (uint64_t tile = tBegin; tile < tEnd; ++tile) {
        const uint32_t n = (uint32_t)(tile / tilesPerImage), tin = (uint32_t)(tile % tilesPerImage);
        const int x0 = (int)(tin % tilesX) * kGW, y0 = (int)(tin / tilesX) * kGH;
        const size_t img = (size_t)n * H * W;
        // dZ tile, transposed to [co][pixel]
        for (int e = t; e < kGH * kGW * 4; e += 256) {
            const int pix = e >> 2, part = e & 3;
            const int gy = y0 + pix / kGW, gx = x0 + pix % kGW;
            union { uint4 u; bf16 h[8]; } v;
            v.u = make_uint4(0u, 0u, 0u, 0u);
            if (gy < (int)H && gx < (int)W)
                v.u = *(const uint4*)(dz + (img + (size_t)gy * W + gx) * dzPitch + co0 + part * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) sZ[(part * 8 + j) * (kGH * kGW) + pix] = v.h[j];
        }
        // X rows y0-1 .. y0+kGH, three copies shifted by dx - 1
        for (int e = t; e < 3 * (kGH + 2) * kGW * 4; e += 256) {
            const int part = e & 3, q = e >> 2;
            const int col = q % kGW, row = (q / kGW) % (kGH + 2), d = q / (kGW * (kGH + 2));
            const int gy = y0 + row - 1, gx = x0 + col + d - 1;
            union { uint4 u; bf16 h[8]; } v;
            v.u = make_uint4(0u, 0u, 0u, 0u);
            if (gy >= 0 && gy < (int)H && gx >= 0 && gx < (int)W)
                v.u = *(const uint4*)(x + (img + (size_t)gy * W + gx) * xPitch + ci0 + part * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                sX[((d * 32 + part * 8 + j) * (kGH + 2) + row) * kGW + col] = v.h[j];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kGH; ++r) {
            const bf16x8 a =
                *(const bf16x8*)(sZ + (coSub + (lane & 15)) * (kGH * kGW) + r * kGW + (lane >> 4) * 8);
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int dy = tap / 3, dx = tap % 3;
                const bf16x8 b = *(const bf16x8*)(sX + ((dx * 32 + ciSub + (lane & 15)) * (kGH + 2) + r + dy) * kGW +
                                                  (lane >> 4) * 8);
                acc[tap] = mfma16(a, b, acc[tap]);
            }
        }
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.y * 9 * cout * cin;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = co0 + coSub + (lane >> 4) * 4 + i, ci = ci0 + ciSub + (lane & 15);
            out[((size_t)tap * cout + co) * cin + ci] = acc[tap][i];
        }
}

// out[i] = sum_s partial[s][i], in split order
__global__ __launch_bounds__(256) void k_reduce_splits(const float* __restrict__ partial, uint32_t splits,
                                                       uint64_t n, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < splits; ++k) s += partial[(uint64_t)k * n + i];
    out[i] = s;
}

// bias gradient: per-block column sums over a pixel run, then k_reduce_splits
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* __restrict__ dz, uint32_t dzPitch, uint32_t cout,
                                                   uint64_t npix, uint64_t pixPerBlock, float* __restrict__ partial) {
    const uint32_t c = threadIdx.x;
    if (c >= cout) return;
    const uint64_t p0 = (uint64_t)blockIdx.x * pixPerBlock;
    const uint64_t p1 = p0 + pixPerBlock < npix ? p0 + pixPerBlock : npix;
    float s = 0.0f;
    for (uint64_t p = p0; p < p1; ++p) s += (float)dz[p * dzPitch + c];
    partial[(uint64_t)blockIdx.x * cout + c] = s;
}

// ------------------------------------------------------------ elementwise
// Each thread handles 8 channels (16 bytes) of one pixel; C % 8 == 0.

// dY = Y > 0 ? dY : 0
__global__ __launch_bounds__(256) void k_relu_bwd(bf16* __restrict__ dy, uint32_t dyPitch,
                                                  const bf16* __restrict__ y, uint32_t yPitch, uint32_t C,
                                                  uint64_t npix) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t groups = C / 8;
    if (e >= npix * groups) return;
    const uint64_t p = e / groups;
    const uint32_t c = (uint32_t)(e % groups) * 8;
    union { uint4 u; bf16 h[8]; } g, v;
    g.u = *(const uint4*)(dy + p * dyPitch + c);
    v.u = *(const uint4*)(y + p * yPitch + c);
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (!((float)v.h[j] > 0.0f)) g.h[j] = f2bf(0.0f);
    *(uint4*)(dy + p * dyPitch + c) = g.u;
}

// 2x2 average pool: (n, H, W) -> (n, H/2, W/2)
__global__ __launch_bounds__(256) void k_pool2_fwd(const bf16* __restrict__ x, uint32_t xPitch, uint32_t C,
                                                   bf16* __restrict__ y, uint32_t yPitch, uint32_t N, uint32_t H,
                                                   uint32_t W) {
    const uint32_t Ho = H / 2, Wo = W / 2, groups = C / 8;
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)N * Ho * Wo * groups) return;
    const uint64_t q = e / groups;
    const uint32_t c = (uint32_t)(e % groups) * 8;
    const uint32_t ox = (uint32_t)(q % Wo), oy = (uint32_t)((q / Wo) % Ho), n = (uint32_t)(q / ((uint64_t)Wo * Ho));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t p = ((uint64_t)n * H + 2 * oy + (k >> 1)) * W + 2 * ox + (k & 1);
        union { uint4 u; bf16 h[8]; } v;
        v.u = *(const uint4*)(x + p * xPitch + c);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] += (float)v.h[j];
    }
    union { uint4 u; bf16 h[8]; } o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.h[j] = f2bf(s[j] * 0.25f);
    *(uint4*)(y + q * yPitch + c) = o.u;
}

// dX (n, H, W) = dY(n, H/2, W/2) / 4 (+ skip gradient when given)
__global__ __launch_bounds__(256) void k_pool2_bwd(const bf16* __restrict__ dy, uint32_t dyPitch, uint32_t C,
                                                   const bf16* __restrict__ skip, uint32_t skipPitch,
                                                   bf16* __restrict__ dx, uint32_t dxPitch, uint32_t N, uint32_t H,
                                                   uint32_t W) {
    const uint32_t groups = C / 8;
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)N * H * W * groups) return;
    const uint64_t p = e / groups;
    const uint32_t c = (uint32_t)(e % groups) * 8;
    const uint32_t xx = (uint32_t)(p % W), yy = (uint32_t)((p / W) % H), n = (uint32_t)(p / ((uint64_t)W * H));
    const uint64_t q = ((uint64_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2;
    union { uint4 u; bf16 h[8]; } g, s, o;
    g.u = *(const uint4*)(dy + q * dyPitch + c);
    s.u = skip ? *(const uint4*)(skip + p * skipPitch + c) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int j = 0; j < 8; ++j) o.h[j] = f2bf((float)g.h[j] * 0.25f + (float)s.h[j]);
    *(uint4*)(dx + p * dxPitch + c) = o.u;
}

// nearest 2x upsample: (n, H/2, W/2) -> (n, H, W)
__global__ __launch_bounds__(256) void k_up2_fwd(const bf16* __restrict__ x, uint32_t xPitch, uint32_t C,
                                                 bf16* __restrict__ y, uint32_t yPitch, uint32_t N, uint32_t H,
                                                 uint32_t W) {
    const uint32_t groups = C / 8;
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)N * H * W * groups) return;
    const uint64_t p = e / groups;
    const uint32_t c = (uint32_t)(e % groups) * 8;
    const uint32_t xx = (uint32_t)(p % W), yy = (uint32_t)((p / W) % H), n = (uint32_t)(p / ((uint64_t)W * H));
    const uint64_t q = ((uint64_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2;
    *(uint4*)(y + p * yPitch + c) = *(const uint4*)(x + q * xPitch + c);
}

// dX (n, H/2, W/2) = sum of the 4 dY it was copied to
__global__ __launch_bounds__(256) void k_up2_bwd(const bf16* __restrict__ dy, uint32_t dyPitch, uint32_t C,
                                                 bf16* __restrict__ dx, uint32_t dxPitch, uint32_t N, uint32_t H,
                                                 uint32_t W) {
    const uint32_t Ho = H / 2, Wo = W / 2, groups = C / 8;
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)N * Ho * Wo * groups) return;
    const uint64_t q = e / groups;
    const uint32_t c = (uint32_t)(e % groups) * 8;
    const uint32_t ox = (uint32_t)(q % Wo), oy = (uint32_t)((q / Wo) % Ho), n = (uint32_t)(q / ((uint64_t)Wo * Ho));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t p = ((uint64_t)n * H + 2 * oy + (k >> 1)) * W + 2 * ox + (k & 1);
        union { uint4 u; bf16 h[8]; } v;
        v.u = *(const uint4*)(dy + p * dyPitch + c);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] += (float)v.h[j];
    }
    union { uint4 u; bf16 h[8]; } o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.h[j] = f2bf(s[j]);
    *(uint4*)(dx + q * dxPitch + c) = o.u;
}

// ---------------------------------------------------- input / output / loss
// 9 guide channels into a 32-channel bf16 tensor: log(1 + max(colour, 0)),
// albedo, normal, zeros. Inputs are FP32 (n, H, W, 3), rows of srcW pixels
// starting at (sx0, sy0) of an image of srcH x srcW (crops for training).
// Positions past the source's last row or column repeat it (edge padding of
// images whose sides are not multiples of 8).
__global__ __launch_bounds__(256) void k_unet_pack(const float* __restrict__ color, const float* __restrict__ albedo,
                                                   const float* __restrict__ normal, const float* __restrict__ noise,
                                                   uint32_t srcH, uint32_t srcW,
                                                   const uint32_t* __restrict__ origins,
                                                   const float* __restrict__ feature, uint32_t numFeature,
                                                   bf16* __restrict__ out, uint32_t N, uint32_t H, uint32_t W) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (uint64_t)N * H * W) return;
    const uint32_t xx = (uint32_t)(p % W), yy = (uint32_t)((p / W) % H), n = (uint32_t)(p / ((uint64_t)W * H));
    uint32_t sx = xx, sy = yy, img = n;
    if (origins) {  // crop n: image origins[3n] at (origins[3n+1], origins[3n+2])
        img = origins[3 * n];
        sx += origins[3 * n + 1];
        sy += origins[3 * n + 2];
    }
    sx = min(sx, srcW - 1u);
    sy = min(sy, srcH - 1u);
    const uint64_t s = ((uint64_t)img * srcH + sy) * srcW + sx;
    float v[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) v[c] = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = log1pf(fmaxf(color[s * 3 + c], 0.0f));
        v[3 + c] = albedo[s * 3 + c];
        v[6 + c] = normal[s * 3 + c];
    }
    // per-image feature (the sample count, log2(spp) / 16): tells the
    // network how noisy its input is
    if (feature && img < numFeature) v[9] = feature[img];
    // per-pixel noise: the standard error of the mean y = (r + g + b) / 3
    // relative to 1 + y, about the noise of the log colour channels
    if (noise) {
        const float y = (color[s * 3 + 0] + color[s * 3 + 1] + color[s * 3 + 2]) / 3.0f;
        v[10] = log1pf(fmaxf(noise[s], 0.0f) / (1.0f + fmaxf(y, 0.0f)));
    }
    union { uint4 u[4]; bf16 h[32]; } o;
#pragma unroll
    for (int c = 0; c < 32; ++c) o.h[c] = f2bf(v[c]);
    uint4* dst = (uint4*)(out + p * 32);
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[k] = o.u[k];
}

// out = exp(q) - 1 for the first 3 channels, q = pred (+ base when given):
// the network predicts log(1 + L), or its correction to the input's
// log(1 + L) when base is the packed input
__global__ __launch_bounds__(256) void k_unet_unpack(const bf16* __restrict__ pred, uint32_t pitch,
                                                     const bf16* __restrict__ base, uint32_t basePitch,
                                                     float* __restrict__ out, uint64_t npix) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float q = (float)pred[p * pitch + c] + (base ? (float)base[p * basePitch + c] : 0.0f);
        out[p * 3 + c] = fmaxf(expm1f(q), 0.0f);
    }
}

// Per pixel and channel c < 3, with t the target and y = exp(pred) - 1:
//   mode 0: (pred - log(1 + t))^2            (log-space MSE)
//   mode 1: (y - t)^2 / (t^2 + 0.01)          (relMSE, the evaluation metric)
//   mode 2: (y - t)^2 / (y^2 + 0.01), the denominator held constant in the
//           gradient (Lehtinen et al. 2018, Noise2Noise)
//   mode 3: (y - t)^2 / (t2^2 + 0.01) with t2 = target2, a second target
//           over disjoint samples: numerator and denominator noise are
//           independent, so the minimiser is E[t] (mode 1 biases it low)
// averaged over pixels and 3 channels; dPred is its gradient (0 for the
// padding channels). Per-block partial loss sums.
__global__ __launch_bounds__(256) void k_unet_loss(const bf16* __restrict__ pred, uint32_t pitch,
                                                   const float* __restrict__ target, uint32_t srcH, uint32_t srcW,
                                                   const uint32_t* __restrict__ origins,
                                                   const float* __restrict__ target2,
                                                   const bf16* __restrict__ base, uint32_t basePitch,
                                                   bf16* __restrict__ dpred, uint32_t N, uint32_t H, uint32_t W,
                                                   uint32_t mode, float* __restrict__ partial) {
    __shared__ float red[256];
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t npix = (uint64_t)N * H * W;
    float l = 0.0f;
    if (p < npix) {
        const uint32_t xx = (uint32_t)(p % W), yy = (uint32_t)((p / W) % H), n = (uint32_t)(p / ((uint64_t)W * H));
        uint32_t sx = xx, sy = yy, img = n;
        if (origins) {
            img = origins[3 * n];
            sx += origins[3 * n + 1];
            sy += origins[3 * n + 2];
        }
        const uint64_t s = ((uint64_t)img * srcH + sy) * srcW + sx;
        const float scale = 2.0f / (float)(npix * 3);
        for (uint32_t c = 0; c < pitch; ++c) {
            float g = 0.0f;
            if (c < 3) {
                const float t = fmaxf(target[s * 3 + c], 0.0f);
                const float q = (float)pred[p * pitch + c] + (base ? (float)base[p * basePitch + c] : 0.0f);
                if (mode >= 1) {
                    const float e = expf(fminf(q, 12.0f));
                    const float y = e - 1.0f;
                    const float t2 = mode == 3 ? fmaxf(target2[s * 3 + c], 0.0f) : t;
                    const float inv = 1.0f / (mode == 2 ? y * y + 0.01f : t2 * t2 + 0.01f);
                    const float d = y - t;
                    l += d * d * inv;
                    g = q < 12.0f ? d * inv * e * scale : 0.0f;
                } else {
                    const float d = q - log1pf(t);
                    l += d * d;
                    g = d * scale;
                }
            }
            dpred[p * pitch + c] = f2bf(g);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ------------------------------------------------------ kernel prediction
// Kernel-predicting output (Bako et al. 2017, "Kernel-Predicting
// Convolutional Networks for Denoising Monte Carlo Renderings"), applied in
// the log domain as their specular branch is: channels 0 .. kKpnTaps - 1 of
// the network output are the logits of a (2 kKpnRadius + 1)^2 filter per
// pixel (row-major, dy then dx), normalised by a softmax and applied to
// l = log(1 + max(c, 0)) of the noisy colour around the pixel's source
// position; channels kKpnTaps .. kKpnTaps + 2 add a per-channel offset d
// (the log-domain average of noisy samples sits below the log of their
// mean). Output: exp(sum_k w_k l_k + d) - 1. Neighbours past the source
// image's edges repeat its edge pixels. Filtering log colours keeps an
// emitter's radiance from spreading into its dark neighbours at any weight.
constexpr int kKpnRadius = 3;
constexpr int kKpnSide = 2 * kKpnRadius + 1;
constexpr int kKpnTaps = kKpnSide * kKpnSide;
constexpr int kKpnChannels = kKpnTaps + 3;

struct SrcPix {
    uint32_t img, sx, sy;
};
// source image and position of output pixel p of n crops of h x w (origins,
// 3 per crop) or of whole images (origins NULL; positions past the last
// source row or column repeat it)
__device__ __forceinline__ SrcPix src_pix(uint64_t p, uint32_t H, uint32_t W, const uint32_t* __restrict__ origins,
                                          uint32_t srcH, uint32_t srcW) {
    const uint32_t xx = (uint32_t)(p % W), yy = (uint32_t)((p / W) % H), n = (uint32_t)(p / ((uint64_t)W * H));
    SrcPix r{n, xx, yy};
    if (origins) {
        r.img = origins[3 * n];
        r.sx += origins[3 * n + 1];
        r.sy += origins[3 * n + 2];
    }
    r.sx = min(r.sx, srcW - 1u);
    r.sy = min(r.sy, srcH - 1u);
    return r;
}

__device__ __forceinline__ const float* kpn_row(const float* color, SrcPix q, int dy, uint32_t srcH, uint32_t srcW) {
    const uint32_t y = (uint32_t)min(max((int)q.sy + dy, 0), (int)srcH - 1);
    return color + ((uint64_t)q.img * srcH + y) * srcW * 3;
}
__device__ __forceinline__ uint32_t kpn_col(SrcPix q, int dx, uint32_t srcW) {
    return (uint32_t)min(max((int)q.sx + dx, 0), (int)srcW - 1);
}

// Filtered log colour of one pixel: u = sum_k w_k l_k + d, with w the
// softmax of the logits; returns the softmax's max logit and 1 / sum exp.
__device__ __forceinline__ void kpn_filter(const bf16* __restrict__ z, const float* __restrict__ color, SrcPix q,
                                           uint32_t srcH, uint32_t srcW, float u[3], float* zmax, float* invSum) {
    float m = -INFINITY;
    for (int k = 0; k < kKpnTaps; ++k) m = fmaxf(m, (float)z[k]);
    float sum = 0.0f, a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int dy = -kKpnRadius; dy <= kKpnRadius; ++dy) {
        const float* row = kpn_row(color, q, dy, srcH, srcW);
        for (int dx = -kKpnRadius; dx <= kKpnRadius; ++dx) {
            const uint32_t x = kpn_col(q, dx, srcW);
            const float w = expf((float)z[(dy + kKpnRadius) * kKpnSide + dx + kKpnRadius] - m);
            sum += w;
            a0 += w * log1pf(fmaxf(row[x * 3 + 0], 0.0f));
            a1 += w * log1pf(fmaxf(row[x * 3 + 1], 0.0f));
            a2 += w * log1pf(fmaxf(row[x * 3 + 2], 0.0f));
        }
    }
    const float inv = 1.0f / sum;
    u[0] = a0 * inv + (float)z[kKpnTaps + 0];
    u[1] = a1 * inv + (float)z[kKpnTaps + 1];
    u[2] = a2 * inv + (float)z[kKpnTaps + 2];
    *zmax = m;
    *invSum = inv;
}

__global__ __launch_bounds__(256) void k_kpn_apply(const bf16* __restrict__ pred, uint32_t pitch,
                                                   const float* __restrict__ color, uint32_t srcH, uint32_t srcW,
                                                   const uint32_t* __restrict__ origins, float* __restrict__ out,
                                                   uint32_t N, uint32_t H, uint32_t W) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (uint64_t)N * H * W) return;
    const SrcPix q = src_pix(p, H, W, origins, srcH, srcW);
    float u[3], m, inv;
    kpn_filter(pred + p * pitch, color, q, srcH, srcW, u, &m, &inv);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[p * 3 + c] = fmaxf(expm1f(u[c]), 0.0f);
}

// The losses of k_unet_loss (modes 0..3) with q = u, the filtered log colour
// (y = exp(u) - 1), and their gradient with respect to the network output:
// G_c = dL/du_c; logits dL/dz_k = w_k sum_c G_c (l_kc - sum_j w_j l_jc);
// offsets dL/dd_c = G_c.
__global__ __launch_bounds__(256) void k_kpn_loss(const bf16* __restrict__ pred, uint32_t pitch,
                                                  const float* __restrict__ color, const float* __restrict__ target,
                                                  uint32_t srcH, uint32_t srcW, const uint32_t* __restrict__ origins,
                                                  const float* __restrict__ target2, bf16* __restrict__ dpred,
                                                  uint32_t N, uint32_t H, uint32_t W, uint32_t mode,
                                                  float* __restrict__ partial) {
    __shared__ float red[256];
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t npix = (uint64_t)N * H * W;
    float l = 0.0f;
    if (p < npix) {
        const SrcPix q = src_pix(p, H, W, origins, srcH, srcW);
        const bf16* z = pred + p * pitch;
        float u[3], m, inv;
        kpn_filter(z, color, q, srcH, srcW, u, &m, &inv);
        const uint64_t s = ((uint64_t)q.img * srcH + q.sy) * srcW + q.sx;
        const float scale = 2.0f / (float)(npix * 3);
        float G[3], mean[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            mean[c] = u[c] - (float)z[kKpnTaps + c];
            const float t = fmaxf(target[s * 3 + c], 0.0f);
            if (mode >= 1) {
                const float e = expf(fminf(u[c], 12.0f));
                const float y = e - 1.0f;
                const float t2 = mode == 3 ? fmaxf(target2[s * 3 + c], 0.0f) : t;
                const float den = 1.0f / (mode == 2 ? y * y + 0.01f : t2 * t2 + 0.01f);
                const float d = y - t;
                l += d * d * den;
                G[c] = u[c] < 12.0f ? d * den * e * scale : 0.0f;
            } else {
                const float d = u[c] - log1pf(t);
                l += d * d;
                G[c] = d * scale;
            }
        }
        bf16* dz = dpred + p * pitch;
        for (int dy = -kKpnRadius; dy <= kKpnRadius; ++dy) {
            const float* row = kpn_row(color, q, dy, srcH, srcW);
            for (int dx = -kKpnRadius; dx <= kKpnRadius; ++dx) {
                const uint32_t xx = kpn_col(q, dx, srcW);
                const int k = (dy + kKpnRadius) * kKpnSide + dx + kKpnRadius;
                const float w = expf((float)z[k] - m) * inv;
                const float e = G[0] * (log1pf(fmaxf(row[xx * 3 + 0], 0.0f)) - mean[0]) +
                                G[1] * (log1pf(fmaxf(row[xx * 3 + 1], 0.0f)) - mean[1]) +
                                G[2] * (log1pf(fmaxf(row[xx * 3 + 2], 0.0f)) - mean[2]);
                dz[k] = f2bf(w * e);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) dz[kKpnTaps + c] = f2bf(G[c]);
        for (uint32_t c = kKpnChannels; c < pitch; ++c) dz[c] = f2bf(0.0f);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------- weights
// FP32 master [9][cout][cin] -> bf16 forward layout and the tap-flipped
// transposed layout used by the data gradient.
__global__ __launch_bounds__(256) void k_weights_bf16(const float* __restrict__ w32, uint32_t cout, uint32_t cin,
                                                      bf16* __restrict__ w, bf16* __restrict__ wt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 9ull * cout * cin) return;
    const uint32_t ci = (uint32_t)(i % cin), co = (uint32_t)((i / cin) % cout), tap = (uint32_t)(i / ((uint64_t)cin * cout));
    const bf16 v = f2bf(w32[i]);
    w[i] = v;
    if (wt) wt[((uint64_t)(8 - tap) * cin + ci) * cout + co] = v;
}

// Adam (Kingma & Ba 2015) on FP32 parameters; bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, uint64_t n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

}  // namespace wrk

using namespace wrk;

static int conv_check(uint32_t cin, uint32_t cout, uint32_t xPitch, uint32_t yPitch) {
    WRK_REQUIRE(cin % 32 == 0 && cin > 0, "conv3x3: input channels must be a positive multiple of 32");
    WRK_REQUIRE(cout % 32 == 0 && cout > 0, "conv3x3: output channels must be a positive multiple of 32");
    WRK_REQUIRE(xPitch % 8 == 0 && yPitch % 8 == 0, "conv3x3: channel pitches must be multiples of 8");
    WRK_REQUIRE(xPitch >= cin && yPitch >= cout, "conv3x3: pitch smaller than the channel count");
    return WRK_OK;
}

// Resident blocks of a persistent launch: blocks per CU the kernel's
// registers and LDS allow, times the CUs, shared over gridY channel tiles.
// Queried once per device and kernel.
static uint32_t resident_blocks(const void* kernel, uint32_t gridY) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, uint32_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024u / gridY + 1u;
    uint32_t total = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find({dev, kernel});
        if (it != cache.end()) {
            total = it->second;
        } else {
            int cus = 0, perCu = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, 256, 0) != hipSuccess || cus <= 0 ||
                perCu <= 0)
                return 1024u / gridY + 1u;
            total = (uint32_t)cus * (uint32_t)perCu;
            cache[{dev, kernel}] = total;
        }
    }
    return total > gridY ? total / gridY : 1u;
}

template <int CT>
static void conv_fwd_launch(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* xUp, uint32_t upPitch,
                            uint32_t cUp, const uint16_t* w, uint32_t wCout, const float* bias, uint16_t* y,
                            uint32_t yPitch, uint32_t cout, uint32_t relu, uint16_t* yPool, uint32_t poolPitch,
                            uint32_t n, uint32_t h, uint32_t wd, hipStream_t stream, const uint16_t* base = nullptr,
                            uint32_t basePitch = 0, float* rgb = nullptr) {
    const uint32_t tilesX = ceil_div(wd, kTW), tilesY = ceil_div(h, kTH);
    const uint32_t coTiles = ceil_div(cout, 16 * CT);
    if (cin == 32 && cUp == 0) {
        const uint32_t tilesPerImage = tilesX * tilesY, nTiles = tilesPerImage * n;
        const uint32_t blocks = std::min(nTiles, resident_blocks((const void*)&k_conv_fwd_c32<CT>, coTiles));
        hipLaunchKernelGGL(k_conv_fwd_c32<CT>, dim3(blocks, coTiles), dim3(256), 0, stream, (const bf16*)x, xPitch,
                           (const bf16*)w, wCout, bias, (bf16*)y, yPitch, cout, relu, h, wd, tilesX, tilesPerImage,
                           nTiles, (bf16*)yPool, poolPitch, (const bf16*)base, basePitch, rgb);
    } else {
        hipLaunchKernelGGL(k_conv_fwd<CT>, dim3(tilesX * tilesY, coTiles, n), dim3(256), 0, stream, (const bf16*)x,
                           xPitch, cin, (const bf16*)xUp, upPitch, cUp, (const bf16*)w, wCout, bias, (bf16*)y, yPitch,
                           cout, relu, h, wd, tilesX, (bf16*)yPool, poolPitch);
    }
}

// Output channels [0, cout) of a convolution whose weights have wCout;
// input channels below cUp upsampled from xUp; the 2x2 average pool of
// the output into yPool when it is given.
static int conv_fwd(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* xUp, uint32_t upPitch,
                    uint32_t cUp, const uint16_t* w, uint32_t wCout, const float* bias, uint16_t* y, uint32_t yPitch,
                    uint32_t cout, uint32_t relu, uint16_t* yPool, uint32_t poolPitch, uint32_t n, uint32_t h,
                    uint32_t wd, void* stream) {
    if (n == 0 || h == 0 || wd == 0) return WRK_OK;
    WRK_REQUIRE(n <= 65535, "wrk_conv3x3_fwd: batch too large");
    WRK_REQUIRE((uint64_t)n * ceil_div(h, kTH) * ceil_div(wd, kTW) <= 0xFFFFFFFFull,
                "wrk_conv3x3_fwd: too many pixel tiles");
    const hipStream_t s = as_stream(stream);
    if (cout % 64 == 0)
        conv_fwd_launch<4>(x, xPitch, cin, xUp, upPitch, cUp, w, wCout, bias, y, yPitch, cout, relu, yPool, poolPitch, n,
                           h, wd, s);
    else if (cout % 32 == 0)
        conv_fwd_launch<2>(x, xPitch, cin, xUp, upPitch, cUp, w, wCout, bias, y, yPitch, cout, relu, yPool, poolPitch, n,
                           h, wd, s);
    else
        conv_fwd_launch<1>(x, xPitch, cin, xUp, upPitch, cUp, w, wCout, bias, y, yPitch, cout, relu, yPool, poolPitch, n,
                           h, wd, s);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_conv3x3_fwd(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w,
                               const float* bias, uint16_t* y, uint32_t yPitch, uint32_t cout, uint32_t relu,
                               uint32_t n, uint32_t h, uint32_t wd, void* stream) {
    WRK_REQUIRE(x && w && y, "wrk_conv3x3_fwd: null argument");
    if (int rc = conv_check(cin, cout, xPitch, yPitch)) return rc;
    return conv_fwd(x, xPitch, cin, nullptr, 0, 0, w, cout, bias, y, yPitch, cout, relu, nullptr, 0, n, h, wd,
                    stream);
}

extern "C" int wrk_conv3x3_fwd_first(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w,
                                     const float* bias, uint16_t* y, uint32_t yPitch, uint32_t cout,
                                     uint32_t coutUsed, uint32_t relu, uint32_t n, uint32_t h, uint32_t wd,
                                     void* stream) {
    WRK_REQUIRE(x && w && y, "wrk_conv3x3_fwd_first: null argument");
    if (int rc = conv_check(cin, cout, xPitch, yPitch)) return rc;
    WRK_REQUIRE(coutUsed > 0 && coutUsed <= cout, "wrk_conv3x3_fwd_first: coutUsed must be in 1..cout");
    const uint32_t computed = std::min(cout, (coutUsed + 15u) / 16u * 16u);
    return conv_fwd(x, xPitch, cin, nullptr, 0, 0, w, cout, bias, y, yPitch, computed, relu, nullptr, 0, n, h, wd,
                    stream);
}

extern "C" int wrk_conv3x3_fwd_up(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* xUp,
                                  uint32_t upPitch, uint32_t cUp, const uint16_t* w, const float* bias, uint16_t* y,
                                  uint32_t yPitch, uint32_t cout, uint32_t relu, uint32_t n, uint32_t h, uint32_t wd,
                                  void* stream) {
    WRK_REQUIRE(x && xUp && w && y, "wrk_conv3x3_fwd_up: null argument");
    if (int rc = conv_check(cin, cout, xPitch, yPitch)) return rc;
    WRK_REQUIRE(cUp > 0 && cUp <= cin && cUp % 32 == 0, "wrk_conv3x3_fwd_up: cUp must be a multiple of 32 in 32..cin");
    WRK_REQUIRE(upPitch % 8 == 0 && upPitch >= cUp, "wrk_conv3x3_fwd_up: bad upsampled-input pitch");
    WRK_REQUIRE(h % 2 == 0 && wd % 2 == 0, "wrk_conv3x3_fwd_up: height and width must be even");
    return conv_fwd(x, xPitch, cin, xUp, upPitch, cUp, w, cout, bias, y, yPitch, cout, relu, nullptr, 0, n, h, wd,
                    stream);
}

extern "C" int wrk_conv3x3_fwd_unpack(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w,
                                      const float* bias, uint32_t cout, const uint16_t* base, uint32_t basePitch,
                                      float* out, uint32_t n, uint32_t h, uint32_t wd, void* stream) {
    WRK_REQUIRE(x && w && base && out, "wrk_conv3x3_fwd_unpack: null argument");
    WRK_REQUIRE(cin == 32, "wrk_conv3x3_fwd_unpack: 32 input channels");
    if (int rc = conv_check(cin, cout, xPitch, cout)) return rc;
    WRK_REQUIRE(basePitch >= 3, "wrk_conv3x3_fwd_unpack: base pitch below 3");
    if (n == 0 || h == 0 || wd == 0) return WRK_OK;
    WRK_REQUIRE(n <= 65535, "wrk_conv3x3_fwd_unpack: batch too large");
    WRK_REQUIRE((uint64_t)n * ceil_div(h, kTH) * ceil_div(wd, kTW) <= 0xFFFFFFFFull,
                "wrk_conv3x3_fwd_unpack: too many pixel tiles");
    conv_fwd_launch<1>(x, xPitch, cin, nullptr, 0, 0, w, cout, bias, nullptr, 16, 16, 0u, nullptr, 0, n, h, wd,
                       as_stream(stream), base, basePitch, out);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_conv3x3_fwd_pool(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w,
                                    const float* bias, uint16_t* y, uint32_t yPitch, uint32_t cout, uint32_t relu,
                                    uint16_t* yPool, uint32_t poolPitch, uint32_t n, uint32_t h, uint32_t wd,
                                    void* stream) {
    WRK_REQUIRE(x && w && y && yPool, "wrk_conv3x3_fwd_pool: null argument");
    if (int rc = conv_check(cin, cout, xPitch, yPitch)) return rc;
    WRK_REQUIRE(poolPitch % 8 == 0 && poolPitch >= cout, "wrk_conv3x3_fwd_pool: bad pooled-output pitch");
    WRK_REQUIRE(h % 2 == 0 && wd % 2 == 0, "wrk_conv3x3_fwd_pool: height and width must be even");
    return conv_fwd(x, xPitch, cin, nullptr, 0, 0, w, cout, bias, y, yPitch, cout, relu, yPool, poolPitch, n, h, wd,
                    stream);
}

static void wgrad_plan(uint32_t cin, uint32_t cout, uint32_t n, uint32_t h, uint32_t w, uint32_t* splits,
                       uint32_t* tilesPerBlock, uint32_t* tilesX, uint32_t* tilesPerImage) {
    *tilesX = ceil_div(w, kGW);
    *tilesPerImage = *tilesX * ceil_div(h, kGH);
    const uint64_t total = (uint64_t)n * *tilesPerImage;
    const uint32_t pairs = (cin / 32) * (cout / 32);
    uint64_t want = pairs >= 1024 ? 1 : 1024 / pairs;  // about 1024 blocks in all
    if (want > total) want = total;
    if (want < 1) want = 1;
    *tilesPerBlock = (uint32_t)((total + want - 1) / want);
    *splits = (uint32_t)((total + *tilesPerBlock - 1) / *tilesPerBlock);
}

extern "C" uint64_t wrk_conv3x3_wgrad_scratch_bytes(uint32_t cin, uint32_t cout, uint32_t n, uint32_t h, uint32_t w) {
    if (cin == 0 || cout == 0 || n == 0 || h == 0 || w == 0) return 0;
    uint32_t splits, tpb, tx, tpi;
    wgrad_plan(cin, cout, n, h, w, &splits, &tpb, &tx, &tpi);
    const uint64_t wBytes = (uint64_t)splits * 9 * cout * cin * 4;
    const uint64_t bBytes = 1024ull * cout * 4;
    return wBytes > bBytes ? wBytes : bBytes;
}

extern "C" int wrk_conv3x3_wgrad(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* dz,
                                 uint32_t dzPitch, uint32_t cout, uint32_t n, uint32_t h, uint32_t w, float* dw,
                                 float* db, void* scratch, uint64_t scratchBytes, void* stream) {
    WRK_REQUIRE(x && dz && dw && scratch, "wrk_conv3x3_wgrad: null argument");
    if (int rc = conv_check(cin, cout, xPitch, dzPitch)) return rc;
    WRK_REQUIRE(n > 0 && h > 0 && w > 0, "wrk_conv3x3_wgrad: empty tensor");
    WRK_REQUIRE(scratchBytes >= wrk_conv3x3_wgrad_scratch_bytes(cin, cout, n, h, w), "wrk_conv3x3_wgrad: scratch too small");
    WRK_REQUIRE(!db || cout <= 256, "wrk_conv3x3_wgrad: bias gradient needs cout <= 256");
    uint32_t splits, tpb, tilesX, tpi;
    wgrad_plan(cin, cout, n, h, w, &splits, &tpb, &tilesX, &tpi);
    float* partial = (float*)scratch;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_conv_wgrad, dim3((cin / 32) * (cout / 32), splits), dim3(256), 0, s, (const bf16*)x, xPitch,
                       cin, (const bf16*)dz, dzPitch, cout, n, h, w, tilesX, tpi, tpb, partial);
    const uint64_t nw = 9ull * cout * cin;
    hipLaunchKernelGGL(k_reduce_splits, dim3(ceil_div(nw, 256)), dim3(256), 0, s, partial, splits, nw, dw);
    if (db) {
        // the bias reduction reuses the scratch after the weight reduction
        const uint64_t npix = (uint64_t)n * h * w;
        const uint64_t per = (npix + 1023) / 1024;
        const uint32_t blocks = (uint32_t)((npix + per - 1) / per);
        hipLaunchKernelGGL(k_bias_grad, dim3(blocks), dim3(256), 0, s, (const bf16*)dz, dzPitch, cout, npix, per,
                           partial);
        hipLaunchKernelGGL(k_reduce_splits, dim3(ceil_div(cout, 256)), dim3(256), 0, s, partial, blocks,
                           (uint64_t)cout, db);
    }
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_relu_bwd(uint16_t* dy, uint32_t dyPitch, const uint16_t* y, uint32_t yPitch, uint32_t c,
                            uint64_t npix, void* stream) {
    WRK_REQUIRE(dy && y && c % 8 == 0 && dyPitch % 8 == 0 && yPitch % 8 == 0, "wrk_relu_bwd: bad argument");
    const uint64_t n = npix * (c / 8);
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_relu_bwd, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), (bf16*)dy, dyPitch,
                       (const bf16*)y, yPitch, c, npix);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_pool2_fwd(const uint16_t* x, uint32_t xPitch, uint32_t c, uint16_t* y, uint32_t yPitch,
                             uint32_t n, uint32_t h, uint32_t w, void* stream) {
    WRK_REQUIRE(x && y && c % 8 == 0 && h % 2 == 0 && w % 2 == 0, "wrk_pool2_fwd: bad argument");
    const uint64_t e = (uint64_t)n * (h / 2) * (w / 2) * (c / 8);
    if (e == 0) return WRK_OK;
    hipLaunchKernelGGL(k_pool2_fwd, dim3(ceil_div(e, 256)), dim3(256), 0, as_stream(stream), (const bf16*)x, xPitch,
                       c, (bf16*)y, yPitch, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_pool2_bwd(const uint16_t* dy, uint32_t dyPitch, uint32_t c, const uint16_t* skip,
                             uint32_t skipPitch, uint16_t* dx, uint32_t dxPitch, uint32_t n, uint32_t h, uint32_t w,
                             void* stream) {
    WRK_REQUIRE(dy && dx && c % 8 == 0 && h % 2 == 0 && w % 2 == 0, "wrk_pool2_bwd: bad argument");
    const uint64_t e = (uint64_t)n * h * w * (c / 8);
    if (e == 0) return WRK_OK;
    hipLaunchKernelGGL(k_pool2_bwd, dim3(ceil_div(e, 256)), dim3(256), 0, as_stream(stream), (const bf16*)dy,
                       dyPitch, c, (const bf16*)skip, skipPitch, (bf16*)dx, dxPitch, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_up2_fwd(const uint16_t* x, uint32_t xPitch, uint32_t c, uint16_t* y, uint32_t yPitch, uint32_t n,
                           uint32_t h, uint32_t w, void* stream) {
    WRK_REQUIRE(x && y && c % 8 == 0 && h % 2 == 0 && w % 2 == 0, "wrk_up2_fwd: bad argument");
    const uint64_t e = (uint64_t)n * h * w * (c / 8);
    if (e == 0) return WRK_OK;
    hipLaunchKernelGGL(k_up2_fwd, dim3(ceil_div(e, 256)), dim3(256), 0, as_stream(stream), (const bf16*)x, xPitch, c,
                       (bf16*)y, yPitch, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_up2_bwd(const uint16_t* dy, uint32_t dyPitch, uint32_t c, uint16_t* dx, uint32_t dxPitch,
                           uint32_t n, uint32_t h, uint32_t w, void* stream) {
    WRK_REQUIRE(dy && dx && c % 8 == 0 && h % 2 == 0 && w % 2 == 0, "wrk_up2_bwd: bad argument");
    const uint64_t e = (uint64_t)n * (h / 2) * (w / 2) * (c / 8);
    if (e == 0) return WRK_OK;
    hipLaunchKernelGGL(k_up2_bwd, dim3(ceil_div(e, 256)), dim3(256), 0, as_stream(stream), (const bf16*)dy, dyPitch,
                       c, (bf16*)dx, dxPitch, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_unet_pack(const float* color, const float* albedo, const float* normal, const float* noise,
                             uint32_t srcH, uint32_t srcW, const uint32_t* origins, const float* feature,
                             uint32_t numFeature, uint16_t* out, uint32_t n, uint32_t h, uint32_t w, void* stream) {
    WRK_REQUIRE(color && albedo && normal && out, "wrk_unet_pack: null argument");
    WRK_REQUIRE(origins || (h >= srcH && w >= srcW), "wrk_unet_pack: crops need origins");
    WRK_REQUIRE(srcH > 0 && srcW > 0, "wrk_unet_pack: empty source");
    const uint64_t np = (uint64_t)n * h * w;
    if (np == 0) return WRK_OK;
    hipLaunchKernelGGL(k_unet_pack, dim3(ceil_div(np, 256)), dim3(256), 0, as_stream(stream), color, albedo, normal,
                       noise, srcH, srcW, origins, feature, numFeature, (bf16*)out, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_unet_unpack(const uint16_t* pred, uint32_t pitch, const uint16_t* base, uint32_t basePitch,
                               float* out, uint64_t npix, void* stream) {
    WRK_REQUIRE(pred && out && pitch >= 3 && (!base || basePitch >= 3), "wrk_unet_unpack: bad argument");
    if (npix == 0) return WRK_OK;
    hipLaunchKernelGGL(k_unet_unpack, dim3(ceil_div(npix, 256)), dim3(256), 0, as_stream(stream), (const bf16*)pred,
                       pitch, (const bf16*)base, basePitch, out, npix);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" uint64_t wrk_unet_loss_scratch_bytes(uint32_t n, uint32_t h, uint32_t w) {
    return (uint64_t)ceil_div((uint64_t)n * h * w, 256) * 4;
}

extern "C" int wrk_unet_loss(const uint16_t* pred, uint32_t pitch, const float* target, uint32_t srcH, uint32_t srcW,
                             const uint32_t* origins, const float* target2, const uint16_t* base, uint32_t basePitch,
                             uint16_t* dpred, uint32_t n, uint32_t h, uint32_t w, uint32_t mode, float* partial,
                             void* stream) {
    WRK_REQUIRE(pred && target && dpred && partial && pitch >= 3 && (!base || basePitch >= 3),
                "wrk_unet_loss: bad argument");
    WRK_REQUIRE(mode <= 3 && (mode != 3 || target2), "wrk_unet_loss: mode 3 needs target2");
    WRK_REQUIRE(origins || (srcH == h && srcW == w), "wrk_unet_loss: crops need origins");
    const uint64_t np = (uint64_t)n * h * w;
    if (np == 0) return WRK_OK;
    hipLaunchKernelGGL(k_unet_loss, dim3(ceil_div(np, 256)), dim3(256), 0, as_stream(stream), (const bf16*)pred,
                       pitch, target, srcH, srcW, origins, target2, (const bf16*)base, basePitch, (bf16*)dpred, n, h, w,
                       mode, partial);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" uint32_t wrk_kpn_taps(void) { return (uint32_t)kKpnTaps; }

extern "C" int wrk_kpn_apply(const uint16_t* pred, uint32_t pitch, const float* color, uint32_t srcH, uint32_t srcW,
                             const uint32_t* origins, float* out, uint32_t n, uint32_t h, uint32_t w, void* stream) {
    WRK_REQUIRE(pred && color && out, "wrk_kpn_apply: null argument");
    WRK_REQUIRE(pitch >= (uint32_t)kKpnChannels, "wrk_kpn_apply: pitch below the filter's channel count");
    WRK_REQUIRE(srcH > 0 && srcW > 0, "wrk_kpn_apply: empty source image");
    WRK_REQUIRE(origins || (h >= srcH && w >= srcW), "wrk_kpn_apply: whole images must cover the source");
    const uint64_t npix = (uint64_t)n * h * w;
    if (npix == 0) return WRK_OK;
    hipLaunchKernelGGL(k_kpn_apply, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, as_stream(stream),
                       (const bf16*)pred, pitch, color, srcH, srcW, origins, out, n, h, w);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_kpn_loss(const uint16_t* pred, uint32_t pitch, const float* color, const float* target,
                            uint32_t srcH, uint32_t srcW, const uint32_t* origins, const float* target2,
                            uint16_t* dpred, uint32_t n, uint32_t h, uint32_t w, uint32_t mode, float* partial,
                            void* stream) {
    WRK_REQUIRE(pred && color && target && dpred && partial, "wrk_kpn_loss: null argument");
    WRK_REQUIRE(pitch >= (uint32_t)kKpnChannels, "wrk_kpn_loss: pitch below the filter's channel count");
    WRK_REQUIRE(mode <= 3, "wrk_kpn_loss: unknown loss mode");
    WRK_REQUIRE(mode != 3 || target2, "wrk_kpn_loss: mode 3 needs target2");
    WRK_REQUIRE(srcH > 0 && srcW > 0, "wrk_kpn_loss: empty source image");
    const uint64_t npix = (uint64_t)n * h * w;
    if (npix == 0) return WRK_OK;
    hipLaunchKernelGGL(k_kpn_loss, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, as_stream(stream),
                       (const bf16*)pred, pitch, color, target, srcH, srcW, origins, target2, (bf16*)dpred, n, h, w,
                       mode, partial);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_weights_bf16(const float* w32, uint32_t cout, uint32_t cin, uint16_t* w, uint16_t* wt,
                                void* stream) {
    WRK_REQUIRE(w32 && w, "wrk_weights_bf16: null argument");
    const uint64_t n = 9ull * cout * cin;
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_weights_bf16, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), w32, cout, cin,
                       (bf16*)w, (bf16*)wt);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                        float eps, uint32_t step, void* stream) {
    WRK_REQUIRE(p && g && m && v && step > 0, "wrk_adam: bad argument");
    if (n == 0) return WRK_OK;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), p, g, m, v, n, lr, b1, b2,
                       eps, bc1, bc2);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}
