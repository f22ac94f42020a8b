// wrk_device.h — device-side path-tracing math for gfx950.
//
// Restates libWR's kernel math (libWR/kernels/sampling.cuh, ggx.cuh,
// materials/*.cuh, utils/*.cuh) in FP32 with the reference's operation
// order. All translation units that include this header are compiled with
// -ffp-contract=off, so every a*b+c below rounds twice, exactly like the
// reference source reads; fmaf appears only where a fused operation is
// intended on every platform. Deliberate deviations from the reference:
//   * normalize() divides by sqrtf (IEEE) instead of multiplying by the
//     approximate rsqrtf of --use_fast_math (vector_math.cuh:47-50);
//   * sinf/cosf are wr_sincos() (FP64 reduction, FP32 Taylor/Horner with
//     fmaf); powf(x,5) is repeated multiplication;
//   * textures are sampled in FP32 software bilinear with wrap addressing
//     instead of CUDA's hardware filter (scene.cpp:403-410).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wrk.h"

namespace wrk {

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ f3 operator*(float a, f3 b) { return {a * b.x, a * b.y, a * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// vector_math.cuh:37-50
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(f3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ f3 normalize(f3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

constexpr float kPi = 3.14159265f;

// sin and cos: reduction by pi/2 in FP64, Taylor polynomials on
// [-pi/4, pi/4] in FP32 Horner form with fmaf.
__device__ __forceinline__ void wr_sincos(float x, float* s, float* c) {
    double xd = (double)x;
    double kd = rint(xd * 0.63661977236758134308);
    double rd = fma(-kd, 1.57079632679489661923, xd);
    float r = (float)rd;
    int q = ((int)kd) & 3;
    float r2 = r * r;
    float ps = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);
    ps = fmaf(r2, ps, 8.3333333e-3f);
    ps = fmaf(r2, ps, -1.6666667e-1f);
    float sn = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);
    pc = fmaf(r2, pc, -1.3888889e-3f);
    pc = fmaf(r2, pc, 4.1666668e-2f);
    pc = fmaf(r2, pc, -0.5f);
    float cs = fmaf(r2, pc, 1.0f);
    float so = (q & 1) ? cs : sn;
    float co = (q & 1) ? sn : cs;
    *s = (q == 2 || q == 3) ? -so : so;
    *c = (q == 1 || q == 2) ? -co : co;
}
__device__ __forceinline__ float pow5(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x; }

// ----------------------------------------------------------------- RNG
// sampling.cuh:17-38
__device__ __forceinline__ uint32_t pcg_hash2(uint32_t seed, uint32_t dim) {
    uint32_t state = seed * 747796405u + 2891336453u + dim * 1013904223u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float rnd_dim(uint32_t seed, uint32_t dim) {
    return float(pcg_hash2(seed, dim)) / 4294967296.0f;
}
// utils/atomic_ops.cuh:16-19
__device__ __forceinline__ float randf(uint32_t& seed) {
    seed = seed * 1664525u + 1013904223u;
    return float(seed >> 8) / 16777216.0f;
}

// ------------------------------------------------------------- textures
// FP32 bilinear, wrap addressing, texel centres at integer + 0.5.
__device__ __forceinline__ f3 sample_tex(const wrk_texture& tex, float u, float v) {
    if (tex.texels == 0 || tex.width == 0 || tex.height == 0) return mk(1.0f, 1.0f, 1.0f);
    const uint8_t* px = (const uint8_t*)tex.texels;
    float uf = u - floorf(u);
    float vf = v - floorf(v);
    if (!(uf >= 0.0f && uf < 1.0f)) uf = 0.0f;
    if (!(vf >= 0.0f && vf < 1.0f)) vf = 0.0f;
    float x = uf * (float)tex.width - 0.5f;
    float y = vf * (float)tex.height - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float fx = x - x0, fy = y - y0;
    int ix = (int)x0, iy = (int)y0;
    int w = (int)tex.width, h = (int)tex.height;
    int i0 = ix < 0 ? ix + w : ix;
    int j0 = iy < 0 ? iy + h : iy;
    int i1 = i0 + 1 == w ? 0 : i0 + 1;
    int j1 = j0 + 1 == h ? 0 : j0 + 1;
    uchar4 t00 = *(const uchar4*)(px + ((size_t)j0 * w + i0) * 4);
    uchar4 t10 = *(const uchar4*)(px + ((size_t)j0 * w + i1) * 4);
    uchar4 t01 = *(const uchar4*)(px + ((size_t)j1 * w + i0) * 4);
    uchar4 t11 = *(const uchar4*)(px + ((size_t)j1 * w + i1) * 4);
    float a00[3] = {float(t00.x) / 255.0f, float(t00.y) / 255.0f, float(t00.z) / 255.0f};
    float a10[3] = {float(t10.x) / 255.0f, float(t10.y) / 255.0f, float(t10.z) / 255.0f};
    float a01[3] = {float(t01.x) / 255.0f, float(t01.y) / 255.0f, float(t01.z) / 255.0f};
    float a11[3] = {float(t11.x) / 255.0f, float(t11.y) / 255.0f, float(t11.z) / 255.0f};
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float top = a00[c] * (1.0f - fx) + a10[c] * fx;
        float bot = a01[c] * (1.0f - fx) + a11[c] * fx;
        o[c] = top * (1.0f - fy) + bot * fy;
    }
    return mk(o[0], o[1], o[2]);
}

// ------------------------------------------------------------- BSDF math
// sampling.cuh:44-60
__device__ __forceinline__ void make_frame(f3 n, f3& t, f3& b) {
    f3 up = fabsf(n.y) < 0.999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    t = normalize(cross(up, n));
    b = cross(n, t);
}
__device__ __forceinline__ f3 to_world(f3 v, f3 n, f3 t, f3 b) {
    return mk(t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z,
              t.z * v.x + b.z * v.y + n.z * v.z);
}
// sampling.cuh:66-75
__device__ __forceinline__ f3 sample_cosine(float u1, float u2, float& pdf) {
    float r = sqrtf(u1);
    float theta = (2.0f * kPi) * u2;
    float sn, cs;
    wr_sincos(theta, &sn, &cs);
    float x = r * cs;
    float y = r * sn;
    float z = sqrtf(fmaxf(0.0f, 1.0f - u1));
    pdf = z / kPi;
    return mk(x, y, z);
}
// sampling.cuh:145-149
__device__ __forceinline__ float power_heuristic(float a, float b) {
    float p1 = a * a;
    float p2 = b * b;
    return p1 / (p1 + p2);
}
// sampling.cuh:199-210
__device__ __forceinline__ f3 clamp_fireflies(f3 L, float maxL) {
    if (maxL <= 0.0f) return L;
    float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
    if (lum > maxL) {
        float scale = maxL / lum;
        return mk(L.x * scale, L.y * scale, L.z * scale);
    }
    return L;
}
// utils/ray_offset.cuh:9-21
__device__ __forceinline__ f3 offset_origin(f3 pos, f3 n, f3 dir) {
    const float off = 1e-4f;
    f3 od = dot(dir, n) > 0.0f ? n : -n;
    return mk(pos.x + od.x * off, pos.y + od.y * off, pos.z + od.z * off);
}
// material_common.cuh:14-29
__device__ __forceinline__ float asym_correction(f3 wi, f3 sn, f3 gn) {
    float cs = fabsf(dot(wi, sn));
    float cg = fabsf(dot(wi, gn));
    if (cs < 1e-6f || cg < 1e-6f) return 1.0f;
    return cg / cs;
}
// utils/fresnel.cuh:8-16
__device__ __forceinline__ float fresnel_dielectric(float cosI, float etaI, float etaT) {
    float sinT2 = etaI / etaT * etaI / etaT * (1.0f - cosI * cosI);
    if (sinT2 > 1.0f) return 1.0f;
    float cosT = sqrtf(1.0f - sinT2);
    float rs = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    float rp = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    return (rs * rs + rp * rp) * 0.5f;
}
// ggx.cuh:18-30
__device__ __forceinline__ float ggx_D(float NoH, float alpha) {
    float a2 = alpha * alpha;
    float NoH2 = NoH * NoH;
    float denom = NoH2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * denom * denom);
}
__device__ __forceinline__ float ggx_G1(float NoV, float alpha) {
    float a2 = alpha * alpha;
    float NoV2 = NoV * NoV;
    return 2.0f * NoV / (NoV + sqrtf(a2 + (1.0f - a2) * NoV2));
}
// ggx.cuh:113-176 with ggx.cuh:38-63
__device__ __forceinline__ f3 eval_ggx(f3 wo, f3 wi, f3 n, f3 albedo, float rough, float metal, float ior) {
    float NoV = fmaxf(0.0f, dot(n, wo));
    float NoL = fmaxf(0.0f, dot(n, wi));
    if (NoV < 1e-5f || NoL < 1e-5f) return mk(0.0f, 0.0f, 0.0f);
    f3 h = normalize(mk(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z));
    float NoH = fmaxf(0.0f, dot(n, h));
    float VoH = fmaxf(0.0f, dot(wo, h));
    float alpha = rough * rough;
    float D = ggx_D(NoH, alpha);
    float G = ggx_G1(NoV, alpha) * ggx_G1(NoL, alpha);
    float f0 = (ior - 1.0f) / (ior + 1.0f);
    f0 = f0 * f0;
    f3 F0 = mk(f0 * (1.0f - metal) + albedo.x * metal, f0 * (1.0f - metal) + albedo.y * metal,
               f0 * (1.0f - metal) + albedo.z * metal);
    float f = pow5(1.0f - VoH);
    f3 F = mk(F0.x + (1.0f - F0.x) * f, F0.y + (1.0f - F0.y) * f, F0.z + (1.0f - F0.z) * f);
    f3 spec = mk(D * G * F.x / (4.0f * NoV * NoL), D * G * F.y / (4.0f * NoV * NoL),
                 D * G * F.z / (4.0f * NoV * NoL));
    f3 kD = mk((1.0f - F.x) * (1.0f - metal), (1.0f - F.y) * (1.0f - metal), (1.0f - F.z) * (1.0f - metal));
    f3 diff = mk(kD.x * albedo.x / kPi, kD.y * albedo.y / kPi, kD.z * albedo.z / kPi);
    return mk(diff.x + spec.x, diff.y + spec.y, diff.z + spec.z);
}
// ggx.cuh:182-215
__device__ __forceinline__ float ggx_pdf(f3 wo, f3 wi, f3 n, float rough) {
    float NoV = fmaxf(0.0f, dot(n, wo));
    float NoL = fmaxf(0.0f, dot(n, wi));
    if (NoV < 1e-5f || NoL < 1e-5f) return 0.0f;
    f3 h = normalize(mk(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z));
    float NoH = fmaxf(0.0f, dot(n, h));
    float VoH = fmaxf(0.0f, dot(wo, h));
    if (VoH < 1e-5f) return 0.0f;
    float alpha = rough * rough;
    float G1 = ggx_G1(NoV, alpha);
    float D = ggx_D(NoH, alpha);
    return G1 * D / (4.0f * NoV);
}
// ggx.cuh:222-317 (libVLR-style VNDF sampling)
__device__ __forceinline__ void sample_ggx(f3 wo, f3 n, f3 tg, f3 bt, float rough, float u0, float u1,
                                           f3& wi, float& pdf) {
    float alpha = rough * rough;
    f3 v = mk(dot(wo, tg), dot(wo, bt), dot(wo, n));
    f3 sv = normalize(mk(alpha * v.x, alpha * v.y, v.z));
    float dist2D = sqrtf(sv.x * sv.x + sv.y * sv.y);
    f3 T1, T2;
    if (dist2D > 1e-6f) {
        float rec = 1.0f / dist2D;
        T1 = mk(sv.y * rec, -sv.x * rec, 0.0f);
        T2 = mk(T1.y * sv.z, -T1.x * sv.z, dist2D);
    } else {
        T1 = mk(1.0f, 0.0f, 0.0f);
        T2 = mk(0.0f, 1.0f, 0.0f);
    }
    float a = 1.0f / (1.0f + sv.z);
    float r = sqrtf(u0);
    float phi = kPi * ((u1 < a) ? u1 / a : 1.0f + (u1 - a) / (1.0f - a));
    float sinPhi, cosPhi;
    wr_sincos(phi, &sinPhi, &cosPhi);
    float P1 = r * cosPhi;
    float P2 = r * sinPhi * ((u1 < a) ? 1.0f : sv.z);
    float pp = P1 * P1 + P2 * P2;
    float sq = sqrtf(fmaxf(0.0f, 1.0f - pp));
    f3 mr = mk(P1 * T1.x + P2 * T2.x + sq * sv.x, P1 * T1.y + P2 * T2.y + sq * sv.y,
               P1 * T1.z + P2 * T2.z + sq * sv.z);
    mr = normalize(mk(alpha * mr.x, alpha * mr.y, mr.z));
    f3 h = mk(tg.x * mr.x + bt.x * mr.y + n.x * mr.z, tg.y * mr.x + bt.y * mr.y + n.y * mr.z,
              tg.z * mr.x + bt.z * mr.y + n.z * mr.z);
    float VoH = dot(wo, h);
    if (VoH <= 0.0f) { pdf = 0.0f; return; }
    wi = mk(2.0f * VoH * h.x - wo.x, 2.0f * VoH * h.y - wo.y, 2.0f * VoH * h.z - wo.z);
    float NoL = dot(n, wi);
    if (NoL <= 0.0f) { pdf = 0.0f; return; }
    float NoV = fmaxf(0.0f, dot(n, wo));
    float NoH = fmaxf(0.0f, dot(n, h));
    float G1 = ggx_G1(NoV, alpha);
    float D = ggx_D(NoH, alpha);
    pdf = G1 * D / (4.0f * NoV);
}

// --------------------------------------------------- watertight triangle
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2013.
struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float o[3];
};
__device__ __forceinline__ RayPre ray_pre(float4 o, float4 d) {
    RayPre p;
    float dd[3] = {d.x, d.y, d.z};
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (dd[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.Sx = dd[kx] / dd[kz];
    p.Sy = dd[ky] / dd[kz];
    p.Sz = 1.0f / dd[kz];
    p.o[0] = o.x; p.o[1] = o.y; p.o[2] = o.z;
    return p;
}
__device__ __forceinline__ float sel3(const float* v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }
// Returns t when hit with tmin <= t <= tmax, else a negative sentinel flag.
__device__ __forceinline__ bool intersect_tri(const RayPre& r, float4 v0, float4 v1, float4 v2, float tmin,
                                              float tmax, float* tout) {
    float A[3] = {v0.x - r.o[0], v0.y - r.o[1], v0.z - r.o[2]};
    float B[3] = {v1.x - r.o[0], v1.y - r.o[1], v1.z - r.o[2]};
    float C[3] = {v2.x - r.o[0], v2.y - r.o[1], v2.z - r.o[2]};
    float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
    float Ax = sel3(A, r.kx) - r.Sx * Akz;
    float Ay = sel3(A, r.ky) - r.Sy * Akz;
    float Bx = sel3(B, r.kx) - r.Sx * Bkz;
    float By = sel3(B, r.ky) - r.Sy * Bkz;
    float Cx = sel3(C, r.kx) - r.Sx * Ckz;
    float Cy = sel3(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz;
    float Bz = r.Sz * Bkz;
    float Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    *tout = t;
    return true;
}

// ----------------------------------------------------------- path flags
constexpr uint32_t FLAG_DEPTH_MASK = 0xFFFFu;
constexpr uint32_t FLAG_FIRST_HIT = 1u << 16;
constexpr uint32_t FLAG_PREV_DELTA = 1u << 17;
constexpr uint32_t FLAG_NEE_DONE = 1u << 18;
// host-loop iterations the path took part in (renderer.cpp:213 cap), 5 bits
constexpr uint32_t FLAG_ITER_SHIFT = 19;
constexpr uint32_t FLAG_ITER_MASK = 31u << FLAG_ITER_SHIFT;
// request slot of the path (index into the per-slot parameter table)
constexpr uint32_t FLAG_SLOT_SHIFT = 24;
constexpr uint32_t kMaxSlots = 256;

// Throughput (xyz) and previous-bounce pdf (w) of pool entry p with flags z.
// A path not yet shaded (FLAG_FIRST_HIT) has throughput 1 and pdf 1, which
// the camera-ray pass leaves unstored; its lanes issue no load.
__device__ __forceinline__ float4 path_throughput(const float4* thr, uint32_t p, uint32_t z) {
    return (z & FLAG_FIRST_HIT) ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : thr[p];
}

}  // namespace wrk
