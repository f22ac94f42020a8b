// wr_oracle.hpp — CPU restatement of libWR's path-tracing kernel math.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library includes, links or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load it (through oracle/build/liboracle.so), as the checker.
//
// Every function names the reference file:line it restates. The arithmetic is
// written operation by operation in the same order as the reference source so
// that the FP32 instantiation reproduces the HIP kernels bit for bit when both
// are compiled without FMA contraction. Deliberate deviations from the
// reference (all documented and applied identically in the HIP kernels):
//   * normalize() uses 1/sqrt (IEEE) instead of CUDA's approximate rsqrtf
//     (vector_math.cuh:47-50; the reference builds with --use_fast_math).
//   * sinf/cosf come from wr_sincos() below (argument reduction in FP64,
//     Taylor polynomials in FP32 with fmaf); powf(x,5) is x*x*x*x*x by
//     repeated multiplication. The reference's fast-math transcendentals are
//     not reproducible off NVIDIA hardware ("parity unpinned" at ulp level).
//   * Texture fetches are FP32 software bilinear with wrap addressing instead
//     of CUDA's 9-bit-weight hardware filter (scene.cpp:403-410).
//   * Ray/triangle intersection is the watertight test of Woop, Benthin and
//     Wald (JCGT 2013) in place of OptiX's closed-source intersector, with the
//     tie rule "smallest t, then smallest primitive id".
//   * In the paired draws rnd_dim(seed, dim++) passed as two arguments
//     (lambertian.cuh:45-48,124-127; ggx_material.cuh:59-62,148-151) the first
//     argument takes the lower dimension.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/wrk.h"

namespace orc {

// ---------------------------------------------------------------- vectors
template <typename T> struct V3 { T x, y, z; };
template <typename T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T b) { return {a.x * b, a.y * b, a.z * b}; }
template <typename T> inline V3<T> operator*(T a, V3<T> b) { return {a * b.x, a * b.y, a * b.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// vector_math.cuh:37-39
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vector_math.cuh:41-44
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T sqrtT(T v) { return std::sqrt(v); }
// vector_math.cuh:45
template <typename T> inline T length(V3<T> v) { return sqrtT(dot(v, v)); }
// vector_math.cuh:47-50 (rsqrtf -> 1/sqrt, see header note)
template <typename T> inline V3<T> normalize(V3<T> v) {
    T inv = T(1) / sqrtT(dot(v, v));
    return v * inv;
}
template <typename T> inline T fmaxT(T a, T b) { return std::fmax(a, b); }
template <typename T> inline T fminT(T a, T b) { return std::fmin(a, b); }

// ------------------------------------------------- transcendentals (own)
// sin and cos of x: Cody–Waite style reduction by pi/2 done in FP64 (exact
// enough for |x| < 1e5), then Taylor polynomials on [-pi/4, pi/4] evaluated
// with fused multiply-add (Horner). Truncation error < 2e-9 relative.
inline void wr_sincos(float x, float* s, float* c) {
    double xd = (double)x;
    double kd = std::rint(xd * 0.63661977236758134308);
    double rd = std::fma(-kd, 1.57079632679489661923, xd);
    float r = (float)rd;
    int q = ((int)kd) & 3;
    float r2 = r * r;
    float ps = std::fma(r2, 2.7557319e-6f, -1.9841270e-4f);
    ps = std::fma(r2, ps, 8.3333333e-3f);
    ps = std::fma(r2, ps, -1.6666667e-1f);
    float sn = std::fma(r * r2, ps, r);
    float pc = std::fma(r2, -2.7557319e-7f, 2.4801587e-5f);
    pc = std::fma(r2, pc, -1.3888889e-3f);
    pc = std::fma(r2, pc, 4.1666668e-2f);
    pc = std::fma(r2, pc, -0.5f);
    float cs = std::fma(r2, pc, 1.0f);
    switch (q) {
        case 0: *s = sn;  *c = cs;  break;
        case 1: *s = cs;  *c = -sn; break;
        case 2: *s = -sn; *c = -cs; break;
        default: *s = -cs; *c = sn; break;
    }
}
inline void wr_sincos(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
// powf(x, 5.0f) of ggx.cuh:39 as repeated multiplication.
template <typename T> inline T pow5(T x) { T x2 = x * x; T x4 = x2 * x2; return x4 * x; }

// Constants: kPi (material_common.cuh:7, trace.cu:9) and M_PI (sampling.cuh:7,
// ggx.cuh:6) both round to the same float.
template <typename T> inline T kPiT() { return sizeof(T) == 4 ? T(3.14159265f) : T(3.14159265358979323846); }

// ------------------------------------------------------------------- RNG
// sampling.cuh:17-21
inline uint32_t pcg_hash(uint32_t seed) {
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// sampling.cuh:24-28
inline uint32_t pcg_hash(uint32_t seed, uint32_t dimension) {
    uint32_t state = seed * 747796405u + 2891336453u + dimension * 1013904223u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// sampling.cuh:36-38 (may return exactly 1.0f)
template <typename T> inline T rnd_dim(uint32_t seed, uint32_t dim) {
    if (sizeof(T) == 4) return T(float(pcg_hash(seed, dim)) / 4294967296.0f);
    return T(double(pcg_hash(seed, dim)) / 4294967296.0);
}
// utils/atomic_ops.cuh:16-19 (LCG used for the camera jitter)
template <typename T> inline T randf(uint32_t& seed) {
    seed = seed * 1664525u + 1013904223u;
    if (sizeof(T) == 4) return T(float(seed >> 8) / 16777216.0f);
    return T(double(seed >> 8) / 16777216.0);
}
// trace.cu:58
inline uint32_t camera_seed(uint32_t pixel, uint32_t sample) {
    return (pixel * 1664525u + sample * 1013904223u) ^ 0x9e3779b9u;
}

// ------------------------------------------------------------- scene
struct Texture { uint32_t w = 0, h = 0; std::vector<uint8_t> rgba; };

// Host BVH node (binary, oracle-only acceleration for large scenes).
struct HNode { double lo[3], hi[3]; uint32_t left, right, start, count; };

struct Scene {
    std::vector<float> vertices, normals, uvs;
    std::vector<uint32_t> indices, triMat;
    std::vector<wrk_material> materials;
    std::vector<Texture> textures;
    bool hasNormals = false, hasUVs = false;
    float env[3] = {0, 0, 0};
    std::vector<uint32_t> emissiveTris;
    std::vector<float> emissiveCdf;
    // host BVH
    std::vector<HNode> nodes;
    std::vector<uint32_t> order;
    double extent = 0;
    uint32_t numTris() const { return (uint32_t)(indices.size() / 3); }
};

// scene.cpp:423-491 — emissive list and power CDF, sequential FP32 sum.
void build_light_cdf(Scene& s);
// Oracle-only binary BVH (binned SAH); never changes a result, only speed.
void build_host_bvh(Scene& s);

// ----------------------------------------------------------- intersection
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float o[3];
};
inline RayPre ray_pre(const float o[3], const float d[3]) {
    RayPre p;
    float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.Sx = d[kx] / d[kz];
    p.Sy = d[ky] / d[kz];
    p.Sz = 1.0f / d[kz];
    p.o[0] = o[0]; p.o[1] = o[1]; p.o[2] = o[2];
    return p;
}
// Returns true and t when the triangle is hit with tmin <= t <= tmax.
inline bool intersect_tri(const RayPre& r, const float* v0, const float* v1, const float* v2,
                          float tmin, float tmax, float* tout) {
    float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    float Ax = A[r.kx] - r.Sx * A[r.kz];
    float Ay = A[r.ky] - r.Sy * A[r.kz];
    float Bx = B[r.kx] - r.Sx * B[r.kz];
    float By = B[r.ky] - r.Sy * B[r.kz];
    float Cx = C[r.kx] - r.Sx * C[r.kz];
    float Cy = C[r.ky] - r.Sy * C[r.kz];
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
    float Az = r.Sz * A[r.kz];
    float Bz = r.Sz * B[r.kz];
    float Cz = r.Sz * C[r.kz];
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    *tout = t;
    return true;
}

struct Hit { float t; uint32_t prim; };
// Closest hit by brute force over all triangles. prim = 0xFFFFFFFF on miss.
Hit closest_brute(const Scene& s, const float o[3], const float d[3], float tmin, float tmax);
bool any_brute(const Scene& s, const float o[3], const float d[3], float tmin, float tmax);
// Same answers through the host BVH.
Hit closest_bvh(const Scene& s, const float o[3], const float d[3], float tmin, float tmax);
bool any_bvh(const Scene& s, const float o[3], const float d[3], float tmin, float tmax);

// ------------------------------------------------------------ path state
// Field-for-field the reference RayState (gpu_types.h:48-76).
template <typename T> struct Ray {
    V3<T> origin, direction, throughput, radiance;
    uint32_t pixelIndex, depth, seed, stage;
    T tMin, tMax;
    bool isFirstHit;
    T prevPdf;
    bool prevWasDelta;
    V3<T> shadowContribution, savedDirection;
    uint32_t neeDone, rngDimension;
};
// gpu_types.h:79-86
template <typename T> struct HitInfo {
    V3<T> position, normal, geometricNormal;
    T u, v;
    uint32_t materialId, primIndex;
};
enum Stage : uint32_t { Trace = 0, Shade = 1, Shadow = 2, Terminated = 3 };

struct Camera {
    float position[3], forward[3], right[3], up[3];
    float tanHalfFovY, aspect, focalDistance, lensRadius;
};
struct Params {
    uint32_t width, height, useNEE, maxBounces;
    float rrStartDepth, fireflyClamp;
    uint32_t iterationCap;  // 8 reproduces renderer.cpp:213; 0 = unlimited
};

// Per-path outcome record for parity tests.
struct PathOut {
    float L[3];           // value added to the accumulation buffer (0 if none)
    uint32_t iterations;  // host-loop iterations the path took part in
    uint32_t flushed;     // 1 if the path reached terminateRay / miss flush
    uint32_t depth;       // final depth
    uint32_t rngDim;      // final rngDimension
};

template <typename T>
void trace_path(const Scene& s, const Camera& cam, const Params& p, uint32_t pixel,
                uint32_t sample, bool useBvh, PathOut* out, double* outL64 = nullptr);

// Denoiser guide values of one (pixel, sample) (trace.cu:267-286): albedo
// (texture or constant) and ray-facing geometric normal at the camera ray's
// first hit; zero for a miss or an out-of-range material id.
void first_hit_guides(const Scene& s, const Camera& cam, const Params& p, uint32_t pixel, uint32_t sample,
                      bool useBvh, float albedo[3], float normal[3]);

// ------------------------------------------------ single-step test hooks
// These run one piece of the FP32 restatement on explicit inputs so that the
// tests can pin each piece against closed forms and invariants.

// One shade() visit (shade.cu:24-85) on an explicit ray state and hit record.
// Field meanings follow RayState / HitInfo (gpu_types.h:48-86); `flushed`
// and `flushL` report a terminateRay() flush (material_common.cuh:33-37).
struct EventIO {
    float origin[3], direction[3], throughput[3], radiance[3];
    float tMin, tMax, prevPdf;
    uint32_t seed, rngDimension, depth, stage, neeDone, prevWasDelta;
    float shadowContribution[3], savedDirection[3];
    float position[3], normal[3], geometricNormal[3], u, v;
    uint32_t materialId;
    uint32_t flushed;
    float flushL[3];
};
void shade_event(const Scene& s, const Params& p, EventIO* e);
// Closest-hit attribute reconstruction (trace.cu:167-264) for the ray (o, d)
// at distance t on triangle prim: out = position[3], shading normal[3],
// geometric normal[3] (both facing the ray), u, v.
void hit_attributes(const Scene& s, const float o[3], const float d[3], float t, uint32_t prim, float out[11]);
// Camera ray of (pixel, sample) (trace.cu:53-125).
void camera_ray(const Camera& cam, const Params& p, uint32_t pixel, uint32_t sample, float origin[3],
                float dir[3]);
// Albedo texture fetch of texture texId (1-based) at (u, v) (materials.cuh:9-16).
void sample_texture(const Scene& s, uint32_t texId, float u, float v, float out[3]);
// getEffectiveNormal (normal_map.cuh:36-60) of material mat for shading
// normal n at (u, v).
void effective_normal(const Scene& s, uint32_t mat, const float n[3], float u, float v, float out[3]);

}  // namespace orc
