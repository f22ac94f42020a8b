// oracle_capi.cpp — extern "C" surface of the oracle for the Python tests and
// the CPU baseline (test infrastructure only; see wr_oracle.hpp).
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "wr_oracle.hpp"

using namespace orc;

namespace {
struct OScene {
    Scene s;
    bool finalized = false;
};

// Camera basis as in libWR/include/wr/types.h:37-50 (glm::normalize is
// v * (1/sqrt(dot(v,v))), glm::cross has the same products) and
// renderer.cpp:144-156.
Camera make_camera(const float* pos, const float* target, const float* up, float fovY, float aspect,
                   float focal, float lens) {
    auto norm = [](float v[3]) {
        float d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        float inv = 1.0f / std::sqrt(d);
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    };
    auto crs = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    Camera c;
    float f[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    norm(f);
    float r[3];
    crs(f, up, r);
    norm(r);
    float u[3];
    crs(r, f, u);
    norm(u);
    for (int i = 0; i < 3; ++i) {
        c.position[i] = pos[i]; c.forward[i] = f[i]; c.right[i] = r[i]; c.up[i] = u[i];
    }
    c.tanHalfFovY = std::tan(fovY * 0.5f);
    c.aspect = aspect;
    c.focalDistance = focal;
    c.lensRadius = lens;
    return c;
}
}  // namespace

extern "C" {

// Camera record as seen by ctypes: position[3], target[3], up[3], fovY,
// aspect, focalDistance, lensRadius.
struct orc_camera_in { float position[3], target[3], up[3], fovY, aspect, focalDistance, lensRadius; };
struct orc_params_in { uint32_t width, height, useNEE, maxBounces; float rrStartDepth, fireflyClamp; uint32_t iterationCap; };

void* orc_scene_new(void) { return new OScene(); }
void orc_scene_free(void* h) { delete static_cast<OScene*>(h); }

uint32_t orc_scene_add_material(void* h, const wrk_material* m) {
    OScene* o = static_cast<OScene*>(h);
    o->s.materials.push_back(*m);
    return (uint32_t)o->s.materials.size() - 1;
}

// scene.cpp:170-175; returns -1 for an invalid material id ("Invalid material ID").
int orc_scene_set_normal_map(void* h, uint32_t mat, uint32_t tex) {
    OScene* o = static_cast<OScene*>(h);
    if (mat >= o->s.materials.size()) return -1;
    o->s.materials[mat].normalTextureId = tex;
    return 0;
}

// RGBA8 row-major texture; returns the 1-based id (scene.cpp:246-264).
uint32_t orc_scene_add_texture(void* h, const uint8_t* rgba, uint32_t w, uint32_t ht) {
    OScene* o = static_cast<OScene*>(h);
    Texture t;
    t.w = w; t.h = ht;
    t.rgba.assign(rgba, rgba + (size_t)w * ht * 4);
    o->s.textures.push_back(std::move(t));
    return (uint32_t)o->s.textures.size();
}

// scene.cpp:190-244 — append a mesh; normals/uvs may be null.
void orc_scene_add_mesh(void* h, const float* verts, uint32_t numFloats, const uint32_t* inds,
                        uint32_t numInds, const float* normals, uint32_t numNormalFloats,
                        const float* uvs, uint32_t numUvFloats, uint32_t materialId) {
    Scene& s = static_cast<OScene*>(h)->s;
    uint32_t vOff = (uint32_t)(s.vertices.size() / 3);
    s.vertices.insert(s.vertices.end(), verts, verts + numFloats);
    size_t nv = numFloats / 3;
    if (normals && numNormalFloats >= nv * 3) {
        s.hasNormals = true;
        s.normals.insert(s.normals.end(), normals, normals + numNormalFloats);
    } else {
        s.normals.insert(s.normals.end(), nv * 3, 0.0f);
    }
    if (uvs && numUvFloats >= nv * 2) {
        s.hasUVs = true;
        s.uvs.insert(s.uvs.end(), uvs, uvs + nv * 2);
    } else {
        s.uvs.insert(s.uvs.end(), nv * 2, 0.0f);
    }
    for (uint32_t i = 0; i < numInds; ++i) s.indices.push_back(vOff + inds[i]);
    for (uint32_t i = 0; i < numInds / 3; ++i) s.triMat.push_back(materialId);
}

void orc_scene_set_env(void* h, float r, float g, float b) {
    Scene& s = static_cast<OScene*>(h)->s;
    s.env[0] = r; s.env[1] = g; s.env[2] = b;
}

// scene.cpp:350-491 (CDF) plus the oracle's own host BVH.
int orc_scene_finalize(void* h, int buildBvh) {
    OScene* o = static_cast<OScene*>(h);
    if (o->s.vertices.empty() || o->s.indices.empty()) return -1;
    build_light_cdf(o->s);
    if (buildBvh) build_host_bvh(o->s);
    o->finalized = true;
    return 0;
}

uint32_t orc_scene_num_tris(void* h) { return static_cast<OScene*>(h)->s.numTris(); }
uint32_t orc_scene_num_emissive(void* h) { return (uint32_t)static_cast<OScene*>(h)->s.emissiveTris.size(); }
void orc_scene_emissive(void* h, uint32_t* tris, float* cdf) {
    Scene& s = static_cast<OScene*>(h)->s;
    for (size_t i = 0; i < s.emissiveTris.size(); ++i) { tris[i] = s.emissiveTris[i]; cdf[i] = s.emissiveCdf[i]; }
}

// --- RNG known-answer helpers
uint32_t orc_pcg_hash(uint32_t seed) { return pcg_hash(seed); }
uint32_t orc_pcg_hash2(uint32_t seed, uint32_t dim) { return pcg_hash(seed, dim); }
float orc_rnd_dim(uint32_t seed, uint32_t dim) { return rnd_dim<float>(seed, dim); }
uint32_t orc_camera_seed(uint32_t pixel, uint32_t sample) { return camera_seed(pixel, sample); }
void orc_sincos(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) wr_sincos(x[i], &s[i], &c[i]);
}

// --- intersection: rays are float4 (o.xyz, tmin) and (d.xyz, tmax)
void orc_trace_closest(void* h, const float* org, const float* dir, uint32_t n, int useBvh, int threads,
                       float* tOut, uint32_t* primOut) {
    const Scene& s = static_cast<OScene*>(h)->s;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* o = org + i * 4;
        const float* d = dir + i * 4;
        Hit hh = useBvh ? closest_bvh(s, o, d, o[3], d[3]) : closest_brute(s, o, d, o[3], d[3]);
        tOut[i] = hh.t;
        primOut[i] = hh.prim;
    }
}

void orc_trace_any(void* h, const float* org, const float* dir, uint32_t n, int useBvh, int threads,
                   uint8_t* occluded) {
    const Scene& s = static_cast<OScene*>(h)->s;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* o = org + i * 4;
        const float* d = dir + i * 4;
        occluded[i] = (useBvh ? any_bvh(s, o, d, o[3], d[3]) : any_brute(s, o, d, o[3], d[3])) ? 1 : 0;
    }
}

// --- camera basis as computed by the host API
void orc_camera_basis(const orc_camera_in* ci, float* out16) {
    Camera c = make_camera(ci->position, ci->target, ci->up, ci->fovY, ci->aspect, ci->focalDistance, ci->lensRadius);
    for (int i = 0; i < 3; ++i) {
        out16[i] = c.position[i]; out16[4 + i] = c.forward[i]; out16[8 + i] = c.right[i]; out16[12 + i] = c.up[i];
    }
    out16[3] = c.tanHalfFovY; out16[7] = c.aspect; out16[11] = c.focalDistance; out16[15] = c.lensRadius;
}

// --- per-path outcomes for explicit (pixel, sample) pairs
void orc_render_paths(void* h, const orc_camera_in* ci, const orc_params_in* pi, const uint32_t* pixels,
                      const uint32_t* samples, uint32_t n, int useBvh, int threads, float* outL,
                      uint32_t* outInfo) {
    const Scene& s = static_cast<OScene*>(h)->s;
    Camera cam = make_camera(ci->position, ci->target, ci->up, ci->fovY, ci->aspect, ci->focalDistance, ci->lensRadius);
    Params p{pi->width, pi->height, pi->useNEE, pi->maxBounces, pi->rrStartDepth, pi->fireflyClamp, pi->iterationCap};
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        PathOut po;
        trace_path<float>(s, cam, p, pixels[i], samples[i], useBvh != 0, &po);
        outL[i * 3 + 0] = po.L[0]; outL[i * 3 + 1] = po.L[1]; outL[i * 3 + 2] = po.L[2];
        if (outInfo) {
            outInfo[i * 4 + 0] = po.iterations; outInfo[i * 4 + 1] = po.flushed;
            outInfo[i * 4 + 2] = po.depth; outInfo[i * 4 + 3] = po.rngDim;
        }
    }
}

// --- whole image: accum[p] = sum over samples [s0, s0+spp) in sample order
// (renderer.cpp:190 serial samples). fp64 != 0 runs the FP64 instantiation
// and accumulates in double (accum64 must then be non-null).
// Guide layers averaged over samples 0..spp-1: per pixel summed in sample
// order, then scaled by 1/spp in FP32 (trace.cu:267-286, renderer.cpp:308-313).
void orc_render_guides(void* h, const orc_camera_in* ci, const orc_params_in* pi, uint32_t spp, int useBvh,
                       int threads, float* albedo, float* normal) {
    const Scene& s = static_cast<OScene*>(h)->s;
    Camera cam = make_camera(ci->position, ci->target, ci->up, ci->fovY, ci->aspect, ci->focalDistance, ci->lensRadius);
    Params p{pi->width, pi->height, pi->useNEE, pi->maxBounces, pi->rrStartDepth, pi->fireflyClamp, pi->iterationCap};
    const int64_t npix = (int64_t)p.width * p.height;
    const float inv = 1.0f / (float)spp;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t px = 0; px < npix; ++px) {
        float a[3] = {0, 0, 0}, n[3] = {0, 0, 0};
        for (uint32_t k = 0; k < spp; ++k) {
            float ga[3], gn[3];
            first_hit_guides(s, cam, p, (uint32_t)px, k, useBvh != 0, ga, gn);
            for (int c = 0; c < 3; ++c) { a[c] = a[c] + ga[c]; n[c] = n[c] + gn[c]; }
        }
        for (int c = 0; c < 3; ++c) {
            albedo[px * 3 + c] = a[c] * inv;
            normal[px * 3 + c] = n[c] * inv;
        }
    }
}

void orc_render_image(void* h, const orc_camera_in* ci, const orc_params_in* pi, uint32_t s0, uint32_t spp,
                      int useBvh, int threads, int fp64, float* accum, double* accum64) {
    const Scene& s = static_cast<OScene*>(h)->s;
    Camera cam = make_camera(ci->position, ci->target, ci->up, ci->fovY, ci->aspect, ci->focalDistance, ci->lensRadius);
    Params p{pi->width, pi->height, pi->useNEE, pi->maxBounces, pi->rrStartDepth, pi->fireflyClamp, pi->iterationCap};
    const int64_t npix = (int64_t)p.width * p.height;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t px = 0; px < npix; ++px) {
        if (fp64) {
            double a[3] = {0, 0, 0};
            for (uint32_t k = 0; k < spp; ++k) {
                PathOut po; double L[3];
                trace_path<double>(s, cam, p, (uint32_t)px, s0 + k, useBvh != 0, &po, L);
                a[0] += L[0]; a[1] += L[1]; a[2] += L[2];
            }
            accum64[px * 3 + 0] = a[0]; accum64[px * 3 + 1] = a[1]; accum64[px * 3 + 2] = a[2];
        } else {
            float a[3] = {accum[px * 3 + 0], accum[px * 3 + 1], accum[px * 3 + 2]};
            for (uint32_t k = 0; k < spp; ++k) {
                PathOut po;
                trace_path<float>(s, cam, p, (uint32_t)px, s0 + k, useBvh != 0, &po);
                a[0] += po.L[0]; a[1] += po.L[1]; a[2] += po.L[2];
            }
            accum[px * 3 + 0] = a[0]; accum[px * 3 + 1] = a[1]; accum[px * 3 + 2] = a[2];
        }
    }
}

// --- stable compaction (compact.cu:6-17 keeps stage != Terminated; the
// build's queues are ordered, i.e. std::copy_if)
uint32_t orc_compact(const uint8_t* keep, const uint32_t* in, uint32_t n, uint32_t* out) {
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) if (keep[i]) out[c++] = in ? in[i] : i;
    return c;
}

// --- ACES + gamma + uint8 truncation (libWR/test/helpers/image.cpp:14-49),
// gamma evaluated as (float)pow((double)x, (double)(1.0f/2.2f)).
void orc_aces(const float* img, uint32_t npix, uint8_t* out) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    const float exposure = 2.2f;
    const float gamma = 1.0f / 2.2f;
    for (uint32_t i = 0; i < npix * 3; ++i) {
        float x = img[i] * exposure;
        float y = (x * (a * x + b)) / (x * (c * x + d) + e);
        y = y < 0.0f ? 0.0f : (y > 1.0f ? 1.0f : y);
        float g = (float)std::pow((double)y, (double)gamma);
        out[i] = (uint8_t)(g * 255.0f);
    }
}

// --- single-step hooks (see wr_oracle.hpp)
void orc_shade_event(void* h, const orc_params_in* pi, EventIO* e) {
    Params p{pi->width, pi->height, pi->useNEE, pi->maxBounces, pi->rrStartDepth, pi->fireflyClamp, pi->iterationCap};
    shade_event(static_cast<OScene*>(h)->s, p, e);
}
void orc_hit_attributes(void* h, const float* o, const float* d, float t, uint32_t prim, float* out11) {
    hit_attributes(static_cast<OScene*>(h)->s, o, d, t, prim, out11);
}
void orc_camera_ray(const orc_camera_in* ci, const orc_params_in* pi, uint32_t pixel, uint32_t sample,
                    float* origin, float* dir) {
    Camera cam = make_camera(ci->position, ci->target, ci->up, ci->fovY, ci->aspect, ci->focalDistance, ci->lensRadius);
    Params p{pi->width, pi->height, pi->useNEE, pi->maxBounces, pi->rrStartDepth, pi->fireflyClamp, pi->iterationCap};
    camera_ray(cam, p, pixel, sample, origin, dir);
}
void orc_sample_texture(void* h, uint32_t texId, float u, float v, float* out) {
    sample_texture(static_cast<OScene*>(h)->s, texId, u, v, out);
}
int orc_effective_normal(void* h, uint32_t mat, const float* n, float u, float v, float* out) {
    const Scene& s = static_cast<OScene*>(h)->s;
    if (mat >= s.materials.size()) return -1;
    effective_normal(s, mat, n, u, v, out);
    return 0;
}

double orc_wall_seconds(void) {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
