// wavefront.hip — camera generation, trace-result classification, the five
// material shaders, accumulation and tonemapping for gfx950.
//
// Restates libWR/kernels/trace.cu (camera, hit attributes, emissive MIS,
// miss and shadow resolution), shade.cu and materials/*.cuh over a
// structure-of-arrays path pool with material-sorted queues. Built with
// -ffp-contract=off: the arithmetic order follows the reference source.
#include <hip/hip_runtime.h>

#include "traverse.h"
#include "wrk_common.h"
#include "wrk_device.h"

namespace wrk {

// Shade categories 0..4 are the material kinds; CAT_DONE ends the path.
// After shading a path continues as one of three queue kinds: trace the
// continuation ray, trace it together with a shadow ray, or trace a shadow
// ray alone (its continuation either ended or, under an iteration cap,
// waits for the shadow result as in libWR's stage machine).
enum : uint32_t { CAT_DONE = kCatDone, NEXT_TRACE = 0, NEXT_TRACE_SHADOW = 1, NEXT_SHADOW = 2, NEXT_DONE = 3 };

struct Pool {
    float4* org;
    float4* dir;
    float4* thr;
    float4* rad;
    uint4* st;
    float4* contrib;
    float4* savedOrg;
    float4* savedDir;
    uint2* hit;
    uint8_t* occl;
    float4* result;
    float4* guides;  // optional: [2p] albedo, [2p + 1] normal of the first hit
};
static Pool pool_of(const wrk_paths* p) {
    Pool q;
    q.org = (float4*)p->org; q.dir = (float4*)p->dir; q.thr = (float4*)p->thr; q.rad = (float4*)p->rad;
    q.st = (uint4*)p->state; q.contrib = (float4*)p->contrib; q.savedOrg = (float4*)p->savedOrg;
    q.savedDir = (float4*)p->savedDir; q.hit = (uint2*)p->hit; q.occl = (uint8_t*)p->occluded;
    q.result = (float4*)p->result;
    q.guides = (float4*)p->guides;
    return q;
}

struct SceneDev {
    const float* vertices;
    const float* normals;
    const float* uvs;
    const uint32_t* indices;
    const uint32_t* triMat;
    const wrk_material* materials;
    const wrk_texture* textures;
    const uint32_t* emissiveTris;
    const float* emissiveCdf;
    const uint32_t* emissiveGuide;
    uint32_t numTris, numMaterials, numTextures, numEmissive, guideSize;
    float env[3];
};
static SceneDev scene_of(const wrk_scene_view* v) {
    SceneDev s;
    s.vertices = (const float*)v->vertices; s.normals = (const float*)v->normals; s.uvs = (const float*)v->uvs;
    s.indices = (const uint32_t*)v->indices; s.triMat = (const uint32_t*)v->triMaterialIds;
    s.materials = (const wrk_material*)v->materials; s.textures = (const wrk_texture*)v->textures;
    s.emissiveTris = (const uint32_t*)v->emissiveTris; s.emissiveCdf = (const float*)v->emissiveCdf;
    s.numTris = v->numTris; s.numMaterials = v->numMaterials; s.numTextures = v->numTextures;
    s.numEmissive = v->numEmissive;
    s.emissiveGuide = (const uint32_t*)v->emissiveGuide;
    s.guideSize = v->emissiveGuideSize;
    s.env[0] = v->envRadiance[0]; s.env[1] = v->envRadiance[1]; s.env[2] = v->envRadiance[2];
    return s;
}

// ------------------------------------------------------------- generate
// trace.cu:53-125: camera ray of global path g = (pixel, sample) into pool
// entry i, path state reset. The throughput (1, 1, 1, prevPdf 1) and the
// result are not stored: a path whose flags still hold FLAG_FIRST_HIT has
// that throughput by definition (its first shade or miss takes it from the
// flag), and every path writes its result where it ends, an unflushed end
// (iteration cap, out-of-range material id) included. Radiance is stored as
// zero because shaders that never add to it leave it in the pool.
__device__ __forceinline__ void generate_path(const wrk_camera& cam, const wrk_render_params& prm, const Pool& P,
                                              uint64_t g, uint32_t i, uint32_t slot) {
    const uint64_t npix = (uint64_t)prm.width * prm.height;
    const uint32_t pixel = (uint32_t)(g % npix);
    const uint32_t sample = (uint32_t)(g / npix);
    const uint32_t px = pixel % prm.width;
    const uint32_t py = pixel / prm.width;
    uint32_t seed = (pixel * 1664525u + sample * 1013904223u) ^ 0x9e3779b9u;
    uint32_t js = seed;
    float r1 = randf(js);
    float r2 = randf(js);
    float ndcX = (2.0f * ((float)px + r1) / (float)prm.width - 1.0f) * cam.aspect;
    float ndcY = 1.0f - 2.0f * ((float)py + r2) / (float)prm.height;
    f3 Pp = mk(cam.position[0], cam.position[1], cam.position[2]);
    f3 F = mk(cam.forward[0], cam.forward[1], cam.forward[2]);
    f3 R = mk(cam.right[0], cam.right[1], cam.right[2]);
    f3 U = mk(cam.up[0], cam.up[1], cam.up[2]);
    float th = cam.tanHalfFovY;
    f3 d = F + R * ndcX * th + U * ndcY * th;
    d = normalize(d);
    f3 o = Pp;
    if (cam.lensRadius > 0.0f && cam.focalDistance > 0.0f) {
        float r3 = randf(js);
        float r4 = randf(js);
        float theta = (2.0f * kPi) * r3;
        float radius = sqrtf(r4) * cam.lensRadius;
        float sn, cs;
        wr_sincos(theta, &sn, &cs);
        f3 lo = mk(R.x * radius * cs + U.x * radius * sn, R.y * radius * cs + U.y * radius * sn,
                   R.z * radius * cs + U.z * radius * sn);
        float fd = cam.focalDistance;
        f3 fp = mk(Pp.x + d.x * fd, Pp.y + d.y * fd, Pp.z + d.z * fd);
        o = mk(Pp.x + lo.x, Pp.y + lo.y, Pp.z + lo.z);
        d = normalize(mk(fp.x - o.x, fp.y - o.y, fp.z - o.z));
    }
    P.org[i] = f4(o, 0.0001f);
    P.dir[i] = f4(d, 1e20f);
    P.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.st[i] = make_uint4(seed, 0u, FLAG_FIRST_HIT | FLAG_PREV_DELTA | (slot << FLAG_SLOT_SHIFT), pixel);
    if (P.guides) {
        P.guides[2 * (size_t)i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.guides[2 * (size_t)i + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_generate(wrk_camera cam, wrk_render_params prm, Pool P, uint64_t first,
                                                  uint32_t n, uint32_t poolBase, uint32_t slot, uint32_t* queue,
                                                  uint8_t* cat, uint32_t catValue) {
    uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = poolBase + k;
    generate_path(cam, prm, P, first + k, i, slot);
    queue[k] = i;
    if (cat) cat[k] = (uint8_t)catValue;
}

// ------------------------------------------------------------- classify
// Queue of one scene: [trace nT | trace + shadow nTS | shadow nS].
// Shadow results first (trace.cu:27-51): add the NEE contribution when
// unoccluded. Then closest-hit results: miss -> environment + flush
// (trace.cu:143-148); hit -> category = material kind of the primitive
// (unknown kinds join the emissive queue, whose shader flushes them like
// the default case of shade.cu:79-81; an out-of-range material id ends the
// path without a flush, shade.cu:35-38). A shadow-only entry either resumes
// its deferred continuation in the Lambertian shader (neeDone set, iteration
// cap mode) or ends: its continuation terminated while the shadow ray was in
// flight, so the radiance is flushed now that the contribution is in.
__global__ __launch_bounds__(256) void k_classify(SceneDev S, const wrk_render_params* __restrict__ prmTable, Pool P,
                                                  const uint32_t* queue, uint32_t nT, uint32_t nTS, uint32_t nS,
                                                  uint8_t* cat, uint32_t presetCats) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nT + nTS + nS) return;
    // the closest-hit launch already wrote the category of tagged hits
    const uint32_t pre = (presetCats && i < nT + nTS) ? cat[i] : 0xFFu;
    if (pre != 0xFFu && i < nT) return;
    uint32_t p = queue[i];
    uint32_t c;
    if (i >= nT) {
        if (!P.occl[p]) {
            float4 L = P.rad[p];
            float4 C = P.contrib[p];
            P.rad[p] = make_float4(L.x + C.x, L.y + C.y, L.z + C.z, L.w);
        }
    }
    if (pre != 0xFFu) {
        return;
    } else if (i < nT + nTS) {
        uint2 h = P.hit[p];
        if (h.y == 0xFFFFFFFFu) {
            const uint32_t z = P.st[p].z;
            float4 T = path_throughput(P.thr, p, z), L = P.rad[p];
            const uint32_t slot = z >> FLAG_SLOT_SHIFT;
            f3 Lr = mk(L.x + T.x * S.env[0], L.y + T.y * S.env[1], L.z + T.z * S.env[2]);
            f3 Lc = clamp_fireflies(Lr, prmTable[slot].fireflyClamp);
            P.rad[p] = f4(Lr, 0.0f);
            P.result[p] = f4(Lc, 0.0f);
            c = CAT_DONE;
        } else {
            c = shade_category(S.triMat[h.y], S.numMaterials, S.materials);
            if (c == CAT_DONE) P.result[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // ends unflushed
        }
    } else if (P.st[p].z & FLAG_NEE_DONE) {
        c = WRK_MAT_LAMBERTIAN;
    } else {
        const uint32_t slot = P.st[p].z >> FLAG_SLOT_SHIFT;
        float4 L = P.rad[p];
        P.result[p] = f4(clamp_fireflies(mk(L.x, L.y, L.z), prmTable[slot].fireflyClamp), 0.0f);
        c = CAT_DONE;
    }
    cat[i] = (uint8_t)c;
}

// ----------------------------------------------------------------- shade
struct Hit {
    f3 position, normal, geometricNormal;
    float u, v;
    uint32_t materialId;
};

struct PathState {
    f3 o, d;
    float tMin, tMax;
    f3 T, L;
    float prevPdf;
    uint32_t seed, dim, depth;
    bool prevDelta, firstHit, neeDone;
    uint32_t pixel;
    bool flushed, done, toShadow;
    f3 flushL;
    f3 contrib, shOrg, shDir;
    float shTMax;
};

__device__ __forceinline__ f3 vert(const SceneDev& S, uint32_t i) {
    return mk(S.vertices[i * 3 + 0], S.vertices[i * 3 + 1], S.vertices[i * 3 + 2]);
}

__device__ __forceinline__ f3 material_albedo(const SceneDev& S, const wrk_material& m, float u, float v) {
    if (m.albedoTextureId != 0 && m.albedoTextureId <= S.numTextures && S.textures) {
        return sample_tex(S.textures[m.albedoTextureId - 1], u, v);
    }
    return mk(m.albedo[0], m.albedo[1], m.albedo[2]);
}

__device__ __forceinline__ void terminate_ray(PathState& r, float maxL) {
    r.flushL = clamp_fireflies(r.L, maxL);
    r.flushed = true;
    r.done = true;
}

// sampling.cuh:162-193 + material_common.cuh:39-52
__device__ __forceinline__ bool apply_rr(PathState& r, const wrk_render_params& prm, bool isDelta) {
    float xi = rnd_dim(r.seed, r.dim++);
    float rrStart = prm.rrStartDepth;
    if ((float)r.depth < rrStart) return true;
    float lum = 0.2126f * r.T.x + 0.7152f * r.T.y + 0.0722f * r.T.z;
    float depthFactor = 1.0f / (1.0f + 0.1f * ((float)r.depth - rrStart));
    float minProb = isDelta ? 0.5f : 0.05f;
    float q = fmaxf(minProb, fminf(0.95f, lum * depthFactor));
    if (xi < q) {
        r.T = mk(r.T.x / q, r.T.y / q, r.T.z / q);
        return true;
    }
    terminate_ray(r, prm.fireflyClamp);
    return false;
}

// material_common.cuh:55-71
__device__ __forceinline__ void setup_next(PathState& r, const Hit& h, f3 nd, const wrk_render_params& prm) {
    r.d = nd;
    r.o = offset_origin(h.position, h.normal, nd);
    r.tMin = 0.0f;
    r.tMax = 1e20f;
    r.depth++;
    if (r.depth >= prm.maxBounces) terminate_ray(r, prm.fireflyClamp);
}

// sampling.cuh:116-135
__device__ __forceinline__ uint32_t sample_emissive(const SceneDev& S, float u) {
    uint32_t left = 0, right = S.numEmissive;
    if (S.emissiveGuide) {
        // the answer lies between the guides of bucket floor(u * G); u * G
        // is exact for a power-of-two G, so the result equals the search
        // over the whole CDF (u = 1 falls in the last bucket)
        const uint32_t k = min((uint32_t)(u * (float)S.guideSize), S.guideSize - 1u);
        left = S.emissiveGuide[k];
        right = S.emissiveGuide[k + 1u];
    }
    while (left < right) {
        uint32_t mid = (left + right) / 2;
        if (S.emissiveCdf[mid] < u) left = mid + 1; else right = mid;
    }
    return min(left, S.numEmissive - 1);
}

struct LightSample { f3 toLight; float distSq, dist, cosTheta, lightPdf; f3 Le; };

// lambertian.cuh:17-70 / ggx_material.cuh:33-109
__device__ __forceinline__ bool sample_light(const SceneDev& S, PathState& r, const Hit& h, LightSample& ls) {
    float lightU = rnd_dim(r.seed, r.dim++);
    uint32_t li = sample_emissive(S, lightU);
    uint32_t tri = S.emissiveTris[li];
    uint32_t i0 = S.indices[tri * 3 + 0], i1 = S.indices[tri * 3 + 1], i2 = S.indices[tri * 3 + 2];
    f3 v0 = vert(S, i0), v1 = vert(S, i1), v2 = vert(S, i2);
    float u1 = rnd_dim(r.seed, r.dim);
    float u2 = rnd_dim(r.seed, r.dim + 1);
    r.dim += 2;
    // sampling.cuh:85-110
    float su = sqrtf(u1);
    float b0 = 1.0f - su;
    float b1 = su * (1.0f - u2);
    float b2 = su * u2;
    f3 lp = mk(v0.x * b0 + v1.x * b1 + v2.x * b2, v0.y * b0 + v1.y * b1 + v2.y * b2,
               v0.z * b0 + v1.z * b1 + v2.z * b2);
    f3 e1 = mk(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
    f3 e2 = mk(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
    f3 cp = cross(e1, e2);
    float area = 0.5f * length(cp);
    f3 ln = normalize(cp);
    f3 tl = mk(lp.x - h.position.x, lp.y - h.position.y, lp.z - h.position.z);
    ls.distSq = dot(tl, tl);
    ls.dist = sqrtf(ls.distSq);
    ls.toLight = mk(tl.x / ls.dist, tl.y / ls.dist, tl.z / ls.dist);
    ls.cosTheta = dot(h.normal, ls.toLight);
    float cosLight = -dot(ln, ls.toLight);
    if (!(ls.cosTheta > 0.0f && cosLight > 0.0f)) return false;
    const wrk_material& lm = S.materials[S.triMat[tri]];
    float pdfArea = 1.0f / ((float)S.numEmissive * area);
    ls.lightPdf = pdfArea * ls.distSq / fmaxf(cosLight, 1e-8f);
    ls.Le = mk(lm.emission[0], lm.emission[1], lm.emission[2]);
    return true;
}

__device__ __forceinline__ bool valid_contrib(f3 c) {
    return !isnan(c.x) && !isnan(c.y) && !isnan(c.z) && !isinf(c.x) && !isinf(c.y) && !isinf(c.z) &&
           c.x >= 0.0f && c.y >= 0.0f && c.z >= 0.0f;
}

// materials/lambertian.cuh:8-147
// The shadow ray goes to shOrg/shDir. Under an iteration cap the shader
// returns here and resumes after the shadow test (libWR's stage machine);
// otherwise it samples the continuation at once and both rays are traced in
// the next iteration (same random dimensions, same arithmetic, and the
// contribution still enters L before anything the continuation adds).
__device__ void shade_lambertian(const SceneDev& S, const wrk_render_params& prm, PathState& r, const Hit& h,
                                 const wrk_material& mat, bool deferContinuation) {
    if (prm.useNEE && S.numEmissive > 0 && !r.neeDone) {
        LightSample ls;
        if (sample_light(S, r, h, ls)) {
            float bsdfPdf = fmaxf(0.0f, ls.cosTheta / kPi);
            float w = power_heuristic(ls.lightPdf, bsdfPdf);
            f3 albedo = material_albedo(S, mat, h.u, h.v);
            f3 bsdf = mk(albedo.x / kPi * ls.cosTheta, albedo.y / kPi * ls.cosTheta, albedo.z / kPi * ls.cosTheta);
            f3 c = mk(r.T.x * bsdf.x * ls.Le.x * w / ls.lightPdf, r.T.y * bsdf.y * ls.Le.y * w / ls.lightPdf,
                      r.T.z * bsdf.z * ls.Le.z * w / ls.lightPdf);
            if (valid_contrib(c)) {
                r.contrib = c;
                r.shDir = ls.toLight;
                r.shOrg = offset_origin(h.position, h.normal, ls.toLight);
                r.shTMax = ls.dist - 1e-5f;
                r.toShadow = true;
                if (deferContinuation) {
                    r.neeDone = true;
                    return;
                }
            }
        }
    }
    r.neeDone = false;
    f3 tg, bt;
    make_frame(h.normal, tg, bt);
    float u1 = rnd_dim(r.seed, r.dim);
    float u2 = rnd_dim(r.seed, r.dim + 1);
    r.dim += 2;
    float pdf;
    f3 local = sample_cosine(u1, u2, pdf);
    f3 wd = to_world(local, h.normal, tg, bt);
    f3 albedo = material_albedo(S, mat, h.u, h.v);
    r.T = r.T * albedo;
    r.prevPdf = pdf;
    r.prevDelta = false;
    if (!apply_rr(r, prm, false)) return;
    setup_next(r, h, wd, prm);
}

// materials/glass.cuh:10-140 == materials/ggx_transmission.cuh:12-107
__device__ void shade_glass(const SceneDev& S, const wrk_render_params& prm, PathState& r, const Hit& h,
                            const wrk_material& mat) {
    f3 wo = -r.d;
    float cosI = dot(wo, h.normal);
    float etaI = 1.0f, etaT = mat.ior;
    f3 n = h.normal;
    if (cosI < 0.0f) { cosI = -cosI; n = -n; etaI = mat.ior; etaT = 1.0f; }
    float F = fresnel_dielectric(cosI, etaI, etaT);
    float rr = rnd_dim(r.seed, r.dim++);
    if (rr < F) {
        float dn = 2.0f * dot(r.d, n);
        r.d = mk(r.d.x - n.x * dn, r.d.y - n.y * dn, r.d.z - n.z * dn);
        r.T = r.T * material_albedo(S, mat, h.u, h.v);
    } else {
        float eta = etaI / etaT;
        float sinT2 = eta * eta * (1.0f - cosI * cosI);
        if (sinT2 >= 1.0f) {
            float dn = 2.0f * dot(r.d, n);
            r.d = mk(r.d.x - n.x * dn, r.d.y - n.y * dn, r.d.z - n.z * dn);
            r.T = r.T * material_albedo(S, mat, h.u, h.v);
        } else {
            float cosT = sqrtf(fmaxf(0.0f, 1.0f - sinT2));
            float k = eta * cosI - cosT;
            f3 rf = mk(eta * r.d.x + n.x * k, eta * r.d.y + n.y * k, eta * r.d.z + n.z * k);
            r.d = rf;
            r.T = r.T * material_albedo(S, mat, h.u, h.v);
            float corr = asym_correction(rf, h.normal, h.geometricNormal);
            r.T = r.T * corr;
        }
    }
    r.prevDelta = true;
    if (!apply_rr(r, prm, true)) return;
    setup_next(r, h, r.d, prm);
}

// materials/ggx_material.cuh:10-213
__device__ void shade_ggx(const SceneDev& S, const wrk_render_params& prm, PathState& r, const Hit& h,
                          const wrk_material& mat) {
    f3 wo = -r.d;
    f3 albedo = mk(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
    if (prm.useNEE && S.numEmissive > 0) {
        LightSample ls;
        if (sample_light(S, r, h, ls)) {
            f3 brdf = eval_ggx(wo, ls.toLight, h.normal, albedo, mat.roughness, mat.metallic, mat.ior);
            float bsdfPdf = ggx_pdf(wo, ls.toLight, h.normal, mat.roughness);
            float w = power_heuristic(ls.lightPdf, bsdfPdf);
            f3 c = mk(r.T.x * brdf.x * ls.Le.x * ls.cosTheta * w / ls.lightPdf,
                      r.T.y * brdf.y * ls.Le.y * ls.cosTheta * w / ls.lightPdf,
                      r.T.z * brdf.z * ls.Le.z * ls.cosTheta * w / ls.lightPdf);
            if (valid_contrib(c)) r.L = mk(r.L.x + c.x, r.L.y + c.y, r.L.z + c.z);
        }
    }
    f3 tg, bt;
    make_frame(h.normal, tg, bt);
    float u0 = rnd_dim(r.seed, r.dim);
    float u1 = rnd_dim(r.seed, r.dim + 1);
    r.dim += 2;
    f3 wi = mk(0.0f, 0.0f, 0.0f);
    float pdf;
    sample_ggx(wo, h.normal, tg, bt, mat.roughness, u0, u1, wi, pdf);
    float cosTheta = dot(h.normal, wi);
    if (pdf < 1e-5f || cosTheta <= 0.0f) { terminate_ray(r, prm.fireflyClamp); return; }
    f3 brdf = eval_ggx(wo, wi, h.normal, albedo, mat.roughness, mat.metallic, mat.ior);
    if (isnan(brdf.x) || isnan(brdf.y) || isnan(brdf.z) || isinf(brdf.x) || isinf(brdf.y) || isinf(brdf.z)) {
        terminate_ray(r, prm.fireflyClamp);
        return;
    }
    r.T = mk(r.T.x * brdf.x * cosTheta / pdf, r.T.y * brdf.y * cosTheta / pdf, r.T.z * brdf.z * cosTheta / pdf);
    r.prevPdf = pdf;
    r.prevDelta = false;
    if (!apply_rr(r, prm, false)) return;
    setup_next(r, h, wi, prm);
}

// trace.cu:155-264 (hit attributes) reconstructed from (origin, direction,
// t, primitive); trace.cu:289-316 emissive MIS is applied by the caller.
__device__ __forceinline__ Hit reconstruct(const SceneDev& S, f3 o, f3 d, float t, uint32_t prim, f3* e1o,
                                           f3* e2o) {
    Hit hit;
    f3 hitPos = mk(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    uint32_t i0 = S.indices[prim * 3 + 0], i1 = S.indices[prim * 3 + 1], i2 = S.indices[prim * 3 + 2];
    f3 v0 = vert(S, i0), v1 = vert(S, i1), v2 = vert(S, i2);
    f3 e1 = v1 - v0, e2 = v2 - v0;
    *e1o = e1; *e2o = e2;
    f3 gN = normalize(cross(e1, e2));
    f3 v0p = hitPos - v0;
    float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
    float d20 = dot(v0p, e1), d21 = dot(v0p, e2);
    float denom = d00 * d11 - d01 * d01;
    float vb = (d11 * d20 - d01 * d21) / fmaxf(denom, 1e-8f);
    float wb = (d00 * d21 - d01 * d20) / fmaxf(denom, 1e-8f);
    float ub = 1.0f - vb - wb;
    f3 sN = gN;
    if (S.normals) {
        f3 n0 = mk(S.normals[i0 * 3 + 0], S.normals[i0 * 3 + 1], S.normals[i0 * 3 + 2]);
        f3 n1 = mk(S.normals[i1 * 3 + 0], S.normals[i1 * 3 + 1], S.normals[i1 * 3 + 2]);
        f3 n2 = mk(S.normals[i2 * 3 + 0], S.normals[i2 * 3 + 1], S.normals[i2 * 3 + 2]);
        float l0 = length(n0), l1 = length(n1), l2 = length(n2);
        if (l0 > 1e-6f && l1 > 1e-6f && l2 > 1e-6f) {
            sN = normalize(ub * n0 + vb * n1 + wb * n2);
            if (dot(sN, gN) < 0.0f) sN = -sN;
        }
    }
    hit.geometricNormal = dot(gN, d) > 0.0f ? -gN : gN;
    hit.normal = dot(sN, d) > 0.0f ? -sN : sN;
    hit.position = hitPos;
    hit.materialId = S.triMat[prim];
    if (S.uvs) {
        float a0 = S.uvs[i0 * 2 + 0], b0 = S.uvs[i0 * 2 + 1];
        float a1 = S.uvs[i1 * 2 + 0], b1 = S.uvs[i1 * 2 + 1];
        float a2 = S.uvs[i2 * 2 + 0], b2 = S.uvs[i2 * 2 + 1];
        hit.u = ub * a0 + vb * a1 + wb * a2;
        hit.v = ub * b0 + vb * b1 + wb * b2;
    } else {
        hit.u = 0.0f;
        hit.v = 0.0f;
    }
    return hit;
}

// utils/normal_map.cuh:36-60
__device__ __forceinline__ f3 effective_normal(const SceneDev& S, const wrk_material& mat, const Hit& h) {
    if (mat.normalTextureId > 0 && mat.normalTextureId <= S.numTextures) {
        const wrk_texture& tex = S.textures[mat.normalTextureId - 1];
        if (tex.texels == 0) return h.normal;
        f3 tn = sample_tex(tex, h.u, h.v);
        f3 tg, bt;
        make_frame(h.normal, tg, bt);
        f3 n = mk(tn.x * 2.0f - 1.0f, tn.y * 2.0f - 1.0f, tn.z * 2.0f - 1.0f);
        f3 w = mk(tg.x * n.x + bt.x * n.y + h.normal.x * n.z, tg.y * n.x + bt.y * n.y + h.normal.y * n.z,
                  tg.z * n.x + bt.z * n.y + h.normal.z * n.z);
        return normalize(w);
    }
    return h.normal;
}

// One shade visit of pool entry p in material queue seg (shade.cu:24-85 and
// the materials/ headers): reads the path state and hit record, runs the
// material's shader and writes the next state. Returns NEXT_TRACE,
// NEXT_TRACE_SHADOW, NEXT_SHADOW or NEXT_DONE.
__device__ __forceinline__ uint32_t shade_path(const SceneDev& S, const wrk_render_params* __restrict__ prmTable,
                                               const Pool& P, uint32_t p, int seg) {
    PathState r;
    uint4 st = P.st[p];
    const uint32_t slot = st.z >> FLAG_SLOT_SHIFT;
    const wrk_render_params prm = prmTable[slot];
    const uint32_t iter = ((st.z & FLAG_ITER_MASK) >> FLAG_ITER_SHIFT) + 1u;
    r.seed = st.x; r.dim = st.y;
    r.depth = st.z & FLAG_DEPTH_MASK;
    r.firstHit = (st.z & FLAG_FIRST_HIT) != 0;
    r.prevDelta = (st.z & FLAG_PREV_DELTA) != 0;
    r.neeDone = (st.z & FLAG_NEE_DONE) != 0;
    r.pixel = st.w;
    const bool defer = prm.iterationCap != 0;
    // Lambertian, glass and GGX-transmission shaders never add to L (their
    // light contribution goes through the shadow ray), so their L is read
    // only for a flush and is not written back
    const bool lazyL = seg == WRK_MAT_LAMBERTIAN || seg == WRK_MAT_GLASS || seg == WRK_MAT_GGX_TRANSMISSION;
    // a path at its first hit has throughput 1 and radiance 0 (the camera
    // pass stores the radiance as zero and no throughput)
    const bool fresh = r.firstHit;
    float4 o4 = P.org[p];
    float4 d4 = P.dir[p];
    float4 T4 = path_throughput(P.thr, p, st.z);
    const float4 L4 = (lazyL || fresh) ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : P.rad[p];
    r.o = xyz(o4); r.tMin = o4.w; r.d = xyz(d4); r.tMax = d4.w;
    r.T = xyz(T4); r.prevPdf = T4.w; r.L = xyz(L4);
    r.flushed = false; r.done = false; r.toShadow = false;
    uint2 hr = P.hit[p];
    float t = __uint_as_float(hr.x);
    f3 e1, e2;
    Hit h = reconstruct(S, r.o, r.d, t, hr.y, &e1, &e2);
    const wrk_material& mat = S.materials[h.materialId];
    if (r.firstHit && P.guides) {
        // trace.cu:267-286: albedo (texture or constant) and the ray-facing
        // geometric normal of the first hit
        P.guides[2 * (size_t)p] = f4(material_albedo(S, mat, h.u, h.v), 0.0f);
        P.guides[2 * (size_t)p + 1] = f4(h.geometricNormal, 0.0f);
    }
    r.firstHit = false;

    if (seg == WRK_MAT_EMISSIVE) {
        // trace.cu:289-316: emission with MIS, then shadeEmissive / default
        if (mat.type == WRK_MAT_EMISSIVE) {
            float w = 1.0f;
            if (prm.useNEE && r.depth > 0 && !r.prevDelta && S.numEmissive > 0) {
                float area = 0.5f * length(cross(e1, e2));
                float distSq = t * t;
                float cosL = fabsf(dot(h.geometricNormal, r.d));
                float pdfArea = 1.0f / ((float)S.numEmissive * area);
                float lightPdf = pdfArea * distSq / fmaxf(cosL, 1e-8f);
                w = power_heuristic(r.prevPdf, lightPdf);
            }
            f3 E = mk(mat.emission[0], mat.emission[1], mat.emission[2]);
            f3 TE = r.T * E;
            r.L = r.L + TE * w;
        }
        terminate_ray(r, prm.fireflyClamp);
    } else {
        Hit eh = h;
        if (mat.normalTextureId > 0) eh.normal = effective_normal(S, mat, h);
        if (seg == WRK_MAT_LAMBERTIAN) shade_lambertian(S, prm, r, eh, mat, defer);
        else if (seg == WRK_MAT_GGX_REFLECTION) shade_ggx(S, prm, r, eh, mat);
        else shade_glass(S, prm, r, eh, mat);
    }

    if (lazyL && r.flushed) r.flushL = clamp_fireflies(fresh ? mk(0.0f, 0.0f, 0.0f) : xyz(P.rad[p]), prm.fireflyClamp);
    uint32_t next;
    if (r.done && r.toShadow) {
        next = NEXT_SHADOW;  // flushed once the shadow result is in
    } else if (r.done) {
        next = NEXT_DONE;
        if (r.flushed) P.result[p] = f4(r.flushL, 0.0f);
    } else if (defer && iter >= prm.iterationCap) {
        next = NEXT_DONE;  // renderer.cpp:213 cap: the path is dropped unflushed
        P.result[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (r.toShadow) {
        next = defer ? NEXT_SHADOW : NEXT_TRACE_SHADOW;
    } else {
        next = NEXT_TRACE;
    }
    if (r.toShadow && next != NEXT_DONE) {
        P.contrib[p] = f4(r.contrib, 0.0f);
        P.savedOrg[p] = f4(r.shOrg, 1e-5f);
        P.savedDir[p] = f4(r.shDir, r.shTMax);
    }
    if (next != NEXT_DONE) {
        P.org[p] = f4(r.o, r.tMin);
        P.dir[p] = f4(r.d, r.tMax);
        P.thr[p] = f4(r.T, r.prevPdf);
        if (!lazyL) P.rad[p] = f4(r.L, 0.0f);
        uint32_t flags = (r.depth & FLAG_DEPTH_MASK) | (r.prevDelta ? FLAG_PREV_DELTA : 0u) |
                         (r.neeDone ? FLAG_NEE_DONE : 0u) | (min(iter, 31u) << FLAG_ITER_SHIFT) |
                         (slot << FLAG_SLOT_SHIFT);
        P.st[p] = make_uint4(r.seed, r.dim, flags, r.pixel);
    }
    return next;
}

// One launch over a scene's queue part in queue order (no material
// partition). Each block takes 256 consecutive entries, which the queue keeps
// in pool order within each admitted chunk, so the block's pool accesses
// stay within a short pool range; it orders them by category in LDS, so a
// wave mostly runs one material's shader. Each entry's next category goes
// back to its own queue position, and queueOut receives the queue as it
// was (the queue rebuild reads it there). Measured on the material scene
// against material-partitioned segments (a stable partition by material,
// then one shader per block): shade 74.3 -> 72.6 ms and partition 14.5 ->
// 8.9 ms per 1080p 256-spp render; without the LDS ordering shade took 79.0
// ms, with 1024-entry windows (4 per thread) 77.1.
__global__ __launch_bounds__(256) void k_shade(SceneDev S, const wrk_render_params* __restrict__ prmTable,
                                                     Pool P, const uint32_t* __restrict__ queueIn,
                                                     uint32_t* __restrict__ queueOut, uint32_t n, uint32_t catBase,
                                                     uint32_t catFinished, uint8_t* category) {
    __shared__ uint32_t sCount[8];
    __shared__ uint32_t sEntry[256];  // pool entry of each sorted position
    __shared__ uint16_t sPos[256];    // block-local queue position of each sorted position
    __shared__ uint8_t sCat[256];
    const uint32_t t = threadIdx.x;
    const uint32_t j = blockIdx.x * 256u + t;
    uint32_t p = 0, c = 7;  // positions past the queue sort last and do nothing
    if (j < n) {
        p = queueIn[j];
        queueOut[j] = p;
        c = min((uint32_t)category[j], (uint32_t)kCatDone);
    }
    if (t < 8) sCount[t] = 0;
    __syncthreads();
    const uint32_t rank = atomicAdd(&sCount[c], 1u);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < c; ++k) base += sCount[k];
    sEntry[base + rank] = p;
    sPos[base + rank] = (uint16_t)t;
    sCat[base + rank] = (uint8_t)c;
    __syncthreads();
    const uint32_t mc = sCat[t];
    if (mc > kCatDone) return;
    const uint32_t q = blockIdx.x * 256u + sPos[t];
    if (mc == kCatDone) {
        category[q] = (uint8_t)catFinished;
        return;
    }
    const uint32_t next = shade_path(S, prmTable, P, sEntry[t], (int)mc);
    category[q] = next == NEXT_DONE ? (uint8_t)catFinished : (uint8_t)(catBase + next);
}

// ----------------------------------------------------------- fused paths
// A chunk of paths rendered to completion in one launch, one path per thread:
// the camera ray, then per visit the stages of one wavefront iteration in
// its order (the shadow ray and its resolution, wrk_trace_any_resolve; the
// continuation ray and its resolution, wrk_trace_closest_resolve; the shade
// step, wrk_shade). Every decision reads only the path's own state, the
// closest hit does not depend on the traversal order (ties go to the smaller
// primitive id) and the arithmetic is the same code, so each path ends with
// the result the wavefront iteration gives it, bit for bit. A short request
// then costs one launch instead of one iteration of launches per bounce.
__global__ __launch_bounds__(256) void k_paths_fused(wrk_camera cam, SceneDev S, const Node8* __restrict__ nodes,
                                                     const float4* __restrict__ tris,
                                                     const wrk_render_params* __restrict__ prmTable, Pool P,
                                                     uint64_t first, uint32_t n, uint32_t poolBase, uint32_t slot) {
    __shared__ uint32_t sNode[kStackLds * 256];
    __shared__ float sT[kStackLds * 256];
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = poolBase + k;
    uint32_t* stN = sNode + threadIdx.x;
    float* stT = sT + threadIdx.x;
    const wrk_render_params prm = prmTable[slot];
    generate_path(cam, prm, P, first + k, i, slot);
    uint32_t next = NEXT_TRACE;
    bool fresh = true;  // not shaded yet: throughput 1, radiance 0, none stored
    for (;;) {
        uint32_t cat = CAT_DONE;
        if (next == NEXT_TRACE_SHADOW || next == NEXT_SHADOW) {
            // trace.cu:27-51: an unoccluded shadow ray adds its contribution
            float th;
            uint32_t ph;
            if (!trace_thread<true>(nodes, tris, P.savedOrg[i], P.savedDir[i], stN, stT, &th, &ph)) {
                const float4 L = P.rad[i];
                const float4 C = P.contrib[i];
                P.rad[i] = make_float4(L.x + C.x, L.y + C.y, L.z + C.z, L.w);
            }
            if (next == NEXT_SHADOW) {
                // the deferred continuation resumes in the Lambertian shader,
                // or the path ends and is flushed
                if (P.st[i].z & FLAG_NEE_DONE) {
                    cat = WRK_MAT_LAMBERTIAN;
                } else {
                    const float4 L = P.rad[i];
                    P.result[i] = f4(clamp_fireflies(mk(L.x, L.y, L.z), prm.fireflyClamp), 0.0f);
                    break;
                }
            }
        }
        if (next == NEXT_TRACE || next == NEXT_TRACE_SHADOW) {
            float tb;
            uint32_t bp;
            if (trace_thread<false>(nodes, tris, P.org[i], P.dir[i], stN, stT, &tb, &bp)) {
                const uint32_t prim = bp & kPrimMask;
                P.hit[i] = make_uint2(__float_as_uint(tb), prim);
                const uint32_t tag = bp >> 28;
                cat = tag != 0u ? tag - 1u : shade_category(S.triMat[prim], S.numMaterials, S.materials);
            } else {
                // trace.cu:143-148: environment term, then the flush
                const float4 T = fresh ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : P.thr[i];
                const float4 L = P.rad[i];
                const f3 Lr = mk(L.x + T.x * S.env[0], L.y + T.y * S.env[1], L.z + T.z * S.env[2]);
                P.rad[i] = f4(Lr, 0.0f);
                P.result[i] = f4(clamp_fireflies(Lr, prm.fireflyClamp), 0.0f);
                break;
            }
        }
        if (cat == CAT_DONE) {  // out-of-range material id: ends unflushed (shade.cu:35-38)
            P.result[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            break;
        }
        next = shade_path(S, prmTable, P, i, (int)cat);
        fresh = false;
        if (next == NEXT_DONE) break;
    }
}

// ----------------------------------------------------------- slot finish
// live[slot] -= paths of that slot among the entries that ended this
// iteration. A block takes 256 * perThread entries; each lane merges runs of
// one slot before adding them to an LDS histogram, and the block subtracts
// its nonzero bins from live[], so device atomics stay near one per block.
// MAP: the slot of pool entry p is slotMap[p >> 8] (a byte per 256 pool
// entries, L2-resident) instead of the flags word of its scattered state.
template <bool MAP>
__global__ __launch_bounds__(256) void k_slot_finish(const uint32_t* __restrict__ finished, uint32_t maxItems,
                                                     const uint32_t* __restrict__ countDev,
                                                     const uint4* __restrict__ st,
                                                     const uint8_t* __restrict__ slotMap, uint32_t bound,
                                                     uint32_t perThread, uint32_t* live) {
    __shared__ uint32_t hist[kMaxSlots];
    const uint32_t t = threadIdx.x;
    hist[t] = 0u;
    __syncthreads();
    uint32_t n = maxItems;
    if (countDev) n = min(*countDev, maxItems);
    const uint64_t start = (uint64_t)blockIdx.x * 256u * perThread;
    uint32_t cur = 0u, run = 0u;
    constexpr uint32_t kBatch = 8;
    for (uint32_t k0 = 0; k0 < perThread; k0 += kBatch) {
        uint32_t p[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            uint64_t i = start + (uint64_t)(k0 + u) * 256u + t;
            p[u] = (k0 + u < perThread && i < n) ? finished[i] : 0xFFFFFFFFu;
        }
        uint32_t sl[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            // entries past the map (or the pool) are skipped: a stale or
            // corrupt queue entry must not read beyond the buffer
            if (p[u] == 0xFFFFFFFFu) sl[u] = 0xFFFFFFFFu;
            else if constexpr (MAP) sl[u] = (p[u] >> 8) < bound ? (uint32_t)slotMap[p[u] >> 8] : 0xFFFFFFFFu;
            else sl[u] = p[u] < bound ? st[p[u]].z >> FLAG_SLOT_SHIFT : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (sl[u] == 0xFFFFFFFFu) continue;
            if (sl[u] != cur && run) {
                atomicAdd(&hist[cur], run);
                run = 0u;
            }
            cur = sl[u];
            ++run;
        }
    }
    if (run) atomicAdd(&hist[cur], run);
    __syncthreads();
    if (hist[t]) atomicSub(&live[t], hist[t]);
}

// accum[pixel] += result[path] over the batch in sample order; the
// first-hit guides likewise when requested.
__global__ __launch_bounds__(256) void k_accumulate(const float4* __restrict__ result,
                                                    const float4* __restrict__ guides, uint32_t poolBase,
                                                    uint64_t first, uint32_t n, uint32_t npix, float* accum,
                                                    float* accAlb, float* accNrm, float2* accNoise) {
    uint32_t k = blockIdx.x * 256 + threadIdx.x;
    uint32_t cover = n < npix ? n : npix;
    if (k >= cover) return;
    uint64_t g0 = first + k;  // first path of this pixel in the chunk
    uint32_t pixel = (uint32_t)(g0 % npix);
    float a0 = accum[pixel * 3 + 0], a1 = accum[pixel * 3 + 1], a2 = accum[pixel * 3 + 2];
    if (accNoise) {
        // first two moments of each sample's mean channel y = (r + g + b) / 3
        float2 m = accNoise[pixel];
        for (uint64_t i = k; i < n; i += npix) {
            float4 v = result[poolBase + i];
            a0 = a0 + v.x; a1 = a1 + v.y; a2 = a2 + v.z;
            const float y = (v.x + v.y + v.z) / 3.0f;
            m.x = m.x + y;
            m.y = m.y + y * y;
        }
        accNoise[pixel] = m;
    } else {
        for (uint64_t i = k; i < n; i += npix) {
            float4 v = result[poolBase + i];
            a0 = a0 + v.x; a1 = a1 + v.y; a2 = a2 + v.z;
        }
    }
    accum[pixel * 3 + 0] = a0; accum[pixel * 3 + 1] = a1; accum[pixel * 3 + 2] = a2;
    if (guides && accAlb && accNrm) {
        float b0 = accAlb[pixel * 3 + 0], b1 = accAlb[pixel * 3 + 1], b2 = accAlb[pixel * 3 + 2];
        float c0 = accNrm[pixel * 3 + 0], c1 = accNrm[pixel * 3 + 1], c2 = accNrm[pixel * 3 + 2];
        for (uint64_t i = k; i < n; i += npix) {
            float4 al = guides[2 * (poolBase + i)], nr = guides[2 * (poolBase + i) + 1];
            b0 = b0 + al.x; b1 = b1 + al.y; b2 = b2 + al.z;
            c0 = c0 + nr.x; c1 = c1 + nr.y; c2 = c2 + nr.z;
        }
        accAlb[pixel * 3 + 0] = b0; accAlb[pixel * 3 + 1] = b1; accAlb[pixel * 3 + 2] = b2;
        accNrm[pixel * 3 + 0] = c0; accNrm[pixel * 3 + 1] = c1; accNrm[pixel * 3 + 2] = c2;
    }
}

__global__ __launch_bounds__(256) void k_average(const float* __restrict__ accum, uint32_t n3, float invSpp,
                                                 float* out) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n3) out[i] = accum[i] * invSpp;
}

// Standard error of a pixel's mean y from its sample moments (sum y, sum
// y^2) over spp samples: sqrt(max(s2 - s1^2 / n, 0) / (n - 1) / n); with one
// sample, |y| itself.
__global__ __launch_bounds__(256) void k_noise_stderr(const float2* __restrict__ moments, uint32_t npix,
                                                      uint32_t spp, float* out) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npix) return;
    const float2 m = moments[i];
    if (spp < 2) {
        out[i] = fabsf(m.x);
        return;
    }
    const float n = (float)spp;
    const float var = fmaxf(m.y - m.x * m.x / n, 0.0f) / (n - 1.0f);
    out[i] = sqrtf(var / n);
}

// libWR/test/helpers/image.cpp:14-49 (gamma in FP64, see oracle_capi.cpp)
__global__ __launch_bounds__(256) void k_aces(const float* __restrict__ img, uint32_t n3, uint8_t* out) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n3) return;
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float x = img[i] * 2.2f;
    float y = (x * (a * x + b)) / (x * (c * x + d) + e);
    y = y < 0.0f ? 0.0f : (y > 1.0f ? 1.0f : y);
    const float gamma = 1.0f / 2.2f;
    float g = (float)pow((double)y, (double)gamma);
    out[i] = (uint8_t)(g * 255.0f);
}

}  // namespace wrk

using namespace wrk;

extern "C" int wrk_generate(const wrk_camera* cam, const wrk_render_params* params, const wrk_paths* paths,
                            uint64_t first, uint32_t n, uint32_t poolBase, uint32_t slot, uint32_t* queue,
                            uint8_t* category, uint32_t categoryValue, void* stream) {
    WRK_REQUIRE(cam && params && paths && queue, "wrk_generate: null argument");
    WRK_REQUIRE((uint64_t)poolBase + n <= paths->capacity, "wrk_generate: range exceeds the path pool capacity");
    WRK_REQUIRE(params->width > 0 && params->height > 0, "wrk_generate: empty image");
    WRK_REQUIRE(slot < kMaxSlots, "wrk_generate: slot out of range");
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_generate, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), *cam, *params,
                       pool_of(paths), first, n, poolBase, slot, queue, category, categoryValue);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_classify(const wrk_scene_view* scene, const wrk_render_params* paramsTable, const wrk_paths* paths,
                            const uint32_t* queue, uint32_t numTrace, uint32_t numTraceShadow, uint32_t numShadow,
                            uint8_t* category, uint32_t flags, void* stream) {
    WRK_REQUIRE(scene && paramsTable && paths && queue && category, "wrk_classify: null argument");
    const uint64_t n = (uint64_t)numTrace + numTraceShadow + numShadow;
    WRK_REQUIRE(n <= paths->capacity, "wrk_classify: queue longer than the path pool");
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_classify, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), scene_of(scene),
                       paramsTable, pool_of(paths), queue, numTrace, numTraceShadow, numShadow, category, flags & 1u);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_shade(const wrk_scene_view* scene, const wrk_render_params* paramsTable, const wrk_paths* paths,
                         const uint32_t* queueIn, uint32_t* queueOut, uint32_t n, uint32_t categoryBase,
                         uint32_t finishedCategory, uint8_t* category, void* stream) {
    WRK_REQUIRE(scene && paramsTable && paths && queueIn && queueOut && category, "wrk_shade: null argument");
    WRK_REQUIRE(categoryBase + 2u <= 0xFFu && finishedCategory <= 0xFFu, "wrk_shade: category out of range");
    WRK_REQUIRE(n <= paths->capacity, "wrk_shade: queue longer than the path pool");
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_shade, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), scene_of(scene),
                       paramsTable, pool_of(paths), queueIn, queueOut, n, categoryBase, finishedCategory, category);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_render_fused(const wrk_bvh* bvh, const wrk_scene_view* scene, const wrk_camera* camera,
                                const wrk_render_params* paramsTable, uint32_t slot, const wrk_paths* paths,
                                uint64_t firstPath, uint32_t n, uint32_t poolBase, void* stream) {
    WRK_REQUIRE(bvh && scene && camera && paramsTable && paths, "wrk_render_fused: null argument");
    WRK_REQUIRE(bvh->nodes && bvh->tris && bvh->nodeBytes == sizeof(Node8), "wrk_render_fused: unknown BVH format");
    WRK_REQUIRE((kBvhWidth - 1u) * bvh->maxDepth + 8u <= kStackTotal,
                "wrk_render_fused: BVH deeper than the traversal stack");
    WRK_REQUIRE(slot < WRK_MAX_SLOTS, "wrk_render_fused: slot out of range");
    WRK_REQUIRE((uint64_t)poolBase + n <= paths->capacity, "wrk_render_fused: range exceeds the path pool");
    WRK_REQUIRE(paths->org && paths->dir && paths->thr && paths->rad && paths->state && paths->contrib &&
                    paths->savedOrg && paths->savedDir && paths->hit && paths->result,
                "wrk_render_fused: incomplete path pool");
    if (n == 0) return WRK_OK;
    hipLaunchKernelGGL(k_paths_fused, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), *camera,
                       scene_of(scene), (const Node8*)bvh->nodes, (const float4*)bvh->tris, paramsTable,
                       pool_of(paths), firstPath, n, poolBase, slot);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_slot_finish(const uint32_t* finished, uint32_t maxItems, const uint32_t* countDev,
                               const wrk_paths* paths, uint32_t* live, void* stream) {
    WRK_REQUIRE(paths && live, "wrk_slot_finish: null argument");
    if (maxItems == 0) return WRK_OK;
    WRK_REQUIRE(finished, "wrk_slot_finish: null argument");
    // about 2048 blocks for long lists, one entry per lane for short ones
    uint32_t perThread = ceil_div(maxItems, 256ull * 2048ull);
    perThread = perThread < 1u ? 1u : (perThread > 256u ? 256u : perThread);
    uint32_t blocks = ceil_div(maxItems, 256ull * perThread);
    hipLaunchKernelGGL(k_slot_finish<false>, dim3(blocks), dim3(256), 0, as_stream(stream), finished, maxItems,
                       countDev, (const uint4*)paths->state, nullptr, paths->capacity, perThread, live);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_slot_finish_map(const uint32_t* finished, uint32_t maxItems, const uint32_t* countDev,
                                   const uint8_t* slotMap, uint32_t numGranules, uint32_t* live, void* stream) {
    WRK_REQUIRE(slotMap && live, "wrk_slot_finish_map: null argument");
    if (maxItems == 0) return WRK_OK;
    WRK_REQUIRE(finished, "wrk_slot_finish_map: null argument");
    WRK_REQUIRE((uint64_t)maxItems <= 256ull * numGranules,
                "wrk_slot_finish_map: more finished entries than the slot map covers");
    uint32_t perThread = ceil_div(maxItems, 256ull * 2048ull);
    perThread = perThread < 1u ? 1u : (perThread > 256u ? 256u : perThread);
    uint32_t blocks = ceil_div(maxItems, 256ull * perThread);
    hipLaunchKernelGGL(k_slot_finish<true>, dim3(blocks), dim3(256), 0, as_stream(stream), finished, maxItems,
                       countDev, nullptr, slotMap, numGranules, perThread, live);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_accumulate(const wrk_paths* paths, uint32_t poolBase, uint64_t first, uint32_t n,
                              uint32_t npix, float* accum, float* accumAlbedo, float* accumNormal, float* accumNoise,
                              void* stream) {
    WRK_REQUIRE(paths && accum && npix > 0, "wrk_accumulate: bad argument");
    WRK_REQUIRE((uint64_t)poolBase + n <= paths->capacity, "wrk_accumulate: range exceeds the path pool");
    if (n == 0) return WRK_OK;
    uint32_t cover = n < npix ? n : npix;
    hipLaunchKernelGGL(k_accumulate, dim3(ceil_div(cover, 256)), dim3(256), 0, as_stream(stream),
                       (const float4*)paths->result, (const float4*)paths->guides, poolBase, first, n, npix, accum,
                       accumAlbedo, accumNormal, (float2*)accumNoise);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_average(const float* accum, uint32_t npix, uint32_t spp, float* out, void* stream) {
    WRK_REQUIRE(accum && out && spp > 0, "wrk_average: bad argument");
    uint32_t n3 = npix * 3;
    if (n3 == 0) return WRK_OK;
    hipLaunchKernelGGL(k_average, dim3(ceil_div(n3, 256)), dim3(256), 0, as_stream(stream), accum, n3,
                       1.0f / (float)spp, out);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_noise_stderr(const float* moments, uint32_t npix, uint32_t spp, float* out, void* stream) {
    WRK_REQUIRE(moments && out && spp > 0, "wrk_noise_stderr: bad argument");
    if (npix == 0) return WRK_OK;
    hipLaunchKernelGGL(k_noise_stderr, dim3(ceil_div(npix, 256)), dim3(256), 0, as_stream(stream),
                       (const float2*)moments, npix, spp, out);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}

extern "C" int wrk_tonemap_aces(const float* image, uint32_t npix, uint8_t* rgb8, void* stream) {
    WRK_REQUIRE(image && rgb8, "wrk_tonemap_aces: null argument");
    uint32_t n3 = npix * 3;
    if (n3 == 0) return WRK_OK;
    hipLaunchKernelGGL(k_aces, dim3(ceil_div(n3, 256)), dim3(256), 0, as_stream(stream), image, n3, rgb8);
    WRK_CHECK_LAUNCH();
    return WRK_OK;
}
