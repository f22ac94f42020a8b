// wr_capi.cpp — C ABI over the wr:: C++ API (see include/wr_c.h).
#include <stdexcept>
#include <string>

#include "../../include/wr/wr.h"
#include "../../include/wr_c.h"

namespace {
thread_local std::string g_err;
// Null handles and null required pointers become errors, not crashes.
template <typename T> T* need(T* p, const char* what) {
    if (!p) throw std::runtime_error(std::string(what) + ": null pointer");
    return p;
}
wr::Camera to_camera(const wr_c_camera* c) {
    need(c, "camera");
    wr::Camera k;
    k.position = wr::Vec3(c->position[0], c->position[1], c->position[2]);
    k.target = wr::Vec3(c->target[0], c->target[1], c->target[2]);
    k.up = wr::Vec3(c->up[0], c->up[1], c->up[2]);
    k.fovY = c->fovY;
    k.aspect = c->aspect;
    k.focalDistance = c->focalDistance;
    k.lensRadius = c->lensRadius;
    return k;
}
wr::RenderParams to_params(const wr_c_render_params* p) {
    need(p, "render params");
    wr::RenderParams r;
    r.width = p->width;
    r.height = p->height;
    r.spp = p->spp;
    r.useNEE = p->useNEE != 0;
    r.maxBounces = p->maxBounces;
    r.russianRouletteDepth = p->russianRouletteDepth;
    r.fireflyClamp = p->fireflyClamp;
    r.denoiser.enabled = p->denoiser != 0;
    r.firstSample = p->firstSample;
    return r;
}
template <typename F> int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown exception";
        return -1;
    }
}
template <typename F> uint32_t guard_id(F&& f) {
    uint32_t id = 0xFFFFFFFFu;
    if (guard([&] { id = f(); }) != 0) return 0xFFFFFFFFu;
    return id;
}
wr::Scene* S(void* p) { return need(static_cast<wr::Scene*>(p), "scene handle"); }
wr::RenderServer* SV(void* p) { return need(static_cast<wr::RenderServer*>(p), "server handle"); }
void fill_stats(const wr::RenderStats& st, wr_c_stats* out) {
    out->seconds = st.seconds;
    out->paths = st.paths;
    out->closestRays = st.closestRays;
    out->shadowRays = st.shadowRays;
    out->iterations = st.iterations;
    out->batches = st.batches;
    out->msGenerate = st.msGenerate; out->msTraceClosest = st.msTraceClosest; out->msTraceAny = st.msTraceAny;
    out->msClassify = st.msClassify; out->msPartition = st.msPartition; out->msShade = st.msShade;
    out->msAccumulate = st.msAccumulate;
    out->launchesTraceClosest = st.launchesTraceClosest; out->launchesTraceAny = st.launchesTraceAny;
    out->closestNodeFetches = st.closestNodeFetches; out->closestTriFetches = st.closestTriFetches;
    out->anyNodeFetches = st.anyNodeFetches; out->anyTriFetches = st.anyTriFetches;
    out->classifyItems = st.classifyItems; out->partitionItems = st.partitionItems; out->shadeItems = st.shadeItems;
    out->closestStackSpills = st.closestStackSpills; out->anyStackSpills = st.anyStackSpills;
    out->fusedPaths = st.fusedPaths;
    out->msFused = st.msFused;
}
wr::Renderer* R(void* p) { return need(static_cast<wr::Renderer*>(p), "renderer handle"); }
}  // namespace

extern "C" {

const char* wr_c_last_error(void) { return g_err.c_str(); }

int wr_c_context_create(int deviceId, int logLevel) {
    return guard([&] {
        wr::ContextConfig cfg;
        cfg.deviceId = deviceId;
        cfg.logLevel = logLevel;
        wr::Context ctx(cfg);
    });
}

void* wr_c_scene_create(void) {
    void* out = nullptr;
    guard([&] { out = new wr::Scene(); });
    return out;
}
void wr_c_scene_destroy(void* scene) { delete static_cast<wr::Scene*>(scene); }

uint32_t wr_c_scene_add_lambertian(void* s, float r, float g, float b, uint32_t tex) {
    return guard_id([&] { return S(s)->addLambertianMaterial(wr::Vec3(r, g, b), tex); });
}
uint32_t wr_c_scene_add_emissive(void* s, float r, float g, float b) {
    return guard_id([&] { return S(s)->addEmissiveMaterial(wr::Vec3(r, g, b)); });
}
uint32_t wr_c_scene_add_glass(void* s, float r, float g, float b, float ior) {
    return guard_id([&] { return S(s)->addGlassMaterial(wr::Vec3(r, g, b), ior); });
}
uint32_t wr_c_scene_add_ggx_reflection(void* s, float r, float g, float b, float rough, float metal) {
    return guard_id([&] { return S(s)->addGGXReflectionMaterial(wr::Vec3(r, g, b), rough, metal); });
}
uint32_t wr_c_scene_add_ggx_transmission(void* s, float r, float g, float b, float rough, float ior) {
    return guard_id([&] { return S(s)->addGGXTransmissionMaterial(wr::Vec3(r, g, b), rough, ior); });
}
int wr_c_scene_set_normal_map(void* s, uint32_t mat, uint32_t tex) {
    return guard([&] { S(s)->setMaterialNormalMap(mat, tex); });
}
int wr_c_scene_add_mesh(void* s, const float* v, size_t nv, const uint32_t* i, size_t ni, const float* n, size_t nn,
                        const float* uv, size_t nuv, uint32_t mat) {
    return guard([&] {
        if ((!v && nv) || (!i && ni)) throw std::runtime_error("addTriangleMesh: null vertex or index array");
        S(s)->addTriangleMesh(std::span<const float>(v, nv), std::span<const uint32_t>(i, ni),
                              std::span<const float>(n, n ? nn : 0), std::span<const float>(uv, uv ? nuv : 0), mat);
    });
}
uint32_t wr_c_texture_create(void* s, const char* path) {
    uint32_t id = 0;
    if (guard([&] { id = wr::Texture2D::create(S(s), need(path, "texture path")); }) != 0) return 0;
    return id;
}
void wr_c_texture_destroy(void* s, uint32_t id) { guard([&] { wr::Texture2D::destroy(S(s), id); }); }
uint32_t wr_c_scene_add_texture_rgba8(void* s, const uint8_t* rgba, uint32_t w, uint32_t h) {
    uint32_t id = 0;
    if (guard([&] {
            id = S(s)->addTextureRGBA8(w, h, std::span<const uint8_t>(need(rgba, "texels"), (size_t)w * h * 4));
        }) != 0)
        return 0;
    return id;
}
int wr_c_scene_set_environment_radiance(void* s, float r, float g, float b) {
    return guard([&] { S(s)->setEnvironmentRadiance(wr::Vec3(r, g, b)); });
}
int wr_c_scene_set_environment_map(void* s, const char* path) {
    return guard([&] { S(s)->setEnvironmentMap(need(path, "environment map path")); });
}
int wr_c_scene_finalize(void* s) { return guard([&] { S(s)->finalize(); }); }
int wr_c_scene_get_bvh(void* s, wrk_bvh* out) { return guard([&] { *need(out, "bvh") = S(s)->getBVH(); }); }
int wr_c_scene_get_view(void* s, wrk_scene_view* out) {
    return guard([&] { *need(out, "scene view") = S(s)->getDeviceView(); });
}
double wr_c_scene_build_seconds(void* s) {
    double t = -1.0;
    if (guard([&] { t = S(s)->getBuildSeconds(); }) != 0) return -1.0;
    return t;
}

void* wr_c_renderer_create(void) {
    void* out = nullptr;
    guard([&] { out = new wr::Renderer(); });
    return out;
}
void wr_c_renderer_destroy(void* r) { delete static_cast<wr::Renderer*>(r); }
int wr_c_renderer_set_wavefront(void* r, uint32_t cap, uint32_t maxPaths, uint32_t maxChunk, uint32_t flags) {
    return guard([&] {
        wr::WavefrontConfig c;
        c.iterationCap = cap;
        c.maxPathsInFlight = maxPaths;
        c.maxChunkPaths = maxChunk;
        c.timeKernels = (flags & 1u) != 0;
        c.countTraversal = (flags & 2u) != 0;
        c.fusedMaxPaths = R(r)->getWavefrontConfig().fusedMaxPaths;
        R(r)->setWavefrontConfig(c);
    });
}
int wr_c_renderer_set_fused(void* r, uint64_t fusedMaxPaths) {
    return guard([&] {
        wr::WavefrontConfig c = R(r)->getWavefrontConfig();
        c.fusedMaxPaths = fusedMaxPaths;
        R(r)->setWavefrontConfig(c);
    });
}
int wr_c_render(void* r, void* s, const wr_c_camera* cam, const wr_c_render_params* p, float* out) {
    return guard([&] {
        R(r)->render(S(s), to_camera(cam), reinterpret_cast<wr::Vec3*>(need(out, "output image")), to_params(p));
    });
}
int wr_c_render_paths(void* r, void* s, const wr_c_camera* cam, const wr_c_render_params* p, uint64_t first,
                      uint32_t count, float* out) {
    return guard([&] { R(r)->renderPaths(S(s), to_camera(cam), to_params(p), first, count, need(out, "paths")); });
}
int wr_c_render_guides(void* r, void* s, const wr_c_camera* cam, const wr_c_render_params* p, float* color,
                       float* albedo, float* normal, float* noise) {
    return guard([&] {
        R(r)->renderWithGuides(S(s), to_camera(cam), to_params(p), reinterpret_cast<wr::Vec3*>(need(color, "colour")),
                               reinterpret_cast<wr::Vec3*>(albedo), reinterpret_cast<wr::Vec3*>(normal), noise);
    });
}
int wr_c_renderer_load_denoiser(void* r, const char* path) {
    return guard([&] { R(r)->loadDenoiser(need(path, "denoiser path")); });
}
int wr_c_renderer_stats(void* r, wr_c_stats* out) {
    return guard([&] { fill_stats(R(r)->getStats(), need(out, "stats")); });
}

void* wr_c_server_create(const wr_c_server_config* cfg) {
    void* out = nullptr;
    guard([&] {
        need(cfg, "server config");
        wr::ServerConfig c;
        c.maxPathsInFlight = cfg->maxPathsInFlight;
        c.maxChunkPaths = cfg->maxChunkPaths;
        c.memoryBudgetBytes = cfg->memoryBudgetBytes;
        c.iterationCap = cfg->iterationCap;
        c.timeKernels = (cfg->flags & 1u) != 0;
        c.shortestFirst = (cfg->flags & 2u) != 0;
        c.roundRobin = (cfg->flags & 4u) != 0;
        c.expressOverlap = (cfg->flags & 8u) != 0;
        c.expressMaxPaths = cfg->expressMaxPaths;
        if (cfg->expressPathsInFlight) c.expressPathsInFlight = cfg->expressPathsInFlight;
        if (cfg->expressIterations) c.expressIterations = cfg->expressIterations;
        c.fusedMaxPaths = cfg->fusedMaxPaths;
        c.backfillPaths = cfg->backfillPaths;
        out = new wr::RenderServer(c);
    });
    return out;
}
void wr_c_server_destroy(void* sv) { delete static_cast<wr::RenderServer*>(sv); }
uint64_t wr_c_server_submit(void* sv, void* scene, const wr_c_camera* cam, const wr_c_render_params* p) {
    uint64_t id = 0;
    if (guard([&] { id = SV(sv)->submit(S(scene), to_camera(cam), to_params(p)); }) != 0) return 0;
    return id;
}
int wr_c_server_step(void* sv) {
    int more = 0;
    if (guard([&] { more = SV(sv)->step() ? 1 : 0; }) != 0) return -1;
    return more;
}
int wr_c_server_status(void* sv, uint64_t id, wr_c_request_status* out) {
    return guard([&] {
        wr::RequestStatus st = SV(sv)->status(id);
        need(out, "status");
        out->done = st.done;
        out->admitted = st.admitted;
        out->submitTime = st.submitTime;
        out->admitTime = st.admitTime;
        out->doneTime = st.doneTime;
        out->paths = st.paths;
    });
}
int wr_c_server_take_completed(void* sv, uint64_t* ids, uint32_t maxIds) {
    int n = 0;
    if (guard([&] { n = (int)SV(sv)->takeCompleted(maxIds ? need(ids, "ids") : ids, maxIds); }) != 0) return -1;
    return n;
}
int wr_c_server_fetch(void* sv, uint64_t id, float* out) {
    return guard([&] { SV(sv)->fetch(id, reinterpret_cast<wr::Vec3*>(out)); });
}
int wr_c_server_stats(void* sv, wr_c_stats* out) {
    return guard([&] { fill_stats(SV(sv)->getStats(), need(out, "stats")); });
}
uint32_t wr_c_server_pending(void* sv) {
    uint32_t n = 0;
    guard([&] { n = SV(sv)->numPending(); });
    return n;
}
uint32_t wr_c_server_in_flight(void* sv) {
    uint32_t n = 0;
    guard([&] { n = SV(sv)->numInFlight(); });
    return n;
}
double wr_c_server_clock(void) { return wr::RenderServer::clock(); }

}  // extern "C"
