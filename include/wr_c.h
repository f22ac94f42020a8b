/*
 * wr_c.h — C ABI over the wr:: C++ API, for foreign-function callers
 * (Python ctypes, Go cgo, JNI, N-API). Each entry point forwards to the
 * wr:: method named in its comment (signatures of the libWR/include/wr/ headers) and
 * turns C++ exceptions into a negative return code, with the exception's
 * message available from wr_c_last_error().
 */
#ifndef WR_C_H
#define WR_C_H

#include <stdint.h>
#include <stddef.h>

#include "wrk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wr_c_camera {
    float position[3], target[3], up[3];
    float fovY, aspect, focalDistance, lensRadius;
} wr_c_camera;

typedef struct wr_c_render_params {
    uint32_t width, height, spp, useNEE, maxBounces;
    float russianRouletteDepth, fireflyClamp;
    uint32_t denoiser;
    uint32_t firstSample;
} wr_c_render_params;

typedef struct wr_c_stats {
    double seconds;
    uint64_t paths, closestRays, shadowRays;
    uint32_t iterations, batches;
    double msGenerate, msTraceClosest, msTraceAny, msClassify, msPartition, msShade, msAccumulate;
    uint32_t launchesTraceClosest, launchesTraceAny;
    uint64_t closestNodeFetches, closestTriFetches, anyNodeFetches, anyTriFetches;
    uint64_t classifyItems, partitionItems, shadeItems;
    uint64_t closestStackSpills, anyStackSpills;
    uint64_t fusedPaths;
    double msFused;
} wr_c_stats;

const char* wr_c_last_error(void);
int wr_c_context_create(int deviceId, int logLevel);                 /* wr::Context(ContextConfig) */
void* wr_c_scene_create(void);                                       /* Context::createScene */
void wr_c_scene_destroy(void* scene);
uint32_t wr_c_scene_add_lambertian(void* scene, float r, float g, float b, uint32_t albedoTextureId);
uint32_t wr_c_scene_add_emissive(void* scene, float r, float g, float b);
uint32_t wr_c_scene_add_glass(void* scene, float r, float g, float b, float ior);
uint32_t wr_c_scene_add_ggx_reflection(void* scene, float r, float g, float b, float roughness, float metallic);
uint32_t wr_c_scene_add_ggx_transmission(void* scene, float r, float g, float b, float roughness, float ior);
int wr_c_scene_set_normal_map(void* scene, uint32_t materialId, uint32_t textureId);
int wr_c_scene_add_mesh(void* scene, const float* vertices, size_t numVertexFloats, const uint32_t* indices,
                        size_t numIndices, const float* normals, size_t numNormalFloats, const float* uvs,
                        size_t numUvFloats, uint32_t materialId);
/* returns the 1-based texture id, 0 on error */
uint32_t wr_c_texture_create(void* scene, const char* pngPath);      /* Texture2D::create */
void wr_c_texture_destroy(void* scene, uint32_t textureId);          /* Texture2D::destroy */
uint32_t wr_c_scene_add_texture_rgba8(void* scene, const uint8_t* rgba, uint32_t width, uint32_t height);
int wr_c_scene_set_environment_radiance(void* scene, float r, float g, float b);
int wr_c_scene_set_environment_map(void* scene, const char* hdrPath);
int wr_c_scene_finalize(void* scene);                                /* Scene::finalize */
int wr_c_scene_get_bvh(void* scene, wrk_bvh* out);
int wr_c_scene_get_view(void* scene, wrk_scene_view* out);
double wr_c_scene_build_seconds(void* scene);

void* wr_c_renderer_create(void);                                    /* Context::createRenderer */
void wr_c_renderer_destroy(void* renderer);
/* flags: bit0 time kernels with HIP events, bit1 count traversal fetches */
int wr_c_renderer_set_wavefront(void* renderer, uint32_t iterationCap, uint32_t maxPathsInFlight, uint32_t maxChunkPaths,
                                uint32_t flags);
/* WavefrontConfig::fusedMaxPaths of the renderer (kept by later
 * wr_c_renderer_set_wavefront calls) */
int wr_c_renderer_set_fused(void* renderer, uint64_t fusedMaxPaths);
/* Renderer::render; out = width*height*3 floats (averaged linear RGB) */
int wr_c_render(void* renderer, void* scene, const wr_c_camera* cam, const wr_c_render_params* params, float* out);
/* Renderer::renderPaths; out = count*4 floats */
int wr_c_render_paths(void* renderer, void* scene, const wr_c_camera* cam, const wr_c_render_params* params,
                      uint64_t firstPath, uint32_t count, float* out);
int wr_c_renderer_stats(void* renderer, wr_c_stats* out);
/* Renderer::loadDenoiser */
int wr_c_renderer_load_denoiser(void* renderer, const char* path);
/* Renderer::renderWithGuides; colour, albedo and normal width*height*3
 * floats each; noise (NULL skips) width*height floats */
int wr_c_render_guides(void* renderer, void* scene, const wr_c_camera* cam, const wr_c_render_params* params,
                       float* color, float* albedo, float* normal, float* noise);

/* --- continuous-batching render server (wr::RenderServer) --- */
typedef struct wr_c_server_config {
    uint32_t maxPathsInFlight, maxChunkPaths;
    uint64_t memoryBudgetBytes;
    uint32_t iterationCap, flags; /* flags bit0: time kernels, bit1: shortest request first,
                                     bit2: round-robin admission (one chunk per request per round),
                                     bit3: express iterations overlap the main iteration */
    uint32_t expressMaxPaths;       /* ServerConfig::expressMaxPaths (0 = no express lane) */
    uint32_t expressPathsInFlight;  /* express pool size (0 = ServerConfig's default) */
    uint32_t expressIterations;     /* express iterations per step (0 = ServerConfig's default) */
    uint32_t fusedMaxPaths;         /* ServerConfig::fusedMaxPaths (0 = bounce iterations only) */
    uint32_t backfillPaths;         /* ServerConfig::backfillPaths (0 = no limit) */
} wr_c_server_config;

typedef struct wr_c_request_status {
    uint32_t done, admitted;
    double submitTime, admitTime, doneTime;
    uint64_t paths;
} wr_c_request_status;

void* wr_c_server_create(const wr_c_server_config* cfg);            /* RenderServer(ServerConfig) */
void wr_c_server_destroy(void* server);
/* returns the request id, 0 on error */
uint64_t wr_c_server_submit(void* server, void* scene, const wr_c_camera* cam, const wr_c_render_params* params);
/* 1 = work remains, 0 = idle, -1 = error */
int wr_c_server_step(void* server);
int wr_c_server_status(void* server, uint64_t id, wr_c_request_status* out);
/* RenderServer::takeCompleted: writes up to maxIds finished request ids, returns the count (-1 on error) */
int wr_c_server_take_completed(void* server, uint64_t* ids, uint32_t maxIds);
int wr_c_server_fetch(void* server, uint64_t id, float* out);      /* width*height*3 floats */
int wr_c_server_stats(void* server, wr_c_stats* out);
uint32_t wr_c_server_pending(void* server);
uint32_t wr_c_server_in_flight(void* server);
double wr_c_server_clock(void);

#ifdef __cplusplus
}
#endif

#endif
