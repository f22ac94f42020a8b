/*
 * wrk.h — C ABI of the MI355X wavefront path-tracing kernels.
 *
 * This is the seam that replaces libWR's OptiX/CUDA layer: the OptiX pipeline
 * seam createPipeline/launchPipeline/destroyPipeline
 * (libWR/src/internal/renderer.cpp:18-21, pipeline.cpp:169-190), the CUBIN
 * kernels "shade" and "compact" fetched by name (renderer.cpp:50-59) and the
 * OptiX GAS build (scene.cpp:493-550). The C++ wr::Scene / wr::Renderer layer
 * is the only product caller; tests call it directly through ctypes.
 *
 * Conventions: plain C types only; device memory is passed as pointers the
 * caller allocated (no function here allocates device memory unless its name
 * says so); `stream` is a hipStream_t passed as void* (NULL = default stream);
 * every function returns WRK_OK or a negative code, with a message available
 * from wrk_last_error().
 */
#ifndef WRK_H
#define WRK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    WRK_OK = 0,
    WRK_ERR_INVALID = -1,   /* bad argument (null pointer, size out of range) */
    WRK_ERR_HIP = -2,       /* HIP runtime call failed */
    WRK_ERR_CAPACITY = -3,  /* input exceeds a documented limit */
};

/* Material kinds: libWR/src/internal/gpu_types.h:11-17 */
enum {
    WRK_MAT_LAMBERTIAN = 0,
    WRK_MAT_EMISSIVE = 1,
    WRK_MAT_GLASS = 2,
    WRK_MAT_GGX_REFLECTION = 3,
    WRK_MAT_GGX_TRANSMISSION = 4,
};

/* 64-byte material record, field for field the reference MaterialData
 * (gpu_types.h:28-37): albedo.xyz, emission.xyz, ior, type, roughness,
 * metallic, 1-based texture ids (0 = none). */
typedef struct wrk_material {
    float albedo[4];
    float emission[4];
    float ior;
    uint32_t type;
    float roughness;
    float metallic;
    uint32_t albedoTextureId;
    uint32_t normalTextureId;
    uint32_t _pad[2];
} wrk_material;

/* RGBA8 texture, row-major, row 0 = v in [0, 1/height). */
typedef struct wrk_texture {
    uint64_t texels; /* device pointer to width*height*4 bytes, 0 = destroyed */
    uint32_t width;
    uint32_t height;
} wrk_texture;

/* Device view of a finalized scene (the getters of libWR/include/wr/scene.h:72-95). */
typedef struct wrk_scene_view {
    uint64_t vertices;       /* float[3V] */
    uint64_t normals;        /* float[3V], 0 if no mesh supplied normals */
    uint64_t uvs;            /* float[2V], 0 if no mesh supplied uvs */
    uint64_t indices;        /* uint32[3T] */
    uint64_t triMaterialIds; /* uint32[T] */
    uint64_t materials;      /* wrk_material[numMaterials] */
    uint64_t textures;       /* wrk_texture[numTextures] or 0 */
    uint64_t emissiveTris;   /* uint32[numEmissive] */
    uint64_t emissiveCdf;    /* float[numEmissive] */
    uint32_t numTris;
    uint32_t numMaterials;
    uint32_t numTextures;
    uint32_t numEmissive;
    float envRadiance[4];
    /* Optional guide of the CDF search (0 = plain binary search): for a
     * power of two G = emissiveGuideSize, guide[k] (k = 0..G) is the first
     * index i with emissiveCdf[i] >= k / G. A sample u lies in bucket
     * k = floor(u * G) (exact: G is a power of two), so the search runs over
     * [guide[k], guide[k + 1]] and returns the same index as a search over
     * the whole CDF. */
    uint64_t emissiveGuide;  /* uint32[emissiveGuideSize + 1] or 0 */
    uint32_t emissiveGuideSize;
    uint32_t _pad2[3];
} wrk_scene_view;

/* BVH over leaf-ordered triangles; nodeBytes gives the node format (96).
 * Compressed 8-wide node = float origin[3] and the three biased float
 * exponents e_x, e_y, e_z in bytes 12..14 (16 B), child[8] (32 B), then
 * 8-bit planes qlo.x[8] qlo.y[8] qlo.z[8] qhi.x[8] qhi.y[8] qhi.z[8] (48 B);
 * plane = fma(q, 2^(e - 127), origin), rounded outward at build time.
 * child: bit31 clear -> inner node index; bit31 set -> leaf with
 * count-1 in bits 27..30 and first leaf-ordered triangle in bits 0..26
 * (at most 2 triangles per leaf); 0xFFFFFFFF -> empty slot (its box is
 * inverted; used slots come first). Triangles: 3 float4 each, (v0.xyz,
 * original primitive id as bits), (v1.xyz, 0 or 1 + the triangle's shade
 * category after wrk_bvh_tag_categories), (v2.xyz, 0). */
typedef struct wrk_bvh {
    uint64_t nodes;
    uint64_t tris;
    uint32_t numNodes;
    uint32_t numTris;
    uint32_t maxDepth;
    uint32_t nodeBytes; /* 96: compressed 8-wide nodes */
    float sceneLo[4];
    float sceneHi[4];
} wrk_bvh;

/* Camera as the reference CameraData (gpu_types.h:89-98), basis precomputed
 * on the host (renderer.cpp:144-156). */
typedef struct wrk_camera {
    float position[4];
    float forward[4];
    float right[4];
    float up[4];
    float tanHalfFovY;
    float aspect;
    float focalDistance;
    float lensRadius;
} wrk_camera;

/* Render parameters (libWR/include/wr/renderer.h:25-36 RenderParams plus the
 * host-loop iteration cap of renderer.cpp:213; iterationCap 0 = no cap). */
typedef struct wrk_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t useNEE;
    uint32_t maxBounces;
    float rrStartDepth;
    float fireflyClamp;
    uint32_t iterationCap;
    uint32_t _pad;
} wrk_render_params;

/* Structure-of-arrays path pool (capacity paths). Replaces the 116-byte
 * RayState + 56-byte HitInfo AoS of gpu_types.h:48-86. Paths of one chunk
 * occupy a contiguous range, so queue-ordered accesses stay coalesced. */
typedef struct wrk_paths {
    uint64_t org;      /* float4: origin.xyz, tMin */
    uint64_t dir;      /* float4: direction.xyz, tMax */
    uint64_t thr;      /* float4: throughput.xyz, prevPdf */
    uint64_t rad;      /* float4: radiance.xyz, - */
    uint64_t state;    /* uint4: seed, rngDimension, flags, pixelIndex */
    uint64_t contrib;  /* float4: pending NEE contribution.xyz */
    uint64_t savedOrg; /* float4: shadow ray origin.xyz, tMin */
    uint64_t savedDir; /* float4: shadow ray direction.xyz, tMax */
    uint64_t hit;      /* uint2: t (float bits), primitive id (0xFFFFFFFF = miss) */
    uint64_t occluded; /* uint8 */
    uint64_t result;   /* float4: value added to the accumulation buffer */
    uint64_t guides;   /* optional (0 = none), 2 x float4 per path: first-hit albedo.xyz, normal.xyz */
    uint32_t capacity;
    uint32_t _pad;
} wrk_paths;

const char* wrk_last_error(void);
int wrk_device_count(int* count);

/* ---------------------------------------------------------------- BVH --- */
/* Sizes of the caller-provided buffers for wrk_bvh_build. */
int wrk_bvh_query_sizes(uint32_t numTris, uint64_t* nodeBytes, uint64_t* triBytes,
                        uint64_t* scratchBytes);
/* GPU LBVH: 63-bit Morton codes of triangle centroids (21 bits per axis),
 * the LDS-staged radix sort of wrk_radix_sort_pairs_u64, a binary tree by
 * agglomerative clustering of the Morton-ordered triangles (PLOC, Meister
 * and Bittner 2018, neighbour radius 8), bottom-up boxes and SAH cost rows
 * (Ylitie, Karras and Laine 2017), the SAH-optimal collapse to compressed
 * 8-wide nodes (format above; leaves of at most 2 triangles) and the
 * triangle records in leaf order. Replaces optixAccelComputeMemoryUsage +
 * optixAccelBuild (scene.cpp:493-550). Synchronizes `stream` (reads level
 * and cluster counts back to the host). */
int wrk_bvh_build(const float* vertices, const uint32_t* indices, uint32_t numTris,
                  void* nodes, void* tris, void* scratch, uint64_t scratchBytes,
                  wrk_bvh* out, void* stream);

/* ---------------------------------------------------------- traversal --- */
/* Closest hit for rays queue[i] (or i if queue is NULL), i < n (or
 * i < *countDev when countDev is non-NULL, n then being an upper bound).
 * Replaces optixTrace in the Trace stage (trace.cu:127-151) together with
 * __closesthit__/__miss__trace's hit/miss decision. Hits are t in
 * [tMin, tMax] of the watertight test; ties resolve to the smaller primitive
 * id. hits[ray] = {t bits, prim}; prim = 0xFFFFFFFF on a miss. */
int wrk_trace_closest(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                      uint32_t n, const uint32_t* countDev, uint32_t* hits, void* stream);
/* Any hit in [tMin, tMax]: occluded[ray] = 1/0. Replaces the Shadow-stage
 * optixTrace + __closesthit__trace payload (trace.cu:27-51,162-165). */
int wrk_trace_any(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                  uint32_t n, const uint32_t* countDev, uint8_t* occluded, void* stream);

/* Shade categories in the triangle records: for every leaf-ordered triangle,
 * 1 + its category (wrk_classify's material-kind mapping of its material id)
 * is written into the w word of its second vertex (0 = untagged, as
 * wrk_bvh_build leaves it). Run again after the materials change. No
 * counterpart in libWR: it lets the closest-hit launch hand over the shade
 * category so the classify stage skips the material lookups of hits. */
int wrk_bvh_tag_categories(const wrk_bvh* bvh, const wrk_scene_view* scene, void* stream);
/* Closest hit as wrk_trace_closest, plus categories[i] for queue position i:
 * the tagged category of the hit triangle, or 0xFF for a miss or an untagged
 * triangle. fetches may be NULL (see the counted variants below). */
int wrk_trace_closest_categorized(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                                  uint32_t n, const uint32_t* countDev, uint32_t* hits, uint8_t* categories,
                                  unsigned long long* fetches, void* stream);

/* The same traversals built with fetch counters: fetches[0] += inner nodes
 * visited, fetches[1] += triangles tested, fetches[2] += stack entries moved
 * from the LDS part of the traversal stack to the per-lane scratch spill
 * (device uint64[3]). They give the
 * mean node/triangle fetches per ray that the algorithmic-bytes model of the
 * traversal kernels needs; results are identical to the uncounted kernels. */
int wrk_trace_closest_counted(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                              uint32_t n, const uint32_t* countDev, uint32_t* hits, unsigned long long* fetches,
                              void* stream);
int wrk_trace_any_counted(const wrk_bvh* bvh, const void* org4, const void* dir4, const uint32_t* queue,
                          uint32_t n, const uint32_t* countDev, uint8_t* occluded, unsigned long long* fetches,
                          void* stream);

/* ---------------------------------------------------- sort and scan --- */
/* Building blocks of wrk_bvh_build (inside optixAccelBuild in the
 * reference, scene.cpp:525-550), exported for callers and tests.
 * Stable LSD radix sort of n (key, value) pairs by key bits
 * [beginBit, endBit), 8 bits per pass: per-tile LDS histograms, a scan of
 * the digit x tile counts, and a scatter that ranks each tile's keys with
 * wave ballots and stages them in LDS in digit order. The sorted pairs end
 * in keys/vals (equal to std::stable_sort by key); keysTmp/valsTmp are
 * n-entry ping-pong buffers. */
uint64_t wrk_radix_sort_scratch_bytes(uint32_t n);
int wrk_radix_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n,
                             uint32_t beginBit, uint32_t endBit, void* scratch, uint64_t scratchBytes,
                             void* stream);
/* Exclusive scan: out[i] = in[0] + ... + in[i-1] (in == out allowed);
 * *total (device, may be NULL) = the sum of all n values. */
uint64_t wrk_exclusive_scan_scratch_bytes(uint32_t n);
int wrk_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* scratch,
                           uint64_t scratchBytes, void* stream);

/* -------------------------------------------------------- compaction --- */
/* Scratch bytes for wrk_partition / wrk_compact over n items. */
uint64_t wrk_partition_scratch_bytes(uint32_t n, uint32_t numCategories);
uint64_t wrk_compact_scratch_bytes(uint32_t n);
/* Stable compaction: out = [in[i] (or i) for i in 0..n if keep[i]] in input
 * order, *countOut (device) = its length. Replaces compact.cu:6-17, whose
 * atomic-per-survivor order is nondeterministic; this one is ordered. */
int wrk_compact(const uint8_t* keep, const uint32_t* in, uint32_t n, uint32_t* out,
                uint32_t* countOut, void* scratch, uint64_t scratchBytes, void* stream);
/* Stable K-way partition (K <= 16): items with category k, k < numCategories,
 * land in out[offsets[k] ...] in input order; categories >= numCategories
 * are dropped. offsetsOut (device) receives K+1 exclusive offsets. */
int wrk_partition(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t numCategories,
                  uint32_t* out, uint32_t* offsetsOut, void* scratch, uint64_t scratchBytes,
                  void* stream);
/* wrk_partition without a scratch argument, allocating nothing: no
 * counterpart in libWR. For callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that must not call hipMalloc,
 * harnesses that forbid it). The count matrix lives in a 4,194,816-byte
 * device array of the library, one per GPU, held until the process ends;
 * n may be at most wrk_partition_noalloc_max_items() = 2^26 (larger n is
 * rejected with WRK_ERR_INVALID) and numCategories at most 16. Results equal
 * wrk_partition's. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two threads) race and must be serialised
 * by the caller. */
int wrk_partition_noalloc(const uint8_t* category, const uint32_t* in, uint32_t n, uint32_t numCategories,
                          uint32_t* out, uint32_t* offsetsOut, void* stream);
uint32_t wrk_partition_noalloc_max_items(void);

/* ---------------------------------------------------------- wavefront --- */
/* Path state flags (uint4 state.z): depth bits 0..15, isFirstHit 16,
 * prevWasDelta 17, neeDone 18, host-loop iterations 19..23, request slot
 * 24..31. A slot indexes a device table of wrk_render_params, so paths of
 * different requests (resolution, spp, depth, iteration cap) share launches. */
#define WRK_MAX_SLOTS 256

/* Camera rays for paths k < n of one chunk: global path g = first + k of the
 * request (pixel = g % (width*height), sample = g / (width*height)),
 * stored at pool position poolBase + k, tagged with `slot`
 * (trace.cu:53-125). queue[k] = poolBase + k; if category is non-NULL,
 * category[k] = categoryValue. Writes org, dir, rad (zero), state (and the
 * guides, zero, when allocated); thr and result are left as they are: a
 * path whose state still holds isFirstHit has throughput 1 and pdf 1 in
 * every kernel that reads it, and each path writes its result where it
 * ends, unflushed ends (iteration cap, out-of-range material id) as zero. */
int wrk_generate(const wrk_camera* cam, const wrk_render_params* params, const wrk_paths* paths, uint64_t first,
                 uint32_t n, uint32_t poolBase, uint32_t slot, uint32_t* queue, uint8_t* category,
                 uint32_t categoryValue, void* stream);
/* The traversals with the per-entry decisions of wrk_classify taken where
 * each ray ends, so no classify pass runs (the engine's path). Run
 * wrk_trace_any_resolve before wrk_trace_closest_resolve of the same
 * iteration and queue part, so a trace + shadow entry adds its shadow
 * contribution before a miss adds the environment (trace.cu:27-51 before
 * trace.cu:143-148, as in libWR's stage order).
 * Any hit over queue entries [0, numTraceShadow + numShadow) of the shadow
 * rays in paths->savedOrg / savedDir: an unoccluded ray adds its pending
 * contribution to the path's radiance; for entries i >= numTraceShadow
 * (shadow only) categories[i] = 0 (the path resumes as Lambertian, with the
 * reference-mode iteration cap) or 5 (flushed: result = clamped radiance).
 * Closest hit over queue entries [0, n) of the rays in paths->org / dir: a
 * hit writes paths->hit and categories[i] = the triangle's shade category
 * (tagged records, else from the scene's material table); a miss adds
 * throughput x environment radiance, flushes and writes categories[i] = 5.
 * paramsTable is the device table of per-slot render parameters. */
int wrk_trace_any_resolve(const wrk_bvh* bvh, const wrk_render_params* paramsTable, const wrk_paths* paths,
                          const uint32_t* queue, uint32_t numTraceShadow, uint32_t numShadow, uint8_t* categories,
                          unsigned long long* fetches, void* stream);
int wrk_trace_closest_resolve(const wrk_bvh* bvh, const wrk_scene_view* scene, const wrk_render_params* paramsTable,
                              const wrk_paths* paths, const uint32_t* queue, uint32_t n, uint8_t* categories,
                              unsigned long long* fetches, void* stream);
/* Resolves the trace results of the active queue of one scene,
 * [trace numTrace | trace + shadow numTraceShadow | shadow numShadow]
 * (closest-hit results for the first two parts, any-hit results of the
 * shadow ray held in savedOrg/savedDir for the last two), and assigns each
 * path a shade category 0..4 (material kinds of gpu_types.h:11-17) or
 * 5 = finished (trace.cu:27-51 shadow resolve, trace.cu:143-148 miss flush).
 * paramsTable is a device array indexed by the path's slot.
 * flags bit0: category[i] of the first two parts already holds the category
 * of a tagged hit (wrk_trace_closest_categorized), 0xFF otherwise; those
 * entries are taken as they are. */
int wrk_classify(const wrk_scene_view* scene, const wrk_render_params* paramsTable, const wrk_paths* paths,
                 const uint32_t* queue, uint32_t numTrace, uint32_t numTraceShadow, uint32_t numShadow,
                 uint8_t* category, uint32_t flags, void* stream);
/* Runs the material shaders over one scene's queue part in queue order
 * (shade.cu:24-85 and the materials/ headers; libWR's shade kernel also
 * takes every active ray in one launch and switches per ray). queueIn[i] is
 * a pool entry and category[i] its shade category from the resolving
 * traversals: a material kind 0..4 (gpu_types.h:11-17) or 5 = path ended.
 * category[i] becomes the entry's next category: categoryBase + 0 (trace
 * the continuation), + 1 (trace the continuation and a shadow ray), + 2
 * (trace a shadow ray only) or finishedCategory (path ended); queueOut[i] =
 * queueIn[i], for i < n. Within each run of 256 entries the shaders run in
 * material order, so a wave mostly runs one shader.
 * With iterationCap = 0 a Lambertian hit emits its shadow ray and its
 * continuation together; with a cap the continuation waits for the shadow
 * result as in libWR's stage machine, and paths reaching the cap are
 * dropped unflushed (renderer.cpp:213). */
int wrk_shade(const wrk_scene_view* scene, const wrk_render_params* paramsTable, const wrk_paths* paths,
              const uint32_t* queueIn, uint32_t* queueOut, uint32_t n, uint32_t categoryBase,
              uint32_t finishedCategory, uint8_t* category, void* stream);
/* Renders paths firstPath .. firstPath + n - 1 of a request (global path
 * index = sample * width * height + pixel, as wrk_generate) into pool
 * entries poolBase .. poolBase + n - 1 to completion in one launch, one path
 * per thread: the camera ray (trace.cu:53-125), then per visit the stages of
 * one wavefront iteration in their order (wrk_trace_any_resolve,
 * wrk_trace_closest_resolve, wrk_shade: libWR's Shadow -> Trace -> Shade
 * stage machine, trace.cu:27-151 and shade.cu:24-85). paths->result then
 * holds each path's flushed radiance exactly as the wavefront iteration
 * leaves it; accumulate it with wrk_accumulate. paramsTable[slot] (device)
 * holds the request's parameters. No counterpart in libWR, whose renderer
 * runs every request as bounce launches (renderer.cpp:185-240): for short
 * requests, where launches and host round trips per bounce cost more than
 * the work. */
int wrk_render_fused(const wrk_bvh* bvh, const wrk_scene_view* scene, const wrk_camera* camera,
                     const wrk_render_params* paramsTable, uint32_t slot, const wrk_paths* paths,
                     uint64_t firstPath, uint32_t n, uint32_t poolBase, void* stream);
/* Finished paths per slot (no counterpart in libWR, whose host loop renders
 * one batch at a time): live[slot] -= number of the first
 * n = min(*countDev, maxItems) entries of `finished` (pool entries of paths
 * that ended this iteration; n = maxItems when countDev is NULL) whose
 * state carries that slot. */
int wrk_slot_finish(const uint32_t* finished, uint32_t maxItems, const uint32_t* countDev, const wrk_paths* paths,
                    uint32_t* live, void* stream);
/* The same count with the slot of pool entry p taken from slotMap[p / 256]
 * (numGranules bytes, one per 256 pool entries; chunks start and end on
 * such granules), so the scattered state words are not read. maxItems may
 * be at most 256 * numGranules; entries p >= 256 * numGranules are skipped
 * (wrk_slot_finish likewise skips entries past paths->capacity). */
int wrk_slot_finish_map(const uint32_t* finished, uint32_t maxItems, const uint32_t* countDev,
                        const uint8_t* slotMap, uint32_t numGranules, uint32_t* live, void* stream);
/* accum[pixel] += result[poolBase + k] for the chunk's paths in sample order
 * (renderer.cpp:190 serial samples; per-pixel accumulation order is fixed).
 * With paths->guides and non-NULL accumAlbedo/accumNormal the first-hit
 * guides are summed the same way (trace.cu:267-286: first-hit albedo and
 * ray-facing geometric normal, summed over samples; paths that miss or end
 * before their first shade add zero). accumNoise (2 floats per pixel, NULL
 * skips) sums y and y^2 over the pixel's samples, y = (r + g + b) / 3 of each
 * path's result, for the learned denoiser's noise input; no counterpart in
 * libWR, whose OptiX denoiser takes colour, albedo and normal only. */
int wrk_accumulate(const wrk_paths* paths, uint32_t poolBase, uint64_t first, uint32_t n, uint32_t npix,
                   float* accum, float* accumAlbedo, float* accumNormal, float* accumNoise, void* stream);
/* out = accum * (1/spp) (renderer.cpp:308-313). */
int wrk_average(const float* accum, uint32_t npix, uint32_t spp, float* out, void* stream);
/* out[pixel] = standard error of the pixel's mean y from accumNoise moments
 * over spp samples: sqrt(max(s2 - s1^2 / spp, 0) / (spp - 1) / spp), or
 * |s1| when spp = 1. No counterpart in libWR (see accumNoise). */
int wrk_noise_stderr(const float* moments, uint32_t npix, uint32_t spp, float* out, void* stream);
/* Exposure 2.2, ACES (Narkowicz), gamma 1/2.2, truncation to uint8
 * (libWR/test/helpers/image.cpp:14-49). */
int wrk_tonemap_aces(const float* image, uint32_t npix, uint8_t* rgb8, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WRK_H */
