#pragma once
// libWR/include/wr/renderer.h:1-67 — same structs, defaults and overloads.
// Additive: WavefrontConfig (iteration cap, paths in flight) and
// Renderer::renderPaths / getStats for tests and benchmarks.
#include <cstdint>
#include <string>
#include <vector>

#include "wr/scene.h"
#include "wr/types.h"

namespace wr {

struct RendererConfig {
    uint32_t maxBounces = 8;
    bool useNEE = true;
    float russianRouletteDepth = 3;
};

struct DenoiserConfig {
    bool enabled = false;
    bool useAlbedo = false;
    bool useNormal = false;
    float hdrIntensity = 1.0f;
};

struct RenderParams {
    uint32_t width;
    uint32_t height;
    uint32_t spp = 1;
    DenoiserConfig denoiser;

    bool useNEE = true;
    uint32_t maxBounces = 8;
    float russianRouletteDepth = 3.0f;
    float fireflyClamp = 0.0f;
    // additive: render samples [firstSample, firstSample + spp) (libWR always
    // starts at sample 0); images of disjoint ranges are independent
    uint32_t firstSample = 0;
};

// Wavefront scheduling. iterationCap = 8 reproduces libWR's host loop
// (renderer.cpp:213: at most 8 trace+shade iterations per sample; paths
// still alive are dropped). iterationCap = 0 runs every path to completion
// (obeys maxBounces). maxPathsInFlight bounds the device path pool.
struct WavefrontConfig {
    uint32_t iterationCap = 8;
    uint32_t maxPathsInFlight = 1u << 26;  // ~11 GB of path state
    // paths per admitted chunk (0 = maxPathsInFlight); smaller chunks refill
    // the pool while the deep paths of earlier chunks finish
    uint32_t maxChunkPaths = 0;
    bool timeKernels = false;     // per-kernel device time with HIP events
    bool countTraversal = false;  // node/triangle fetch counters (slower)
    // requests of at most this many paths (width * height * spp) are
    // rendered by one fused launch per chunk, each path to its end on one
    // thread, with the same per-path results; 0 = bounce iterations only
    uint64_t fusedMaxPaths = 0;
};

struct RenderStats {
    double seconds = 0;            // wall time of the last render (device work, synchronized)
    uint64_t paths = 0;            // width * height * spp
    uint64_t closestRays = 0;      // closest-hit queries issued
    uint64_t shadowRays = 0;       // any-hit queries issued
    uint32_t iterations = 0;       // wavefront iterations summed over batches
    uint32_t batches = 0;
    // with WavefrontConfig::timeKernels: summed device milliseconds per kernel
    double msGenerate = 0, msTraceClosest = 0, msTraceAny = 0, msClassify = 0, msPartition = 0, msShade = 0,
           msAccumulate = 0;
    uint32_t launchesTraceClosest = 0, launchesTraceAny = 0;
    // with WavefrontConfig::countTraversal (stack spills: traversal stack
    // entries moved from LDS to the per-lane scratch area)
    uint64_t closestNodeFetches = 0, closestTriFetches = 0, anyNodeFetches = 0, anyTriFetches = 0;
    uint64_t closestStackSpills = 0, anyStackSpills = 0;
    // with WavefrontConfig::timeKernels: items the classify, partition and
    // shade launches processed (queue entries; for shade the entries of the
    // five material queues). The wavefront engine resolves trace results in
    // the traversal kernels, so msClassify and classifyItems stay 0 there.
    uint64_t classifyItems = 0, partitionItems = 0, shadeItems = 0;
    // paths rendered by the fused one-launch kernel (WavefrontConfig::fusedMaxPaths)
    // and, with timeKernels, its device milliseconds
    uint64_t fusedPaths = 0;
    double msFused = 0;
};

struct RendererImpl;
struct ServerImpl;

class Renderer {
public:
    Renderer(const RendererConfig& config = {});
    ~Renderer();

    void render(Scene* scene, const Camera& camera, Vec3* outputBuffer, const RenderParams& params);

    void render(Scene* scene, const Camera& camera, Vec3* outputBuffer, uint32_t width, uint32_t height,
                uint32_t spp = 1, bool denoiser = false);

    // --- additive API
    void setWavefrontConfig(const WavefrontConfig& cfg);
    // Learned denoiser weights (FP32 parameter file written by
    // offlinerenderer_amd.denoiser.UNet.save): render() then denoises when
    // params.denoiser.enabled (in place of libWR's OptiX AI denoiser).
    // Without a call, the weights shipped in offlinerenderer_amd/models are
    // used when present.
    void loadDenoiser(const std::string& path);
    WavefrontConfig getWavefrontConfig() const;
    // Per-path values added to the accumulation buffer for global path
    // indices [firstPath, firstPath + count), path = sample * W*H + pixel;
    // out holds 4 floats per path (rgb, unused).
    void renderPaths(Scene* scene, const Camera& camera, const RenderParams& params, uint64_t firstPath,
                     uint32_t count, float* out);
    // The averaged image plus the denoiser guide layers of libWR
    // (trace.cu:267-286): first-hit albedo and ray-facing geometric normal,
    // averaged over spp; each buffer holds width*height Vec3. noise (optional,
    // width*height floats): the standard error of each pixel's mean
    // (r + g + b) / 3 over its samples, the learned denoiser's noise input
    // (no counterpart in libWR).
    void renderWithGuides(Scene* scene, const Camera& camera, const RenderParams& params, Vec3* color,
                          Vec3* albedo, Vec3* normal, float* noise = nullptr);
    const RenderStats& getStats() const;

private:
    RendererImpl* m_impl;
    RendererConfig m_config;
};

// --- additive: continuous-batching render server -----------------------------
// Requests of different scenes (up to 4 resident), resolutions and spp share
// one wavefront: each iteration traces and shades every in-flight path of
// every admitted request, and new requests are admitted (in chunks of at
// most maxChunkPaths paths) as soon as path slots and the memory budget
// allow. A request's image does not depend on what else is in flight.
struct ServerConfig {
    uint32_t maxPathsInFlight = 1u << 26;  // path pool
    uint32_t maxChunkPaths = 1u << 24;     // admission unit (paths per request chunk)
    uint64_t memoryBudgetBytes = 0;        // accumulation + image buffers of admitted requests; 0 = unlimited
    uint32_t iterationCap = 0;             // as WavefrontConfig::iterationCap
    bool timeKernels = false;
    // admission order of pending requests: false = arrival order, true =
    // fewest remaining paths first (short requests pass long ones)
    bool shortestFirst = false;
    // false: the first pending request takes every chunk that fits before
    // the next one gets any (requests finish in order, as served one after
    // another, while the pool stays full); true: one chunk per pending
    // request per round (all requests progress together)
    bool roundRobin = false;
    // express lane: requests of at most expressMaxPaths paths (0 = off) run
    // in a second pool of expressPathsInFlight paths, so short requests do
    // not wait out the long ones' bounces. Each step() runs up to
    // expressIterations of its iterations before one iteration of the main
    // pool, or, with expressOverlap, while that iteration is on the device
    uint32_t expressMaxPaths = 0;
    uint32_t expressPathsInFlight = 1u << 23;
    uint32_t expressIterations = 32;
    bool expressOverlap = false;
    // requests of at most this many paths render in one fused launch per
    // chunk (as WavefrontConfig::fusedMaxPaths) in the lane they take
    uint64_t fusedMaxPaths = 0;
    // in-order admission: while the oldest unfinished request has paths in
    // flight, the requests behind it share at most backfillPaths of the pool
    // (0 = no limit), so they fill its draining tail without stretching its
    // iterations; fused requests are exempt
    uint64_t backfillPaths = 0;
};

struct RequestStatus {
    bool done = false;
    bool admitted = false;
    double submitTime = 0, admitTime = 0, doneTime = 0;  // RenderServer::clock() seconds
    uint64_t paths = 0;
};

class RenderServer {
public:
    explicit RenderServer(const ServerConfig& config = {});
    ~RenderServer();
    RenderServer(const RenderServer&) = delete;
    RenderServer& operator=(const RenderServer&) = delete;

    uint64_t submit(Scene* scene, const Camera& camera, const RenderParams& params);
    // Admission, one wavefront iteration and completions; true while work remains.
    bool step();
    bool isDone(uint64_t id) const;
    RequestStatus status(uint64_t id) const;
    // Up to maxIds ids of requests that finished since they were last taken,
    // in completion order; returns how many were written.
    uint32_t takeCompleted(uint64_t* ids, uint32_t maxIds);
    // Averaged image (width*height Vec3) of a finished request; releases it.
    void fetch(uint64_t id, Vec3* out);
    uint32_t numPending() const;
    uint32_t numInFlight() const;
    const RenderStats& getStats() const;
    static double clock();

private:
    ServerImpl* m_impl;
};

}  // namespace wr
