#pragma once
// libWR/include/wr/renderer.h:1-67 — same structs, defaults and overloads.
// Additive: WavefrontConfig (iteration cap, paths in flight) and
// Renderer::renderPaths / getStats for tests and benchmarks.
#include <cstdint>

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
};

// Wavefront scheduling. iterationCap = 8 reproduces libWR's host loop
// (renderer.cpp:213: at most 8 trace+shade iterations per sample; paths
// still alive are dropped). iterationCap = 0 runs every path to completion
// (obeys maxBounces). maxPathsInFlight bounds the device path pool.
struct WavefrontConfig {
    uint32_t iterationCap = 8;
    uint32_t maxPathsInFlight = 1u << 26;  // ~11 GB of path state
    bool timeKernels = false;     // per-kernel device time with HIP events
    bool countTraversal = false;  // node/triangle fetch counters (slower)
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
    // with WavefrontConfig::countTraversal
    uint64_t closestNodeFetches = 0, closestTriFetches = 0, anyNodeFetches = 0, anyTriFetches = 0;
};

struct RendererImpl;

class Renderer {
public:
    Renderer(const RendererConfig& config = {});
    ~Renderer();

    void render(Scene* scene, const Camera& camera, Vec3* outputBuffer, const RenderParams& params);

    void render(Scene* scene, const Camera& camera, Vec3* outputBuffer, uint32_t width, uint32_t height,
                uint32_t spp = 1, bool denoiser = false);

    // --- additive API
    void setWavefrontConfig(const WavefrontConfig& cfg);
    WavefrontConfig getWavefrontConfig() const;
    // Per-path values added to the accumulation buffer for global path
    // indices [firstPath, firstPath + count), path = sample * W*H + pixel;
    // out holds 4 floats per path (rgb, unused).
    void renderPaths(Scene* scene, const Camera& camera, const RenderParams& params, uint64_t firstPath,
                     uint32_t count, float* out);
    const RenderStats& getStats() const;

private:
    RendererImpl* m_impl;
    RendererConfig m_config;
};

}  // namespace wr
