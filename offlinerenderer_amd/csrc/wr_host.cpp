// wr_host.cpp — the wr:: C++ API (Context, Scene, Texture2D, Renderer) over
// the wrk_* C ABI. Mirrors libWR/src/internal/{context,scene,renderer}.cpp:
// same method semantics, exception types and messages; the OptiX/CUDA
// machinery is replaced by HIP device buffers, the GPU LBVH and the
// wavefront kernels.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wr/wr.h"
#include "../../include/wrk.h"
#include "engine.h"
#include "../../include/wrk_unet.h"
#include "image_io.h"

namespace wr {

using detail::DevBuf;
using detail::hip_check;
using detail::wrk_check;

namespace {

int g_logLevel = 4;
bool g_contextReady = false;
int g_deviceId = 0;

void ensure_context() {
    if (!g_contextReady) throw std::runtime_error("wr::Context must be created before scenes and renderers");
    hip_check(hipSetDevice(g_deviceId), "hipSetDevice");
}

}  // namespace

// ------------------------------------------------------------- Context
// libWR/src/internal/context.cpp:25-58,79-103
Context::Context(const ContextConfig& config) {
    if (!g_contextReady) {
        g_logLevel = config.logLevel;
        g_deviceId = config.deviceId;
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (config.deviceId < 0 || config.deviceId >= count)
            throw std::runtime_error("wr::Context: device " + std::to_string(config.deviceId) + " not present");
        hip_check(hipSetDevice(config.deviceId), "hipSetDevice");
        hip_check(hipFree(nullptr), "hipFree(0)");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, config.deviceId), "hipGetDeviceProperties");
        if (g_logLevel >= 3) std::cout << "[WR] Using device: " << prop.name << " (" << prop.gcnArchName << ")" << std::endl;
        g_contextReady = true;
    }
}
Context::~Context() {}
Scene* Context::createScene() { return new Scene(); }
Renderer* Context::createRenderer() { return new Renderer(); }
int Context::getDeviceId() const { return g_contextReady ? g_deviceId : 0; }

// ---------------------------------------------------------------- Scene
struct TextureData {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> pixels;
    DevBuf dev;
    bool destroyed = false;
};

struct SceneImpl {
    std::vector<float> vertices, normals, uvs;
    std::vector<uint32_t> indices, triangleMaterialIds;
    std::vector<wrk_material> materials;
    std::vector<TextureData> textures;
    bool hasNormals = false, hasUVs = false;
    uint32_t numTriangles = 0, numVertices = 0;
    Vec3 environmentRadiance{0.0f, 0.0f, 0.0f};
    bool finalized = false;

    std::vector<uint32_t> emissiveTriangles;
    std::vector<float> emissiveTriangleCDF;

    wrhost::ImageF envMap;
    DevBuf d_envMap;

    DevBuf d_vertices, d_normals, d_uvs, d_indices, d_triMat, d_materials, d_textures, d_emissive, d_cdf, d_guide;
    uint32_t guideSize = 0;
    DevBuf d_nodes, d_tris, d_bvhRecord;
    wrk_bvh bvh{};
    wrk_scene_view view{};
    double buildSeconds = 0;

    uint32_t pushMaterial(const wrk_material& m) {
        materials.push_back(m);
        return static_cast<uint32_t>(materials.size() - 1);
    }
    void uploadTextures() {
        std::vector<wrk_texture> desc(textures.size());
        for (size_t i = 0; i < textures.size(); ++i) {
            TextureData& t = textures[i];
            if (!t.destroyed && !t.pixels.empty()) {
                t.dev.upload(t.pixels.data(), t.pixels.size());
                desc[i].texels = t.dev.addr();
                desc[i].width = t.width;
                desc[i].height = t.height;
            } else {
                desc[i].texels = 0; desc[i].width = 0; desc[i].height = 0;
            }
        }
        if (!desc.empty()) d_textures.upload(desc.data(), desc.size() * sizeof(wrk_texture));
        view.textures = desc.empty() ? 0 : d_textures.addr();
        view.numTextures = (uint32_t)desc.size();
    }
};

static wrk_material make_material(const Vec3& albedo, const Vec3& emission, float ior, uint32_t type, float rough,
                                  float metal, uint32_t albedoTex) {
    wrk_material m;
    std::memset(&m, 0, sizeof(m));
    m.albedo[0] = albedo.x; m.albedo[1] = albedo.y; m.albedo[2] = albedo.z; m.albedo[3] = 0.0f;
    m.emission[0] = emission.x; m.emission[1] = emission.y; m.emission[2] = emission.z; m.emission[3] = 0.0f;
    m.ior = ior;
    m.type = type;
    m.roughness = rough;
    m.metallic = metal;
    m.albedoTextureId = albedoTex;
    m.normalTextureId = 0;
    return m;
}

Scene::Scene() : m_impl(new SceneImpl()) {}
Scene::~Scene() { delete m_impl; }

// scene.cpp:91-168
uint32_t Scene::addLambertianMaterial(const Vec3& albedo) { return addLambertianMaterial(albedo, 0); }
uint32_t Scene::addLambertianMaterial(const Vec3& albedo, uint32_t albedoTextureId) {
    return m_impl->pushMaterial(make_material(albedo, Vec3(0.0f), 1.0f, WRK_MAT_LAMBERTIAN, 1.0f, 0.0f, albedoTextureId));
}
uint32_t Scene::addEmissiveMaterial(const Vec3& emission) {
    return m_impl->pushMaterial(make_material(Vec3(0.0f), emission, 1.0f, WRK_MAT_EMISSIVE, 0.0f, 0.0f, 0));
}
uint32_t Scene::addGlassMaterial(const Vec3& albedo, float ior) {
    return m_impl->pushMaterial(make_material(albedo, Vec3(0.0f), ior, WRK_MAT_GLASS, 0.0f, 0.0f, 0));
}
uint32_t Scene::addGGXReflectionMaterial(const Vec3& albedo, float roughness, float metallic) {
    return m_impl->pushMaterial(make_material(albedo, Vec3(0.0f), 1.5f, WRK_MAT_GGX_REFLECTION, roughness, metallic, 0));
}
uint32_t Scene::addGGXTransmissionMaterial(const Vec3& albedo, float roughness, float ior) {
    return m_impl->pushMaterial(make_material(albedo, Vec3(0.0f), ior, WRK_MAT_GGX_TRANSMISSION, roughness, 0.0f, 0));
}
void Scene::setMaterialNormalMap(uint32_t materialId, uint32_t normalTextureId) {
    if (materialId >= m_impl->materials.size()) throw std::runtime_error("Invalid material ID");
    m_impl->materials[materialId].normalTextureId = normalTextureId;
}

// scene.cpp:176-244
void Scene::addTriangleMesh(std::span<const float> verts, std::span<const uint32_t> inds, uint32_t materialId) {
    addTriangleMesh(verts, inds, std::span<const float>(), materialId);
}
void Scene::addTriangleMesh(std::span<const float> verts, std::span<const uint32_t> inds, std::span<const float> uvData,
                            uint32_t materialId) {
    addTriangleMesh(verts, inds, std::span<const float>(), uvData, materialId);
}
void Scene::addTriangleMesh(std::span<const float> verts, std::span<const uint32_t> inds,
                            std::span<const float> normalData, std::span<const float> uvData, uint32_t materialId) {
    SceneImpl& s = *m_impl;
    uint32_t vertexOffset = static_cast<uint32_t>(s.vertices.size() / 3);
    s.vertices.insert(s.vertices.end(), verts.begin(), verts.end());
    size_t numNewVerts = verts.size() / 3;
    if (!normalData.empty() && normalData.size() >= numNewVerts * 3) {
        s.hasNormals = true;
        s.normals.insert(s.normals.end(), normalData.begin(), normalData.end());
    } else {
        s.normals.insert(s.normals.end(), numNewVerts * 3, 0.0f);
    }
    if (!uvData.empty() && uvData.size() >= numNewVerts * 2) {
        s.hasUVs = true;
        s.uvs.insert(s.uvs.end(), uvData.begin(), uvData.begin() + numNewVerts * 2);
    } else {
        s.uvs.insert(s.uvs.end(), numNewVerts * 2, 0.0f);
    }
    for (uint32_t idx : inds) s.indices.push_back(vertexOffset + idx);
    uint32_t numTris = static_cast<uint32_t>(inds.size() / 3);
    s.triangleMaterialIds.insert(s.triangleMaterialIds.end(), numTris, materialId);
    s.numTriangles += numTris;
    s.numVertices = static_cast<uint32_t>(s.vertices.size() / 3);
}

// scene.cpp:246-276
uint32_t Scene::loadTexture(const std::string& path) {
    wrhost::Image8 img;
    if (!wrhost::load_png_rgba8(path, img)) throw std::runtime_error("Failed to load texture: " + path);
    uint32_t id = addTextureRGBA8(img.width, img.height, std::span<const uint8_t>(img.rgba.data(), img.rgba.size()));
    if (g_logLevel >= 4)
        std::cout << "[Scene] Loaded texture: " << path << " (" << img.width << "x" << img.height << "), id=" << id << std::endl;
    return id;
}
uint32_t Scene::addTextureRGBA8(uint32_t width, uint32_t height, std::span<const uint8_t> rgba) {
    if (width == 0 || height == 0 || rgba.size() < (size_t)width * height * 4)
        throw std::runtime_error("addTextureRGBA8: texel buffer smaller than width*height*4");
    TextureData t;
    t.width = width;
    t.height = height;
    t.pixels.assign(rgba.begin(), rgba.begin() + (size_t)width * height * 4);
    m_impl->textures.push_back(std::move(t));
    uint32_t id = static_cast<uint32_t>(m_impl->textures.size());
    if (m_impl->finalized) m_impl->uploadTextures();
    return id;
}
void Scene::destroyTexture(uint32_t textureId) {
    if (textureId == 0 || textureId > m_impl->textures.size()) return;
    TextureData& t = m_impl->textures[textureId - 1];
    t.destroyed = true;
    t.pixels.clear();
    if (m_impl->finalized) m_impl->uploadTextures();
    t.dev.release();
}
uint32_t Texture2D::create(Scene* scene, const std::string& path) { return scene->loadTexture(path); }
void Texture2D::destroy(Scene* scene, uint32_t textureId) { scene->destroyTexture(textureId); }

void Scene::setEnvironmentRadiance(const Vec3& radiance) {
    m_impl->environmentRadiance = radiance;
    m_impl->view.envRadiance[0] = radiance.x;
    m_impl->view.envRadiance[1] = radiance.y;
    m_impl->view.envRadiance[2] = radiance.z;
}
// scene.cpp:289-348. As in libWR the map is loaded and kept on the device
// but the kernels shade misses with the uniform environmentRadiance.
void Scene::setEnvironmentMap(const std::string& hdrPath) {
    wrhost::ImageF img;
    if (!wrhost::load_hdr_rgba32f(hdrPath, img)) throw std::runtime_error("Failed to load HDR environment map: " + hdrPath);
    ensure_context();
    m_impl->envMap = std::move(img);
    m_impl->d_envMap.upload(m_impl->envMap.rgba.data(), m_impl->envMap.rgba.size() * sizeof(float));
}

// scene.cpp:350-557
void Scene::finalize(const SceneBuildConfig& config) {
    (void)config;
    SceneImpl& s = *m_impl;
    if (s.finalized) return;
    if (g_logLevel >= 4)
        std::cout << "[Scene] Finalizing scene: " << s.numTriangles << " triangles, " << s.numVertices << " vertices"
                  << std::endl;
    if (s.vertices.empty() || s.indices.empty()) throw std::runtime_error("Scene has no geometry");
    ensure_context();
    s.d_vertices.upload(s.vertices.data(), s.vertices.size() * sizeof(float));
    s.d_indices.upload(s.indices.data(), s.indices.size() * sizeof(uint32_t));
    s.d_triMat.upload(s.triangleMaterialIds.data(), s.triangleMaterialIds.size() * sizeof(uint32_t));
    if (!s.materials.empty()) s.d_materials.upload(s.materials.data(), s.materials.size() * sizeof(wrk_material));
    if (s.hasNormals) s.d_normals.upload(s.normals.data(), s.normals.size() * sizeof(float));
    if (s.hasUVs) s.d_uvs.upload(s.uvs.data(), s.uvs.size() * sizeof(float));

    // Emissive list and power CDF: sequential FP32 running sum (scene.cpp:423-491).
    s.emissiveTriangles.clear();
    s.emissiveTriangleCDF.clear();
    float totalPower = 0.0f;
    for (uint32_t tri = 0; tri < s.numTriangles; ++tri) {
        uint32_t matId = s.triangleMaterialIds[tri];
        if (matId >= s.materials.size()) continue;
        const wrk_material& mat = s.materials[matId];
        if (mat.type != WRK_MAT_EMISSIVE) continue;
        s.emissiveTriangles.push_back(tri);
        const float* v0 = &s.vertices[s.indices[tri * 3 + 0] * 3];
        const float* v1 = &s.vertices[s.indices[tri * 3 + 1] * 3];
        const float* v2 = &s.vertices[s.indices[tri * 3 + 2] * 3];
        float e1x = v1[0] - v0[0], e1y = v1[1] - v0[1], e1z = v1[2] - v0[2];
        float e2x = v2[0] - v0[0], e2y = v2[1] - v0[1], e2z = v2[2] - v0[2];
        float cx = e1y * e2z - e1z * e2y;
        float cy = e1z * e2x - e1x * e2z;
        float cz = e1x * e2y - e1y * e2x;
        float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
        float power = (mat.emission[0] + mat.emission[1] + mat.emission[2]) * area;
        totalPower += power;
        s.emissiveTriangleCDF.push_back(totalPower);
    }
    if (!s.emissiveTriangleCDF.empty()) {
        for (float& c : s.emissiveTriangleCDF) c /= totalPower;
        s.d_emissive.upload(s.emissiveTriangles.data(), s.emissiveTriangles.size() * sizeof(uint32_t));
        s.d_cdf.upload(s.emissiveTriangleCDF.data(), s.emissiveTriangleCDF.size() * sizeof(float));
        // search guide: guide[k] = first i with cdf[i] >= k / G (G a power of
        // two >= the light count, 256..2^20), same float comparison as the
        // device search, so each sample searches about one entry
        uint32_t G = 256;
        while (G < s.emissiveTriangleCDF.size() && G < (1u << 20)) G <<= 1;
        std::vector<uint32_t> guide(G + 1);
        size_t i = 0;
        for (uint32_t k = 0; k <= G; ++k) {
            const float x = (float)k / (float)G;
            while (i < s.emissiveTriangleCDF.size() && s.emissiveTriangleCDF[i] < x) ++i;
            guide[k] = (uint32_t)i;
        }
        s.d_guide.upload(guide.data(), guide.size() * sizeof(uint32_t));
        s.guideSize = G;
        if (g_logLevel >= 4)
            std::cout << "[Scene] Built light list: " << s.emissiveTriangles.size() << " emissive triangles" << std::endl;
    }

    // GPU LBVH (replaces optixAccelBuild)
    uint64_t nodeBytes = 0, triBytes = 0, scratchBytes = 0;
    wrk_check(wrk_bvh_query_sizes(s.numTriangles, &nodeBytes, &triBytes, &scratchBytes), "wrk_bvh_query_sizes");
    s.d_nodes.alloc(nodeBytes);
    s.d_tris.alloc(triBytes);
    DevBuf scratch;
    scratch.alloc(scratchBytes);
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    auto t0 = std::chrono::steady_clock::now();
    wrk_check(wrk_bvh_build(s.d_vertices.as<float>(), s.d_indices.as<uint32_t>(), s.numTriangles, s.d_nodes.ptr,
                            s.d_tris.ptr, scratch.ptr, scratchBytes, &s.bvh, nullptr),
              "wrk_bvh_build");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    s.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s.d_bvhRecord.upload(&s.bvh, sizeof(wrk_bvh));

    wrk_scene_view& v = s.view;
    v.vertices = s.d_vertices.addr();
    v.normals = s.hasNormals ? s.d_normals.addr() : 0;
    v.uvs = s.hasUVs ? s.d_uvs.addr() : 0;
    v.indices = s.d_indices.addr();
    v.triMaterialIds = s.d_triMat.addr();
    v.materials = s.d_materials.addr();
    v.emissiveTris = s.d_emissive.addr();
    v.emissiveCdf = s.d_cdf.addr();
    v.emissiveGuide = s.guideSize ? s.d_guide.addr() : 0;
    v.emissiveGuideSize = s.guideSize;
    v.numTris = s.numTriangles;
    v.numMaterials = (uint32_t)s.materials.size();
    v.numEmissive = (uint32_t)s.emissiveTriangles.size();
    v.envRadiance[0] = s.environmentRadiance.x;
    v.envRadiance[1] = s.environmentRadiance.y;
    v.envRadiance[2] = s.environmentRadiance.z;
    v.envRadiance[3] = 0.0f;
    // shade categories into the triangle records (hits then carry them)
    wrk_check(wrk_bvh_tag_categories(&s.bvh, &v, nullptr), "wrk_bvh_tag_categories");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    s.uploadTextures();
    if (g_logLevel >= 4)
        std::cout << "[Scene] BVH built: " << s.bvh.numNodes << " nodes, depth " << s.bvh.maxDepth << ", "
                  << s.buildSeconds * 1e3 << " ms" << std::endl;
    s.finalized = true;
}

DevicePtr Scene::getVerticesBuffer() const { return m_impl->d_vertices.addr(); }
DevicePtr Scene::getNormalsBuffer() const { return m_impl->hasNormals ? m_impl->d_normals.addr() : 0; }
DevicePtr Scene::getIndicesBuffer() const { return m_impl->d_indices.addr(); }
DevicePtr Scene::getTriangleMaterialIdsBuffer() const { return m_impl->d_triMat.addr(); }
DevicePtr Scene::getMaterialsBuffer() const { return m_impl->d_materials.addr(); }
TraversableHandle Scene::getGASHandle() const { return m_impl->d_bvhRecord.addr(); }
uint32_t Scene::getNumMaterials() const { return static_cast<uint32_t>(m_impl->materials.size()); }
Vec3 Scene::getEnvironmentRadiance() const { return m_impl->environmentRadiance; }
uint64_t Scene::getEnvironmentMapTexture() const { return m_impl->d_envMap.addr(); }
uint32_t Scene::getEnvironmentMapWidth() const { return m_impl->envMap.width; }
uint32_t Scene::getEnvironmentMapHeight() const { return m_impl->envMap.height; }
DevicePtr Scene::getEmissiveTrianglesBuffer() const { return m_impl->d_emissive.addr(); }
DevicePtr Scene::getEmissiveTriangleCDFBuffer() const { return m_impl->d_cdf.addr(); }
uint32_t Scene::getNumEmissiveTriangles() const { return static_cast<uint32_t>(m_impl->emissiveTriangles.size()); }
DevicePtr Scene::getTexturesBuffer() const { return m_impl->d_textures.addr(); }
uint32_t Scene::getNumTextures() const { return static_cast<uint32_t>(m_impl->textures.size()); }
DevicePtr Scene::getUVsBuffer() const { return m_impl->hasUVs ? m_impl->d_uvs.addr() : 0; }
const wrk_bvh& Scene::getBVH() const { return m_impl->bvh; }
const wrk_scene_view& Scene::getDeviceView() const { return m_impl->view; }
uint32_t Scene::getNumTriangles() const { return m_impl->numTriangles; }
double Scene::getBuildSeconds() const { return m_impl->buildSeconds; }

// -------------------------------------------------------------- Renderer
// Weights shipped beside the library (offlinerenderer_amd/models), used
// when render() is asked to denoise and no other network was loaded.
static std::string default_denoiser_path() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&default_denoiser_path), &info) || !info.dli_fname) return {};
    std::string lib(info.dli_fname);
    const size_t slash = lib.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../models/denoiser_unet.f32";
}

struct RendererImpl {
    WavefrontConfig wf;
    RenderStats stats;
    std::unique_ptr<detail::Engine> eng;
    void* unet = nullptr;  // learned denoiser (wrk_unet_*), loaded by loadDenoiser
    bool triedDefault = false;
    DevBuf denoised;

    void load_denoiser(const std::string& path) {
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) throw std::runtime_error("loadDenoiser: cannot open " + path);
        const std::streamsize bytes = f.tellg();
        f.seekg(0);
        // the file's parameter count names the network's output head
        int head = -1;
        for (uint32_t h = 0; h < 2; ++h)
            if ((uint64_t)bytes == wrk_unet_params_for_head(h) * sizeof(float)) head = (int)h;
        if (head < 0)
            throw std::runtime_error("loadDenoiser: " + path + " does not hold the " +
                                     std::to_string(wrk_unet_params_for_head(0)) + " or " +
                                     std::to_string(wrk_unet_params_for_head(1)) + " FP32 parameters of a network");
        if (unet && wrk_unet_head(unet) != (uint32_t)head) {
            wrk_unet_destroy(unet);
            unet = nullptr;
        }
        if (!unet) unet = wrk_unet_create_head((uint32_t)head);
        if (!unet) throw std::runtime_error(std::string("loadDenoiser: ") + wrk_unet_last_error());
        const uint64_t n = wrk_unet_num_params(unet);
        std::vector<float> p(n);
        f.read(reinterpret_cast<char*>(p.data()), bytes);
        if (wrk_unet_set_params(unet, p.data()) != WRK_OK)
            throw std::runtime_error(std::string("loadDenoiser: ") + wrk_unet_last_error());
    }

    RendererImpl() { ensure_context(); }
    ~RendererImpl() {
        if (unet) wrk_unet_destroy(unet);
    }

    wrk_render_params params_of(const RenderParams& p) const {
        wrk_render_params k;
        std::memset(&k, 0, sizeof(k));
        k.width = p.width; k.height = p.height; k.useNEE = p.useNEE ? 1u : 0u; k.maxBounces = p.maxBounces;
        k.rrStartDepth = p.russianRouletteDepth; k.fireflyClamp = p.fireflyClamp; k.iterationCap = wf.iterationCap;
        return k;
    }
    detail::Engine& engine(uint64_t paths) {
        detail::EngineConfig c;
        c.capacity = (uint32_t)std::max<uint64_t>(1024, std::min<uint64_t>(paths, std::max(wf.maxPathsInFlight, 1024u)));
        c.maxChunkPaths = std::max(wf.maxChunkPaths ? std::min(wf.maxChunkPaths, wf.maxPathsInFlight) : wf.maxPathsInFlight,
                                   1024u);
        c.timeKernels = wf.timeKernels;
        c.countTraversal = wf.countTraversal;
        c.fusedMaxPaths = wf.fusedMaxPaths;
        if (!eng) eng = std::make_unique<detail::Engine>(c);
        c.capacity = std::max(c.capacity, eng->config().capacity);
        eng->setConfig(c);
        eng->ensureCapacity(c.capacity);
        eng->stats() = RenderStats{};
        return *eng;
    }
    // renderer.cpp:132-316 — the whole image as one request, in chunks of
    // at most maxPathsInFlight paths, accumulated per pixel in sample order.
    void render(Scene* scene, const Camera& camera, const RenderParams& rp, float* outHost, float* albedo = nullptr,
                float* normal = nullptr, float* noise = nullptr) {
        if (!scene) throw std::runtime_error("render: null scene");
        if (rp.width == 0 || rp.height == 0 || rp.spp == 0) throw std::runtime_error("render: empty image");
        const uint64_t total = (uint64_t)rp.width * rp.height * rp.spp;
        detail::Engine& e = engine(total);
        auto t0 = std::chrono::steady_clock::now();
        // denoiser.cpp:77-195 takes colour + albedo + normal of any size;
        // the learned network runs on sides padded up to multiples of 8
        if (rp.denoiser.enabled && !unet && !triedDefault) {
            triedDefault = true;
            const std::string path = default_denoiser_path();
            if (!path.empty() && std::ifstream(path).good()) load_denoiser(path);
        }
        if (rp.denoiser.enabled && !unet)
            throw std::runtime_error("render: denoiser requested but no denoiser network is loaded "
                                     "(Renderer::loadDenoiser)");
        const bool denoise = rp.denoiser.enabled;
        const bool guides = albedo || normal || denoise;
        const uint64_t first = (uint64_t)rp.firstSample * rp.width * rp.height;
        uint64_t id = e.submit(scene, detail::camera_record(camera), params_of(rp), rp.spp, first, total,
                               detail::Engine::Mode::Accumulate, nullptr, guides);
        while (e.step()) {
        }
        std::vector<float> tmp;
        float* dst = outHost;
        if (!dst) {
            tmp.resize((size_t)rp.width * rp.height * 3);
            dst = tmp.data();
        }
        if (denoise) {
            // sides padded to multiples of 8 by repeating the last row and
            // column; the caller gets the top-left width x height crop
            const uint32_t pw = (rp.width + 7u) & ~7u, ph = (rp.height + 7u) & ~7u;
            const size_t rowBytes = (size_t)rp.width * 3 * sizeof(float);
            denoised.alloc((size_t)pw * ph * 3 * sizeof(float));
            const uint32_t spp = rp.spp;  // the network is told how noisy its input is
            if (wrk_unet_set_sample_counts(unet, &spp, 1) != WRK_OK ||
                wrk_unet_set_noise(unet, e.deviceNoise(id)) != WRK_OK ||
                wrk_unet_forward(unet, e.deviceImage(id), e.deviceAlbedo(id), e.deviceNormal(id), rp.height,
                                 rp.width, nullptr, 1, ph, pw, denoised.as<float>(), e.stream()) != WRK_OK)
                throw std::runtime_error(std::string("denoiser: ") + wrk_unet_last_error());
            hip_check(hipStreamSynchronize(e.stream()), "hipStreamSynchronize");
            hip_check(hipMemcpy2D(dst, rowBytes, denoised.ptr, (size_t)pw * 3 * sizeof(float), rowBytes, rp.height,
                                  hipMemcpyDeviceToHost),
                      "hipMemcpy2D D2H");
            e.fetch(id, nullptr, albedo, normal, noise);
        } else {
            e.fetch(id, dst, albedo, normal, noise);
        }
        stats = e.stats();
        stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats.paths = total;
    }
    void render_paths(Scene* scene, const Camera& camera, const RenderParams& rp, uint64_t firstPath, uint32_t count,
                      float* out) {
        if (!scene) throw std::runtime_error("renderPaths: null scene");
        detail::Engine& e = engine(count);
        uint64_t id = e.submit(scene, detail::camera_record(camera), params_of(rp), rp.spp, firstPath, count,
                               detail::Engine::Mode::Paths, out);
        while (e.step()) {
        }
        e.release(id);
        stats = e.stats();
    }
};

Renderer::Renderer(const RendererConfig& config) : m_impl(nullptr), m_config(config) {
    try {
        m_impl = new RendererImpl();
    } catch (const std::exception& e) {
        std::cerr << "[Renderer] Initialization failed: " << e.what() << std::endl;
        throw;
    }
}
Renderer::~Renderer() { delete m_impl; }

// renderer.cpp:132-316
void Renderer::render(Scene* scene, const Camera& camera, Vec3* outputBuffer, const RenderParams& params) {
    if (g_logLevel >= 4)
        std::cout << "[Renderer] Starting render: " << params.width << "x" << params.height << " @ " << params.spp
                  << " spp" << std::endl;
    m_impl->render(scene, camera, params, reinterpret_cast<float*>(outputBuffer));
    if (g_logLevel >= 4) std::cout << "[Renderer] Render complete" << std::endl;
}

// renderer.cpp:318-335
void Renderer::render(Scene* scene, const Camera& camera, Vec3* outputBuffer, uint32_t width, uint32_t height,
                      uint32_t spp, bool denoiser) {
    RenderParams params;
    params.width = width;
    params.height = height;
    params.spp = spp;
    params.denoiser.enabled = denoiser;
    params.denoiser.useAlbedo = true;
    params.denoiser.useNormal = true;
    render(scene, camera, outputBuffer, params);
}

void Renderer::setWavefrontConfig(const WavefrontConfig& cfg) { m_impl->wf = cfg; }
void Renderer::loadDenoiser(const std::string& path) { m_impl->load_denoiser(path); }
WavefrontConfig Renderer::getWavefrontConfig() const { return m_impl->wf; }
void Renderer::renderPaths(Scene* scene, const Camera& camera, const RenderParams& params, uint64_t firstPath,
                           uint32_t count, float* out) {
    m_impl->render_paths(scene, camera, params, firstPath, count, out);
}
const RenderStats& Renderer::getStats() const { return m_impl->stats; }
void Renderer::renderWithGuides(Scene* scene, const Camera& camera, const RenderParams& params, Vec3* color,
                                Vec3* albedo, Vec3* normal, float* noise) {
    if (!albedo || !normal) throw std::runtime_error("renderWithGuides: null guide buffer");
    m_impl->render(scene, camera, params, reinterpret_cast<float*>(color), reinterpret_cast<float*>(albedo),
                   reinterpret_cast<float*>(normal), noise);
}

// ----------------------------------------------------------- RenderServer
// Two lanes: the main pool, and (ServerConfig::expressMaxPaths > 0) a small
// express pool for requests of at most that many paths, so that a short
// request takes its few bounce iterations at the express pool's pace instead
// of one large-pool iteration per bounce. A step runs up to
// expressIterations express iterations and then one main iteration
// (time-sliced), or, with expressOverlap, launches the main iteration and
// runs express iterations on the express engine's own streams while it is
// on the device (at least one, at most expressIterations). Both lanes charge
// one memory budget. Server request ids map to (lane, lane request id).
struct ServerImpl {
    ServerConfig cfg;
    uint64_t footprint = 0;  // accumulation + image bytes admitted, both lanes
    std::unique_ptr<detail::Engine> eng, express;
    struct Route {
        detail::Engine* lane;
        uint64_t id;
    };
    std::map<uint64_t, Route> routes;
    std::map<std::pair<detail::Engine*, uint64_t>, uint64_t> back;
    std::deque<uint64_t> completed;
    uint64_t nextId = 1;
    RenderStats stats;

    explicit ServerImpl(const ServerConfig& c) : cfg(c) {
        ensure_context();
        detail::EngineConfig e;
        e.capacity = std::max(c.maxPathsInFlight, 1024u);
        e.maxChunkPaths = c.maxChunkPaths;
        e.memoryBudgetBytes = c.memoryBudgetBytes;
        e.timeKernels = c.timeKernels;
        e.shortestFirst = c.shortestFirst;
        e.roundRobin = c.roundRobin;
        e.sharedFootprint = &footprint;
        e.fusedMaxPaths = c.fusedMaxPaths;
        e.backfillPaths = c.backfillPaths;
        eng = std::make_unique<detail::Engine>(e);
        eng->ensureCapacity(e.capacity);
        if (c.expressMaxPaths) {
            if (c.expressIterations == 0) throw std::runtime_error("RenderServer: expressIterations must be > 0");
            detail::EngineConfig x = e;
            x.capacity = std::max(c.expressPathsInFlight, 1024u);
            x.maxChunkPaths = std::min(c.maxChunkPaths ? c.maxChunkPaths : x.capacity, x.capacity);
            express = std::make_unique<detail::Engine>(x);
            express->ensureCapacity(x.capacity);
        }
    }
    const Route& route(uint64_t id) const {
        auto it = routes.find(id);
        if (it == routes.end()) throw std::runtime_error("unknown request id");
        return it->second;
    }
    void collect(detail::Engine& lane) {
        uint64_t ids[64];
        for (size_t n; (n = lane.takeCompleted(ids, 64)) != 0;)
            for (size_t i = 0; i < n; ++i) completed.push_back(back.at({&lane, ids[i]}));
    }
    bool busy() const { return eng->busy() || (express && express->busy()); }
};

RenderServer::RenderServer(const ServerConfig& config) : m_impl(new ServerImpl(config)) {}
RenderServer::~RenderServer() { delete m_impl; }

uint64_t RenderServer::submit(Scene* scene, const Camera& camera, const RenderParams& params) {
    if (params.width == 0 || params.height == 0 || params.spp == 0) throw std::runtime_error("render: empty image");
    ServerImpl& s = *m_impl;
    wrk_render_params k;
    std::memset(&k, 0, sizeof(k));
    k.width = params.width; k.height = params.height; k.useNEE = params.useNEE ? 1u : 0u;
    k.maxBounces = params.maxBounces; k.rrStartDepth = params.russianRouletteDepth;
    k.fireflyClamp = params.fireflyClamp; k.iterationCap = s.cfg.iterationCap;
    uint64_t total = (uint64_t)params.width * params.height * params.spp;
    const uint64_t first = (uint64_t)params.firstSample * params.width * params.height;
    detail::Engine* lane = s.express && total <= s.cfg.expressMaxPaths ? s.express.get() : s.eng.get();
    const uint64_t laneId = lane->submit(scene, detail::camera_record(camera), k, params.spp, first, total,
                                         detail::Engine::Mode::Accumulate, nullptr);
    const uint64_t id = s.nextId++;
    s.routes[id] = {lane, laneId};
    s.back[{lane, laneId}] = id;
    return id;
}
bool RenderServer::step() {
    ServerImpl& s = *m_impl;
    const bool main = s.eng->busy();
    if (s.express && !s.cfg.expressOverlap) {
        for (uint32_t i = 0; i < s.cfg.expressIterations && s.express->busy(); ++i) s.express->step();
        s.collect(*s.express);
    }
    if (main) s.eng->launch();
    if (s.express && s.cfg.expressOverlap) {
        try {
            for (uint32_t i = 0; i < s.cfg.expressIterations && s.express->busy(); ++i) {
                if (i > 0 && main && s.eng->iterationDone()) break;
                s.express->step();
            }
        } catch (...) {
            // the launched main iteration still completes and is accounted
            if (main) s.eng->finish();
            throw;
        }
        s.collect(*s.express);
    }
    if (main) s.eng->finish();
    s.collect(*s.eng);
    return s.busy();
}
bool RenderServer::isDone(uint64_t id) const {
    const auto& r = m_impl->route(id);
    return r.lane->isDone(r.id);
}
uint32_t RenderServer::takeCompleted(uint64_t* ids, uint32_t maxIds) {
    auto& c = m_impl->completed;
    const size_t n = std::min<size_t>(maxIds, c.size());
    std::copy(c.begin(), c.begin() + n, ids);
    c.erase(c.begin(), c.begin() + n);
    return (uint32_t)n;
}
RequestStatus RenderServer::status(uint64_t id) const {
    const auto& rt = m_impl->route(id);
    const auto& r = rt.lane->request(rt.id);
    RequestStatus s;
    s.done = r.done;
    s.admitted = r.admitted;
    s.submitTime = r.tSubmit;
    s.admitTime = r.tAdmit;
    s.doneTime = r.tDone;
    s.paths = r.endPath - r.firstPath;
    return s;
}
void RenderServer::fetch(uint64_t id, Vec3* out) {
    ServerImpl& s = *m_impl;
    const ServerImpl::Route rt = s.route(id);
    rt.lane->fetch(rt.id, reinterpret_cast<float*>(out));
    s.back.erase({rt.lane, rt.id});
    s.routes.erase(id);
}
uint32_t RenderServer::numPending() const {
    const ServerImpl& s = *m_impl;
    return (uint32_t)(s.eng->numPending() + (s.express ? s.express->numPending() : 0));
}
uint32_t RenderServer::numInFlight() const {
    const ServerImpl& s = *m_impl;
    return (uint32_t)(s.eng->numActiveSlots() + (s.express ? s.express->numActiveSlots() : 0));
}
const RenderStats& RenderServer::getStats() const {
    ServerImpl& s = *m_impl;
    s.stats = s.eng->stats();
    if (s.express) {
        const RenderStats& x = s.express->stats();
        RenderStats& t = s.stats;
        t.seconds += x.seconds; t.paths += x.paths; t.closestRays += x.closestRays; t.shadowRays += x.shadowRays;
        t.iterations += x.iterations; t.batches += x.batches;
        t.msGenerate += x.msGenerate; t.msTraceClosest += x.msTraceClosest; t.msTraceAny += x.msTraceAny;
        t.msClassify += x.msClassify; t.msPartition += x.msPartition; t.msShade += x.msShade;
        t.msAccumulate += x.msAccumulate;
        t.launchesTraceClosest += x.launchesTraceClosest; t.launchesTraceAny += x.launchesTraceAny;
        t.closestNodeFetches += x.closestNodeFetches; t.closestTriFetches += x.closestTriFetches;
        t.anyNodeFetches += x.anyNodeFetches; t.anyTriFetches += x.anyTriFetches;
        t.classifyItems += x.classifyItems; t.partitionItems += x.partitionItems; t.shadeItems += x.shadeItems;
        t.fusedPaths += x.fusedPaths; t.msFused += x.msFused;
        t.closestStackSpills += x.closestStackSpills; t.anyStackSpills += x.anyStackSpills;
    }
    return s.stats;
}
double RenderServer::clock() { return detail::Engine::now(); }

}  // namespace wr
