// wr_host.cpp — the wr:: C++ API (Context, Scene, Texture2D, Renderer) over
// the wrk_* C ABI. Mirrors libWR/src/internal/{context,scene,renderer}.cpp:
// same method semantics, exception types and messages; the OptiX/CUDA
// machinery is replaced by HIP device buffers, the GPU LBVH and the
// wavefront kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wr/wr.h"
#include "../../include/wrk.h"
#include "image_io.h"

namespace wr {

namespace {

int g_logLevel = 4;
bool g_contextReady = false;
int g_deviceId = 0;

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
void wrk_check(int rc, const char* what) {
    if (rc != WRK_OK) throw std::runtime_error(std::string(what) + ": " + wrk_last_error());
}

// Owning device allocation.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && ptr) return;
        release();
        if (n == 0) return;
        hip_check(hipMalloc(&ptr, n), "hipMalloc");
        bytes = n;
    }
    void upload(const void* src, size_t n) {
        alloc(n);
        if (n) hip_check(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    uint64_t addr() const { return (uint64_t)(uintptr_t)ptr; }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

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

    DevBuf d_vertices, d_normals, d_uvs, d_indices, d_triMat, d_materials, d_textures, d_emissive, d_cdf;
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
    v.numTris = s.numTriangles;
    v.numMaterials = (uint32_t)s.materials.size();
    v.numEmissive = (uint32_t)s.emissiveTriangles.size();
    v.envRadiance[0] = s.environmentRadiance.x;
    v.envRadiance[1] = s.environmentRadiance.y;
    v.envRadiance[2] = s.environmentRadiance.z;
    v.envRadiance[3] = 0.0f;
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
struct RendererImpl {
    WavefrontConfig wf;
    RenderStats stats;
    hipStream_t stream = nullptr;
    uint32_t capacity = 0;
    DevBuf org, dir, thr, rad, st, contrib, savedOrg, savedDir, hit, occl, result;
    DevBuf qA, qB, cat, scratch, offs;
    DevBuf accum, image;
    uint32_t* hostCounts = nullptr;
    wrk_paths paths{};
    DevBuf fetchCounters;  // 4 x uint64: closest nodes/tris, any nodes/tris
    hipEvent_t ev[16] = {};

    RendererImpl() {
        ensure_context();
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipHostMalloc((void**)&hostCounts, 64, hipHostMallocDefault), "hipHostMalloc");
        for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        fetchCounters.alloc(64);
    }
    ~RendererImpl() {
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (hostCounts) (void)hipHostFree(hostCounts);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void reserve(uint32_t cap) {
        if (cap <= capacity) return;
        size_t c = cap;
        org.alloc(c * 16); dir.alloc(c * 16); thr.alloc(c * 16); rad.alloc(c * 16); st.alloc(c * 16);
        contrib.alloc(c * 16); savedOrg.alloc(c * 16); savedDir.alloc(c * 16); hit.alloc(c * 8);
        occl.alloc(c); result.alloc(c * 16);
        qA.alloc(c * 4); qB.alloc(c * 4); cat.alloc(c);
        scratch.alloc(wrk_partition_scratch_bytes(cap, 8));
        offs.alloc(64);
        capacity = cap;
        paths.org = org.addr(); paths.dir = dir.addr(); paths.thr = thr.addr(); paths.rad = rad.addr();
        paths.state = st.addr(); paths.contrib = contrib.addr(); paths.savedOrg = savedOrg.addr();
        paths.savedDir = savedDir.addr(); paths.hit = hit.addr(); paths.occluded = occl.addr();
        paths.result = result.addr(); paths.capacity = cap;
    }

    static wrk_camera camera_of(const Camera& c) {
        // renderer.cpp:144-156
        Vec3 f = c.getForward(), r = c.getRight(), u = c.getUp();
        wrk_camera k;
        std::memset(&k, 0, sizeof(k));
        k.position[0] = c.position.x; k.position[1] = c.position.y; k.position[2] = c.position.z;
        k.forward[0] = f.x; k.forward[1] = f.y; k.forward[2] = f.z;
        k.right[0] = r.x; k.right[1] = r.y; k.right[2] = r.z;
        k.up[0] = u.x; k.up[1] = u.y; k.up[2] = u.z;
        k.tanHalfFovY = std::tan(c.fovY * 0.5f);
        k.aspect = c.aspect;
        k.focalDistance = c.focalDistance;
        k.lensRadius = c.lensRadius;
        return k;
    }
    wrk_render_params params_of(const RenderParams& p) const {
        wrk_render_params k;
        std::memset(&k, 0, sizeof(k));
        k.width = p.width; k.height = p.height; k.useNEE = p.useNEE ? 1u : 0u; k.maxBounces = p.maxBounces;
        k.rrStartDepth = p.russianRouletteDepth; k.fireflyClamp = p.fireflyClamp; k.iterationCap = wf.iterationCap;
        return k;
    }

    // One batch of paths [first, first+n) through the wavefront loop
    // (renderer.cpp:190-284 with device-resident, ordered queues).
    float elapsed(int a, int b) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, ev[a], ev[b]), "hipEventElapsedTime");
        return ms;
    }
    void mark(int i) {
        if (wf.timeKernels) hip_check(hipEventRecord(ev[i], stream), "hipEventRecord");
    }

    // One batch of paths [first, first+n) through the wavefront loop
    // (renderer.cpp:190-284 with device-resident, ordered queues). One
    // host synchronisation per iteration reads the next queue sizes.
    void run_batch(const Scene* scene, const wrk_camera& cam, const wrk_render_params& prm, uint64_t first, uint32_t n) {
        const wrk_scene_view& view = scene->getDeviceView();
        const wrk_bvh& bvh = scene->getBVH();
        uint32_t* qa = qA.as<uint32_t>();
        uint32_t* qb = qB.as<uint32_t>();
        uint32_t* off = offs.as<uint32_t>();
        unsigned long long* fc = fetchCounters.as<unsigned long long>();
        mark(0);
        wrk_check(wrk_generate(&cam, &prm, &paths, first, n, qa, stream), "wrk_generate");
        mark(1);
        if (wf.timeKernels) {
            hip_check(hipEventSynchronize(ev[1]), "hipEventSynchronize");
            stats.msGenerate += elapsed(0, 1);
        }
        uint32_t nC = n, nS = 0;
        const uint32_t cap = prm.iterationCap;
        const uint32_t safety = 4 * prm.maxBounces + 16;
        for (uint32_t it = 0;; ++it) {
            if (cap && it >= cap) break;
            if (it >= safety) throw std::runtime_error("wavefront loop did not drain");
            if (nC + nS == 0) break;
            uint32_t active = nC + nS;
            mark(2);
            if (nC) {
                if (wf.countTraversal)
                    wrk_check(wrk_trace_closest_counted(&bvh, org.ptr, dir.ptr, qa, nC, nullptr, hit.as<uint32_t>(), fc,
                                                        stream), "wrk_trace_closest_counted");
                else
                    wrk_check(wrk_trace_closest(&bvh, org.ptr, dir.ptr, qa, nC, nullptr, hit.as<uint32_t>(), stream),
                              "wrk_trace_closest");
            }
            mark(3);
            if (nS) {
                if (wf.countTraversal)
                    wrk_check(wrk_trace_any_counted(&bvh, org.ptr, dir.ptr, qa + nC, nS, nullptr, occl.as<uint8_t>(),
                                                    fc + 2, stream), "wrk_trace_any_counted");
                else
                    wrk_check(wrk_trace_any(&bvh, org.ptr, dir.ptr, qa + nC, nS, nullptr, occl.as<uint8_t>(), stream),
                              "wrk_trace_any");
            }
            mark(4);
            stats.closestRays += nC;
            stats.shadowRays += nS;
            if (nC) stats.launchesTraceClosest++;
            if (nS) stats.launchesTraceAny++;
            wrk_check(wrk_classify(&view, &prm, &paths, qa, nC, nS, cat.as<uint8_t>(), stream), "wrk_classify");
            mark(5);
            wrk_check(wrk_partition(cat.as<uint8_t>(), qa, active, 5, qb, off, scratch.ptr, scratch.bytes, stream),
                      "wrk_partition");
            mark(6);
            uint32_t last = (cap && it + 1 == cap) ? 1u : 0u;
            wrk_check(wrk_shade(&view, &prm, &paths, qb, off, active, last, cat.as<uint8_t>(), stream), "wrk_shade");
            mark(7);
            wrk_check(wrk_partition(cat.as<uint8_t>(), qb, active, 2, qa, off + 8, scratch.ptr, scratch.bytes, stream),
                      "wrk_partition");
            mark(8);
            hip_check(hipMemcpyAsync(hostCounts, off + 8, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            if (wf.timeKernels) {
                stats.msTraceClosest += elapsed(2, 3);
                stats.msTraceAny += elapsed(3, 4);
                stats.msClassify += elapsed(4, 5);
                stats.msPartition += elapsed(5, 6) + elapsed(7, 8);
                stats.msShade += elapsed(6, 7);
            }
            nC = hostCounts[1] - hostCounts[0];
            nS = hostCounts[2] - hostCounts[1];
            stats.iterations++;
        }
        stats.batches++;
    }

    void render(Scene* scene, const Camera& camera, const RenderParams& rp, float* outHost, bool accumulateOnly) {
        if (!scene) throw std::runtime_error("render: null scene");
        if (rp.width == 0 || rp.height == 0 || rp.spp == 0) throw std::runtime_error("render: empty image");
        scene->finalize();
        const uint64_t npix = (uint64_t)rp.width * rp.height;
        const uint64_t total = npix * rp.spp;
        uint32_t cap = (uint32_t)std::min<uint64_t>(total, std::max<uint32_t>(wf.maxPathsInFlight, 1024u));
        reserve(cap);
        accum.alloc(npix * 3 * sizeof(float));
        image.alloc(npix * 3 * sizeof(float));
        hip_check(hipMemsetAsync(accum.ptr, 0, npix * 3 * sizeof(float), stream), "hipMemsetAsync");
        wrk_camera cam = camera_of(camera);
        wrk_render_params prm = params_of(rp);
        stats = RenderStats{};
        hip_check(hipMemsetAsync(fetchCounters.ptr, 0, 32, stream), "hipMemsetAsync");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        auto t0 = std::chrono::steady_clock::now();
        for (uint64_t first = 0; first < total; first += cap) {
            uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - first);
            run_batch(scene, cam, prm, first, n);
            mark(9);
            wrk_check(wrk_accumulate(&paths, first, n, (uint32_t)npix, accum.as<float>(), stream), "wrk_accumulate");
            mark(10);
            if (wf.timeKernels) {
                hip_check(hipEventSynchronize(ev[10]), "hipEventSynchronize");
                stats.msAccumulate += elapsed(9, 10);
            }
        }
        wrk_check(wrk_average(accum.as<float>(), (uint32_t)npix, rp.spp, image.as<float>(), stream), "wrk_average");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats.paths = total;
        if (wf.countTraversal) {
            unsigned long long c[4];
            hip_check(hipMemcpy(c, fetchCounters.ptr, sizeof(c), hipMemcpyDeviceToHost), "hipMemcpy D2H");
            stats.closestNodeFetches = c[0]; stats.closestTriFetches = c[1];
            stats.anyNodeFetches = c[2]; stats.anyTriFetches = c[3];
        }
        if (outHost && !accumulateOnly)
            hip_check(hipMemcpy(outHost, image.ptr, npix * 3 * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }

    void render_paths(Scene* scene, const Camera& camera, const RenderParams& rp, uint64_t firstPath, uint32_t count,
                      float* out) {
        if (!scene) throw std::runtime_error("renderPaths: null scene");
        scene->finalize();
        uint32_t cap = std::max<uint32_t>(std::min<uint32_t>(count, wf.maxPathsInFlight), 1u);
        reserve(cap);
        wrk_camera cam = camera_of(camera);
        wrk_render_params prm = params_of(rp);
        stats = RenderStats{};
        for (uint64_t done = 0; done < count; done += cap) {
            uint32_t n = (uint32_t)std::min<uint64_t>(cap, count - done);
            run_batch(scene, cam, prm, firstPath + done, n);
            hip_check(hipMemcpyAsync(out + done * 4, result.ptr, (size_t)n * 16, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        }
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
    m_impl->render(scene, camera, params, reinterpret_cast<float*>(outputBuffer), false);
    if (params.denoiser.enabled && g_logLevel >= 2)
        std::cout << "[Renderer] denoiser requested: not available in this build, returning the averaged image" << std::endl;
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
WavefrontConfig Renderer::getWavefrontConfig() const { return m_impl->wf; }
void Renderer::renderPaths(Scene* scene, const Camera& camera, const RenderParams& params, uint64_t firstPath,
                           uint32_t count, float* out) {
    m_impl->render_paths(scene, camera, params, firstPath, count, out);
}
const RenderStats& Renderer::getStats() const { return m_impl->stats; }

}  // namespace wr
