#pragma once
// libWR/include/wr/scene.h:1-105 — same class, method names and signatures.
// Additive: Scene::addTextureRGBA8 (in-memory RGBA8 texture) and
// Scene::getBVH/getDeviceView for the kernel layer.
#include <cstdint>
#include <span>
#include <string>

#include "wr/types.h"

struct wrk_bvh;
struct wrk_scene_view;

namespace wr {

struct SceneBuildConfig {
    bool allowUpdate = false;
    bool allowCompaction = true;
    bool preferFastTrace = true;
};

struct SceneImpl;
class Scene;

using DevicePtr = unsigned long long;
using TraversableHandle = unsigned long long;
using TextureHandle = unsigned long long;

// Texture2D - RGBA8 format, supports PNG loading
class Texture2D {
public:
    static uint32_t create(Scene* scene, const std::string& path);
    static void destroy(Scene* scene, uint32_t textureId);
};

class Scene {
public:
    Scene();
    ~Scene();

    uint32_t addLambertianMaterial(const Vec3& albedo);
    uint32_t addLambertianMaterial(const Vec3& albedo, uint32_t albedoTextureId);
    uint32_t addEmissiveMaterial(const Vec3& emission);
    uint32_t addGlassMaterial(const Vec3& albedo, float ior);
    uint32_t addGGXReflectionMaterial(const Vec3& albedo, float roughness, float metallic = 0.0f);
    uint32_t addGGXTransmissionMaterial(const Vec3& albedo, float roughness, float ior);

    void setMaterialNormalMap(uint32_t materialId, uint32_t normalTextureId);

    void addTriangleMesh(std::span<const float> vertices, std::span<const uint32_t> indices, uint32_t materialId);
    void addTriangleMesh(std::span<const float> vertices, std::span<const uint32_t> indices,
                         std::span<const float> uvs, uint32_t materialId);
    void addTriangleMesh(std::span<const float> vertices, std::span<const uint32_t> indices,
                         std::span<const float> normals, std::span<const float> uvs, uint32_t materialId);

    void setEnvironmentRadiance(const Vec3& radiance);
    void setEnvironmentMap(const std::string& hdrPath);

    // Builds the GPU BVH and uploads geometry, materials, textures and the
    // emissive-triangle CDF.
    void finalize(const SceneBuildConfig& config = {});

    DevicePtr getVerticesBuffer() const;
    DevicePtr getNormalsBuffer() const;
    DevicePtr getIndicesBuffer() const;
    DevicePtr getTriangleMaterialIdsBuffer() const;
    DevicePtr getMaterialsBuffer() const;
    TraversableHandle getGASHandle() const;  // address of the device-side wrk_bvh record
    uint32_t getNumMaterials() const;
    Vec3 getEnvironmentRadiance() const;

    uint64_t getEnvironmentMapTexture() const;  // device pointer to RGBA32F texels (0 if none)
    uint32_t getEnvironmentMapWidth() const;
    uint32_t getEnvironmentMapHeight() const;

    DevicePtr getEmissiveTrianglesBuffer() const;
    DevicePtr getEmissiveTriangleCDFBuffer() const;
    uint32_t getNumEmissiveTriangles() const;

    DevicePtr getTexturesBuffer() const;
    uint32_t getNumTextures() const;
    DevicePtr getUVsBuffer() const;

    uint32_t loadTexture(const std::string& path);
    void destroyTexture(uint32_t textureId);

    // --- additive API
    // RGBA8 texels, row-major, width*height*4 bytes; returns a 1-based id.
    uint32_t addTextureRGBA8(uint32_t width, uint32_t height, std::span<const uint8_t> rgba);
    const wrk_bvh& getBVH() const;
    const wrk_scene_view& getDeviceView() const;
    uint32_t getNumTriangles() const;
    double getBuildSeconds() const;  // wall time of the last BVH build

private:
    SceneImpl* m_impl;
};

}  // namespace wr
