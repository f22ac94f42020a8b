// A libWR-style caller: builds a Cornell box through the wr:: API exactly as
// libWR's test programs do (Context -> createScene -> add*Material ->
// addTriangleMesh(std::span) -> finalize -> createRenderer -> render) and
// prints the mean of the returned image. Only the GLM spelling differs
// (wr::radians instead of glm::radians).
#include <wr/wr.h>

#include <cstdio>
#include <vector>

using namespace wr;

static void quad(Scene* s, const float (&v)[12], uint32_t mat) {
    const uint32_t idx[6] = {0, 1, 2, 0, 2, 3};
    s->addTriangleMesh(std::span<const float>(v, 12), std::span<const uint32_t>(idx, 6), mat);
}

int main(int argc, char** argv) {
    const uint32_t W = argc > 1 ? (uint32_t)atoi(argv[1]) : 64;
    const uint32_t spp = argc > 2 ? (uint32_t)atoi(argv[2]) : 16;
    ContextConfig cc;
    cc.logLevel = 0;
    Context context(cc);
    Scene* scene = context.createScene();
    uint32_t white = scene->addLambertianMaterial(Vec3(0.73f, 0.73f, 0.73f));
    uint32_t red = scene->addLambertianMaterial(Vec3(0.65f, 0.05f, 0.05f));
    uint32_t green = scene->addLambertianMaterial(Vec3(0.12f, 0.45f, 0.15f));
    uint32_t light = scene->addEmissiveMaterial(Vec3(17.0f, 12.0f, 4.0f));
    quad(scene, {-1, -1, -1, 1, -1, -1, 1, -1, 1, -1, -1, 1}, white);
    quad(scene, {-1, 1, -1, -1, 1, 1, 1, 1, 1, 1, 1, -1}, white);
    quad(scene, {-1, -1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1}, white);
    quad(scene, {-1, -1, -1, -1, -1, 1, -1, 1, 1, -1, 1, -1}, red);
    quad(scene, {1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1}, green);
    quad(scene, {-0.25f, 0.99f, -0.25f, 0.25f, 0.99f, -0.25f, 0.25f, 0.99f, 0.25f, -0.25f, 0.99f, 0.25f}, light);
    scene->setEnvironmentRadiance(Vec3(0.0f, 0.0f, 0.0f));
    scene->finalize();

    Renderer* renderer = context.createRenderer();
    Camera camera;
    camera.position = Vec3(0.0f, 0.0f, 3.0f);
    camera.target = Vec3(0.0f, 0.0f, 2.0f);
    camera.up = Vec3(0.0f, 1.0f, 0.0f);
    camera.fovY = radians(45.0f);
    camera.aspect = 1.0f;

    std::vector<Vec3> image(W * W);
    RenderParams params;
    params.width = W;
    params.height = W;
    params.spp = spp;
    params.useNEE = true;
    params.maxBounces = 8;
    params.russianRouletteDepth = 3.0f;
    renderer->render(scene, camera, image.data(), params);

    double sum[3] = {0, 0, 0};
    for (const Vec3& p : image) { sum[0] += p.x; sum[1] += p.y; sum[2] += p.z; }
    std::printf("mean %.6f %.6f %.6f\n", sum[0] / image.size(), sum[1] / image.size(), sum[2] / image.size());
    delete renderer;
    delete scene;
    return 0;
}
