// A C++ serving loop over wr::RenderServer: requests of several sizes (the
// small ones take the express lane) are submitted while others are in
// flight, stepped to completion and fetched; each image must equal the same
// request rendered alone with wr::Renderer, bit for bit. Prints
// "equal <k>/<n>" and exits with 0 only if every image matches.
#include <wr/wr.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

using namespace wr;

static void quad(Scene* s, const float (&v)[12], uint32_t mat) {
    const uint32_t idx[6] = {0, 1, 2, 0, 2, 3};
    s->addTriangleMesh(std::span<const float>(v, 12), std::span<const uint32_t>(idx, 6), mat);
}

int main() {
    ContextConfig cc;
    cc.logLevel = 0;
    Context context(cc);
    Scene* scene = context.createScene();
    uint32_t white = scene->addLambertianMaterial(Vec3(0.73f, 0.73f, 0.73f));
    uint32_t red = scene->addLambertianMaterial(Vec3(0.65f, 0.05f, 0.05f));
    uint32_t green = scene->addLambertianMaterial(Vec3(0.12f, 0.45f, 0.15f));
    uint32_t light = scene->addEmissiveMaterial(Vec3(17.0f, 12.0f, 4.0f));
    uint32_t glass = scene->addGlassMaterial(Vec3(1.0f, 1.0f, 1.0f), 1.5f);
    quad(scene, {-1, -1, -1, 1, -1, -1, 1, -1, 1, -1, -1, 1}, white);
    quad(scene, {-1, 1, -1, -1, 1, 1, 1, 1, 1, 1, 1, -1}, white);
    quad(scene, {-1, -1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1}, white);
    quad(scene, {-1, -1, -1, -1, -1, 1, -1, 1, 1, -1, 1, -1}, red);
    quad(scene, {1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1}, green);
    quad(scene, {-0.4f, -0.5f, 0.0f, 0.4f, -0.5f, 0.0f, 0.4f, 0.3f, -0.2f, -0.4f, 0.3f, -0.2f}, glass);
    quad(scene, {-0.25f, 0.99f, -0.25f, 0.25f, 0.99f, -0.25f, 0.25f, 0.99f, 0.25f, -0.25f, 0.99f, 0.25f}, light);
    scene->finalize();

    Camera camera;
    camera.position = Vec3(0.0f, 0.0f, 3.0f);
    camera.target = Vec3(0.0f, 0.0f, 2.0f);
    camera.up = Vec3(0.0f, 1.0f, 0.0f);
    camera.fovY = radians(45.0f);
    camera.aspect = 1.0f;

    const uint32_t sizes[][2] = {{128, 16}, {32, 1}, {64, 4}, {16, 2}, {96, 8}, {48, 1}};  // side, spp
    std::vector<RenderParams> reqs;
    for (auto& s : sizes) {
        RenderParams p;
        p.width = p.height = s[0];
        p.spp = s[1];
        p.maxBounces = 8;
        p.russianRouletteDepth = 3.0f;
        reqs.push_back(p);
    }

    ServerConfig sc;
    sc.maxPathsInFlight = 1u << 18;
    sc.maxChunkPaths = 1u << 16;
    sc.expressMaxPaths = 4096;        // the 32x1, 16x2 and 48x1 requests
    sc.expressPathsInFlight = 4096;
    sc.expressIterations = 4;
    RenderServer server(sc);
    std::map<uint64_t, size_t> which;
    std::map<size_t, std::vector<Vec3>> served;
    size_t next = 0;
    for (int step = 0; step < 100000; ++step) {
        if (next < reqs.size()) {  // one new request per step, while others are in flight
            which[server.submit(scene, camera, reqs[next])] = next;
            ++next;
        }
        const bool more = server.step();
        uint64_t done[16];
        const uint32_t n = server.takeCompleted(done, 16);
        for (uint32_t k = 0; k < n; ++k) {
            const size_t i = which.at(done[k]);
            served[i].resize((size_t)reqs[i].width * reqs[i].height);
            server.fetch(done[k], served[i].data());
        }
        if (!more && next == reqs.size()) break;
    }

    Renderer* alone = context.createRenderer();
    WavefrontConfig wc;
    wc.iterationCap = 0;  // the server's default: paths run to completion
    alone->setWavefrontConfig(wc);
    size_t equal = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        std::vector<Vec3> img((size_t)reqs[i].width * reqs[i].height);
        alone->render(scene, camera, img.data(), reqs[i]);
        if (served.count(i) && std::memcmp(img.data(), served[i].data(), img.size() * sizeof(Vec3)) == 0) ++equal;
    }
    std::printf("equal %zu/%zu\n", equal, reqs.size());
    delete alone;
    delete scene;
    return equal == reqs.size() ? 0 : 1;
}
