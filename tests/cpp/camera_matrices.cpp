// Exercises the value types of libWR's public header (wr/types.h) the way a
// libWR caller does: Camera::getViewMatrix / getProjectionMatrix and the
// Mat3 / Mat4 / Vec4 types, printing the quantities a test checks.
#include <wr/wr.h>

#include <cstdio>

using namespace wr;

static void print4(const char* tag, const Vec4& v) { std::printf("%s %.7g %.7g %.7g %.7g\n", tag, v.x, v.y, v.z, v.w); }

int main() {
    Camera cam;
    cam.position = Vec3(1.0f, 2.0f, 3.0f);
    cam.target = Vec3(0.5f, 1.0f, -1.0f);
    cam.up = Vec3(0.0f, 1.0f, 0.0f);
    cam.fovY = radians(50.0f);
    cam.aspect = 1.5f;
    const Mat4 V = cam.getViewMatrix();
    const Mat4 P = cam.getProjectionMatrix();         // near 0.1, far 100
    const Mat4 P2 = cam.getProjectionMatrix(1.0f, 10.0f);
    print4("eye", V * Vec4(cam.position, 1.0f));
    print4("target", V * Vec4(cam.target, 1.0f));
    const Vec3 f = cam.getForward(), r = cam.getRight(), u = cam.getUp();
    print4("right", V * Vec4(r, 0.0f));
    print4("up", V * Vec4(u, 0.0f));
    print4("forward", V * Vec4(f, 0.0f));
    // points on the axis at the near and far planes, and the corner of the
    // near plane's visible window
    print4("near", P2 * Vec4(0.0f, 0.0f, -1.0f, 1.0f));
    print4("far", P2 * Vec4(0.0f, 0.0f, -10.0f, 1.0f));
    const float t = std::tan(cam.fovY / 2.0f);
    print4("corner", P2 * Vec4(cam.aspect * t * 1.0f, t * 1.0f, -1.0f, 1.0f));
    print4("pv", (P * V) * Vec4(cam.target, 1.0f));
    Mat3 m3;
    const Vec3 x = m3 * Vec3(1.0f, 2.0f, 3.0f);
    std::printf("mat3 %g %g %g sizes %zu %zu %zu\n", x.x, x.y, x.z, sizeof(Mat3), sizeof(Mat4), sizeof(Vec4));
    return 0;
}
