// wr_oracle.cpp — CPU restatement of libWR's wavefront path tracer (test
// infrastructure only; see wr_oracle.hpp). Reference files are cited as
// libWR/<path>:<line>.
#include "wr_oracle.hpp"

#include <cstdio>
#include <limits>

namespace orc {

// ------------------------------------------------------------- scene prep
// libWR/src/internal/scene.cpp:423-491
void build_light_cdf(Scene& s) {
    s.emissiveTris.clear();
    s.emissiveCdf.clear();
    float totalPower = 0.0f;
    const uint32_t nt = s.numTris();
    for (uint32_t tri = 0; tri < nt; ++tri) {
        uint32_t matId = s.triMat[tri];
        if (matId >= s.materials.size()) continue;  // reference indexes out of range here
        const wrk_material& mat = s.materials[matId];
        if (mat.type != WRK_MAT_EMISSIVE) continue;
        s.emissiveTris.push_back(tri);
        uint32_t i0 = s.indices[tri * 3 + 0], i1 = s.indices[tri * 3 + 1], i2 = s.indices[tri * 3 + 2];
        const float* v0 = &s.vertices[i0 * 3];
        const float* v1 = &s.vertices[i1 * 3];
        const float* v2 = &s.vertices[i2 * 3];
        float e1x = v1[0] - v0[0], e1y = v1[1] - v0[1], e1z = v1[2] - v0[2];
        float e2x = v2[0] - v0[0], e2y = v2[1] - v0[1], e2z = v2[2] - v0[2];
        float cx = e1y * e2z - e1z * e2y;
        float cy = e1z * e2x - e1x * e2z;
        float cz = e1x * e2y - e1y * e2x;
        float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
        float power = (mat.emission[0] + mat.emission[1] + mat.emission[2]) * area;
        totalPower += power;
        s.emissiveCdf.push_back(totalPower);
    }
    for (float& c : s.emissiveCdf) c /= totalPower;
}

// ------------------------------------------------------------ host BVH
namespace {
struct Build {
    Scene& s;
    std::vector<double> cen;  // centroid per tri
    std::vector<double> blo, bhi;
    explicit Build(Scene& sc) : s(sc) {}
    void tri_box(uint32_t t, double lo[3], double hi[3]) const {
        for (int a = 0; a < 3; ++a) { lo[a] = blo[t * 3 + a]; hi[a] = bhi[t * 3 + a]; }
    }
    uint32_t build(uint32_t start, uint32_t count) {
        HNode n{};
        for (int a = 0; a < 3; ++a) { n.lo[a] = 1e300; n.hi[a] = -1e300; }
        double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t i = start; i < start + count; ++i) {
            uint32_t t = s.order[i];
            for (int a = 0; a < 3; ++a) {
                n.lo[a] = std::min(n.lo[a], blo[t * 3 + a]);
                n.hi[a] = std::max(n.hi[a], bhi[t * 3 + a]);
                clo[a] = std::min(clo[a], cen[t * 3 + a]);
                chi[a] = std::max(chi[a], cen[t * 3 + a]);
            }
        }
        uint32_t idx = (uint32_t)s.nodes.size();
        s.nodes.push_back(n);
        if (count <= 4) {
            s.nodes[idx].start = start; s.nodes[idx].count = count;
            s.nodes[idx].left = s.nodes[idx].right = 0xFFFFFFFFu;
            return idx;
        }
        int axis = 0;
        double ext = chi[0] - clo[0];
        for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
        uint32_t mid = start + count / 2;
        if (ext > 0) {
            // binned SAH, 16 bins
            const int NB = 16;
            double bl[NB][3], bh[NB][3]; uint32_t bc[NB];
            for (int b = 0; b < NB; ++b) { bc[b] = 0; for (int a = 0; a < 3; ++a) { bl[b][a] = 1e300; bh[b][a] = -1e300; } }
            auto binOf = [&](uint32_t t) {
                int b = (int)((cen[t * 3 + axis] - clo[axis]) / ext * NB);
                return std::min(std::max(b, 0), NB - 1);
            };
            for (uint32_t i = start; i < start + count; ++i) {
                uint32_t t = s.order[i]; int b = binOf(t); bc[b]++;
                for (int a = 0; a < 3; ++a) { bl[b][a] = std::min(bl[b][a], blo[t * 3 + a]); bh[b][a] = std::max(bh[b][a], bhi[t * 3 + a]); }
            }
            auto area = [](const double lo[3], const double hi[3]) {
                double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
                return (dx < 0) ? 0.0 : dx * dy + dy * dz + dz * dx;
            };
            double best = 1e300; int bestB = -1;
            for (int split = 1; split < NB; ++split) {
                double l0[3] = {1e300, 1e300, 1e300}, l1[3] = {-1e300, -1e300, -1e300}; uint32_t lc = 0;
                double r0[3] = {1e300, 1e300, 1e300}, r1[3] = {-1e300, -1e300, -1e300}; uint32_t rc = 0;
                for (int b = 0; b < NB; ++b) {
                    if (!bc[b]) continue;
                    double* lo = b < split ? l0 : r0; double* hi = b < split ? l1 : r1;
                    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], bl[b][a]); hi[a] = std::max(hi[a], bh[b][a]); }
                    (b < split ? lc : rc) += bc[b];
                }
                if (!lc || !rc) continue;
                double c = area(l0, l1) * lc + area(r0, r1) * rc;
                if (c < best) { best = c; bestB = split; }
            }
            if (bestB > 0) {
                auto it = std::partition(s.order.begin() + start, s.order.begin() + start + count,
                                         [&](uint32_t t) { return binOf(t) < bestB; });
                mid = (uint32_t)(it - s.order.begin());
                if (mid == start || mid == start + count) mid = start + count / 2;
            } else {
                std::nth_element(s.order.begin() + start, s.order.begin() + mid, s.order.begin() + start + count,
                                 [&](uint32_t a, uint32_t b) { return cen[a * 3 + axis] < cen[b * 3 + axis]; });
            }
        }
        uint32_t l = build(start, mid - start);
        uint32_t r = build(mid, start + count - mid);
        s.nodes[idx].left = l; s.nodes[idx].right = r; s.nodes[idx].count = 0;
        return idx;
    }
};

inline bool box_hit(const HNode& n, const double o[3], const double inv[3], double tmin, double tmax,
                    double ext) {
    // FP64 slab test on a box padded outward by 2^-20 * (|coordinate| +
    // scene extent), far distance clamped to the best hit and then widened
    // by 1e-6: conservative with respect to the FP32 watertight triangle
    // test, whose rounding can report edge/vertex hits slightly outside the
    // exact bounds and ties at equal t.
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        double pad = (std::max(std::fabs(n.lo[a]), std::fabs(n.hi[a])) + ext) * 9.5367431640625e-07;
        double lo = n.lo[a] - pad, hi = n.hi[a] + pad;
        double ta = (lo - o[a]) * inv[a], tb = (hi - o[a]) * inv[a];
        if (std::isnan(ta) || std::isnan(tb)) { if (o[a] < lo || o[a] > hi) return false; continue; }
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta); t1 = std::min(t1, tb);
    }
    return t0 <= t1 * (1 + 1e-6) + 1e-30;
}
}  // namespace

void build_host_bvh(Scene& s) {
    const uint32_t nt = s.numTris();
    s.nodes.clear();
    s.order.resize(nt);
    for (uint32_t i = 0; i < nt; ++i) s.order[i] = i;
    if (nt == 0) return;
    Build b(s);
    b.cen.resize(nt * 3); b.blo.resize(nt * 3); b.bhi.resize(nt * 3);
    for (uint32_t t = 0; t < nt; ++t) {
        for (int a = 0; a < 3; ++a) {
            double x0 = s.vertices[s.indices[t * 3 + 0] * 3 + a];
            double x1 = s.vertices[s.indices[t * 3 + 1] * 3 + a];
            double x2 = s.vertices[s.indices[t * 3 + 2] * 3 + a];
            b.blo[t * 3 + a] = std::min(x0, std::min(x1, x2));
            b.bhi[t * 3 + a] = std::max(x0, std::max(x1, x2));
            b.cen[t * 3 + a] = (b.blo[t * 3 + a] + b.bhi[t * 3 + a]) * 0.5;
        }
    }
    s.extent = 0;
    for (int a = 0; a < 3; ++a) {
        double lo = 1e300, hi = -1e300;
        for (uint32_t t = 0; t < nt; ++t) { lo = std::min(lo, b.blo[t * 3 + a]); hi = std::max(hi, b.bhi[t * 3 + a]); }
        s.extent = std::max(s.extent, hi - lo);
    }
    s.nodes.reserve(2 * nt / 2 + 8);
    b.build(0, nt);
}

static inline void tri_verts(const Scene& s, uint32_t t, const float** v0, const float** v1, const float** v2) {
    *v0 = &s.vertices[s.indices[t * 3 + 0] * 3];
    *v1 = &s.vertices[s.indices[t * 3 + 1] * 3];
    *v2 = &s.vertices[s.indices[t * 3 + 2] * 3];
}

Hit closest_brute(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    RayPre r = ray_pre(o, d);
    Hit h{tmax, 0xFFFFFFFFu};
    const uint32_t nt = s.numTris();
    for (uint32_t t = 0; t < nt; ++t) {
        const float *v0, *v1, *v2;
        tri_verts(s, t, &v0, &v1, &v2);
        float th;
        if (intersect_tri(r, v0, v1, v2, tmin, h.t, &th)) {
            if (th < h.t || (th == h.t && t < h.prim)) { h.t = th; h.prim = t; }
        }
    }
    return h;
}

bool any_brute(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    RayPre r = ray_pre(o, d);
    const uint32_t nt = s.numTris();
    for (uint32_t t = 0; t < nt; ++t) {
        const float *v0, *v1, *v2;
        tri_verts(s, t, &v0, &v1, &v2);
        float th;
        if (intersect_tri(r, v0, v1, v2, tmin, tmax, &th)) return true;
    }
    return false;
}

template <bool ANY>
static Hit bvh_query(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    Hit h{tmax, 0xFFFFFFFFu};
    if (s.nodes.empty()) return h;
    RayPre r = ray_pre(o, d);
    double od[3] = {o[0], o[1], o[2]};
    double inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0 / (double)d[a];
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const HNode& n = s.nodes[stack[--sp]];
        if (!box_hit(n, od, inv, tmin, h.t, s.extent)) continue;
        if (n.left == 0xFFFFFFFFu) {
            for (uint32_t i = n.start; i < n.start + n.count; ++i) {
                uint32_t t = s.order[i];
                const float *v0, *v1, *v2;
                tri_verts(s, t, &v0, &v1, &v2);
                float th;
                if (intersect_tri(r, v0, v1, v2, tmin, h.t, &th)) {
                    if (ANY) { h.t = th; h.prim = t; return h; }
                    if (th < h.t || (th == h.t && t < h.prim)) { h.t = th; h.prim = t; }
                }
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return h;
}

Hit closest_bvh(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    return bvh_query<false>(s, o, d, tmin, tmax);
}
bool any_bvh(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    return bvh_query<true>(s, o, d, tmin, tmax).prim != 0xFFFFFFFFu;
}

// --------------------------------------------------------- path tracing
// sampling.cuh:116-135 — lower-bound binary search on cdf[mid] < u, clamped
// to n-1.
template <typename T>
uint32_t cdf_search(const float* cdf, uint32_t n, T u) {
    uint32_t left = 0, right = n;
    while (left < right) {
        uint32_t mid = (left + right) / 2;
        if (T(cdf[mid]) < u) left = mid + 1; else right = mid;
    }
    return std::min(left, n - 1);
}

template <typename T>
struct Tracer {
    const Scene& s;
    const Camera& cam;
    const Params& p;
    bool useBvh;
    uint32_t sample;
    // per-path flush record
    bool flushed = false;
    V3<T> flushL{T(0), T(0), T(0)};

    Tracer(const Scene& sc, const Camera& c, const Params& pp, bool bvh, uint32_t smp)
        : s(sc), cam(c), p(pp), useBvh(bvh), sample(smp) {}

    static T kPi() { return kPiT<T>(); }

    V3<T> vert(uint32_t i) const {
        return mk<T>(T(s.vertices[i * 3 + 0]), T(s.vertices[i * 3 + 1]), T(s.vertices[i * 3 + 2]));
    }
    V3<T> nrm(uint32_t i) const {
        return mk<T>(T(s.normals[i * 3 + 0]), T(s.normals[i * 3 + 1]), T(s.normals[i * 3 + 2]));
    }

    // Software bilinear RGBA8 fetch, normalized coordinates, wrap addressing,
    // texel centres at integer + 0.5 (replaces CUDA tex2D, scene.cpp:403-410).
    V3<T> sampleTex(const Texture& tex, T u, T v) const {
        if (tex.w == 0 || tex.h == 0) return mk<T>(T(1), T(1), T(1));
        T uf = u - std::floor(u);
        T vf = v - std::floor(v);
        if (!(uf >= T(0) && uf < T(1))) uf = T(0);
        if (!(vf >= T(0) && vf < T(1))) vf = T(0);
        T x = uf * T(tex.w) - T(0.5f);
        T y = vf * T(tex.h) - T(0.5f);
        T x0 = std::floor(x), y0 = std::floor(y);
        T fx = x - x0, fy = y - y0;
        int ix = (int)x0, iy = (int)y0;
        int i0 = ix < 0 ? ix + (int)tex.w : ix;
        int j0 = iy < 0 ? iy + (int)tex.h : iy;
        int i1 = i0 + 1 == (int)tex.w ? 0 : i0 + 1;
        int j1 = j0 + 1 == (int)tex.h ? 0 : j0 + 1;
        auto tx = [&](int i, int j, int c) {
            return T(tex.rgba[((size_t)j * tex.w + i) * 4 + c]) / T(255.0f);
        };
        T out[3];
        for (int c = 0; c < 3; ++c) {
            T top = tx(i0, j0, c) * (T(1) - fx) + tx(i1, j0, c) * fx;
            T bot = tx(i0, j1, c) * (T(1) - fx) + tx(i1, j1, c) * fx;
            out[c] = top * (T(1) - fy) + bot * fy;
        }
        return mk<T>(out[0], out[1], out[2]);
    }

    // materials.cuh:19-30
    V3<T> materialAlbedo(const wrk_material& m, T u, T v) const {
        if (m.albedoTextureId != 0 && m.albedoTextureId <= s.textures.size()) {
            return sampleTex(s.textures[m.albedoTextureId - 1], u, v);
        }
        return mk<T>(T(m.albedo[0]), T(m.albedo[1]), T(m.albedo[2]));
    }

    // sampling.cuh:199-210
    V3<T> clampFireflies(V3<T> L) const {
        T maxL = T(p.fireflyClamp);
        if (maxL <= T(0)) return L;
        T lum = T(0.2126f) * L.x + T(0.7152f) * L.y + T(0.0722f) * L.z;
        if (lum > maxL) {
            T scale = maxL / lum;
            return mk<T>(L.x * scale, L.y * scale, L.z * scale);
        }
        return L;
    }

    // material_common.cuh:33-37 and trace.cu:146-148
    void terminateRay(Ray<T>& ray) {
        flushL = clampFireflies(ray.radiance);
        flushed = true;
        ray.stage = Terminated;
    }

    // sampling.cuh:44-52
    static void frame(V3<T> n, V3<T>& t, V3<T>& b) {
        V3<T> up = std::fabs(n.y) < T(0.999f) ? mk<T>(T(0), T(1), T(0)) : mk<T>(T(1), T(0), T(0));
        t = normalize(cross(up, n));
        b = cross(n, t);
    }
    // sampling.cuh:54-60
    static V3<T> toWorld(V3<T> v, V3<T> n, V3<T> t, V3<T> b) {
        return mk<T>(t.x * v.x + b.x * v.y + n.x * v.z,
                     t.y * v.x + b.y * v.y + n.y * v.z,
                     t.z * v.x + b.z * v.y + n.z * v.z);
    }
    // sampling.cuh:66-75
    static V3<T> sampleCosine(T u1, T u2, T& pdf) {
        T r = sqrtT(u1);
        T theta = (T(2.0f) * kPi()) * u2;
        T sn, cs;
        wr_sincos(theta, &sn, &cs);
        T x = r * cs;
        T y = r * sn;
        T z = sqrtT(fmaxT(T(0), T(1) - u1));
        pdf = z / kPi();
        return mk<T>(x, y, z);
    }
    // sampling.cuh:85-110
    static V3<T> sampleTriangle(T u1, T u2, V3<T> v0, V3<T> v1, V3<T> v2, V3<T>& normal, T& area) {
        T su = sqrtT(u1);
        T b0 = T(1) - su;
        T b1 = su * (T(1) - u2);
        T b2 = su * u2;
        V3<T> pt = mk<T>(v0.x * b0 + v1.x * b1 + v2.x * b2,
                         v0.y * b0 + v1.y * b1 + v2.y * b2,
                         v0.z * b0 + v1.z * b1 + v2.z * b2);
        V3<T> e1 = mk<T>(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
        V3<T> e2 = mk<T>(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
        V3<T> c = cross(e1, e2);
        area = T(0.5f) * length(c);
        normal = normalize(c);
        return pt;
    }
    uint32_t sampleEmissive(T u) const {
        return cdf_search(s.emissiveCdf.data(), (uint32_t)s.emissiveCdf.size(), u);
    }
    // sampling.cuh:145-149
    static T powerHeuristic(T a, T b) {
        T p1 = a * a;
        T p2 = b * b;
        return p1 / (p1 + p2);
    }
    // sampling.cuh:162-193
    bool russianRoulette(V3<T> thr, uint32_t depth, T xi, V3<T>& newThr, bool isDelta) const {
        T rrStart = T(p.rrStartDepth);
        if (T(depth) < rrStart) { newThr = thr; return true; }
        T lum = T(0.2126f) * thr.x + T(0.7152f) * thr.y + T(0.0722f) * thr.z;
        T depthFactor = T(1) / (T(1) + T(0.1f) * (T(depth) - rrStart));
        T minProb = isDelta ? T(0.5f) : T(0.05f);
        T q = fmaxT(minProb, fminT(T(0.95f), lum * depthFactor));
        if (xi < q) { newThr = mk<T>(thr.x / q, thr.y / q, thr.z / q); return true; }
        return false;
    }
    // material_common.cuh:39-52
    bool applyRR(Ray<T>& ray, bool isDelta) {
        V3<T> nt;
        T xi = rnd_dim<T>(ray.seed, ray.rngDimension++);
        if (!russianRoulette(ray.throughput, ray.depth, xi, nt, isDelta)) {
            terminateRay(ray);
            return false;
        }
        ray.throughput = nt;
        return true;
    }
    // utils/ray_offset.cuh:9-21
    static V3<T> offsetOrigin(V3<T> pos, V3<T> n, V3<T> dir) {
        const T off = T(1e-4f);
        V3<T> od = dot(dir, n) > T(0) ? n : -n;
        return mk<T>(pos.x + od.x * off, pos.y + od.y * off, pos.z + od.z * off);
    }
    // material_common.cuh:55-71
    void setupNextBounce(Ray<T>& ray, const HitInfo<T>& hit, V3<T> dir) {
        ray.direction = dir;
        ray.origin = offsetOrigin(hit.position, hit.normal, dir);
        ray.tMin = T(0);
        ray.tMax = T(1e20f);
        ray.depth++;
        ray.stage = Trace;
        if (ray.depth >= p.maxBounces) terminateRay(ray);
    }
    // material_common.cuh:14-29
    static T asymCorrection(V3<T> wi, V3<T> sn, V3<T> gn) {
        T cs = std::fabs(dot(wi, sn));
        T cg = std::fabs(dot(wi, gn));
        if (cs < T(1e-6f) || cg < T(1e-6f)) return T(1);
        return cg / cs;
    }
    // utils/fresnel.cuh:8-16
    static T fresnelDielectric(T cosI, T etaI, T etaT) {
        T sinT2 = etaI / etaT * etaI / etaT * (T(1) - cosI * cosI);
        if (sinT2 > T(1)) return T(1);
        T cosT = sqrtT(T(1) - sinT2);
        T rs = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
        T rp = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
        return (rs * rs + rp * rp) * T(0.5f);
    }
    // ggx.cuh:18-23
    static T ggxD(T NoH, T alpha) {
        T a2 = alpha * alpha;
        T NoH2 = NoH * NoH;
        T denom = NoH2 * (a2 - T(1)) + T(1);
        return a2 / (kPi() * denom * denom);
    }
    // ggx.cuh:26-30
    static T ggxG1(T NoV, T alpha) {
        T a2 = alpha * alpha;
        T NoV2 = NoV * NoV;
        return T(2) * NoV / (NoV + sqrtT(a2 + (T(1) - a2) * NoV2));
    }
    // ggx.cuh:113-176 (with ggx.cuh:33-63)
    static V3<T> evalGGX(V3<T> wo, V3<T> wi, V3<T> n, V3<T> albedo, T rough, T metal, T ior) {
        T NoV = fmaxT(T(0), dot(n, wo));
        T NoL = fmaxT(T(0), dot(n, wi));
        if (NoV < T(1e-5f) || NoL < T(1e-5f)) return mk<T>(T(0), T(0), T(0));
        V3<T> h = normalize(mk<T>(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z));
        T NoH = fmaxT(T(0), dot(n, h));
        T VoH = fmaxT(T(0), dot(wo, h));
        T alpha = rough * rough;
        T D = ggxD(NoH, alpha);
        T G = ggxG1(NoV, alpha) * ggxG1(NoL, alpha);
        T f0 = (ior - T(1)) / (ior + T(1));
        f0 = f0 * f0;
        V3<T> F0 = mk<T>(f0 * (T(1) - metal) + albedo.x * metal,
                         f0 * (T(1) - metal) + albedo.y * metal,
                         f0 * (T(1) - metal) + albedo.z * metal);
        T f = pow5(T(1) - VoH);
        V3<T> F = mk<T>(F0.x + (T(1) - F0.x) * f, F0.y + (T(1) - F0.y) * f, F0.z + (T(1) - F0.z) * f);
        V3<T> spec = mk<T>(D * G * F.x / (T(4) * NoV * NoL),
                           D * G * F.y / (T(4) * NoV * NoL),
                           D * G * F.z / (T(4) * NoV * NoL));
        V3<T> kD = mk<T>((T(1) - F.x) * (T(1) - metal), (T(1) - F.y) * (T(1) - metal), (T(1) - F.z) * (T(1) - metal));
        V3<T> diff = mk<T>(kD.x * albedo.x / kPi(), kD.y * albedo.y / kPi(), kD.z * albedo.z / kPi());
        return mk<T>(diff.x + spec.x, diff.y + spec.y, diff.z + spec.z);
    }
    // ggx.cuh:182-215
    static T ggxPdf(V3<T> wo, V3<T> wi, V3<T> n, T rough) {
        T NoV = fmaxT(T(0), dot(n, wo));
        T NoL = fmaxT(T(0), dot(n, wi));
        if (NoV < T(1e-5f) || NoL < T(1e-5f)) return T(0);
        V3<T> h = normalize(mk<T>(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z));
        T NoH = fmaxT(T(0), dot(n, h));
        T VoH = fmaxT(T(0), dot(wo, h));
        if (VoH < T(1e-5f)) return T(0);
        T alpha = rough * rough;
        T G1 = ggxG1(NoV, alpha);
        T D = ggxD(NoH, alpha);
        return G1 * D / (T(4) * NoV);
    }
    // ggx.cuh:222-317; returns false where the reference sets pdf = 0 and
    // leaves wi unwritten.
    static void sampleGGX(V3<T> wo, V3<T> n, V3<T> tg, V3<T> bt, T rough, T u0, T u1, V3<T>& wi, T& pdf, bool& wiSet) {
        T alpha = rough * rough;
        V3<T> v = mk<T>(dot(wo, tg), dot(wo, bt), dot(wo, n));
        V3<T> sv = normalize(mk<T>(alpha * v.x, alpha * v.y, v.z));
        T dist2D = sqrtT(sv.x * sv.x + sv.y * sv.y);
        V3<T> T1, T2;
        if (dist2D > T(1e-6f)) {
            T rec = T(1) / dist2D;
            T1 = mk<T>(sv.y * rec, -sv.x * rec, T(0));
            T2 = mk<T>(T1.y * sv.z, -T1.x * sv.z, dist2D);
        } else {
            T1 = mk<T>(T(1), T(0), T(0));
            T2 = mk<T>(T(0), T(1), T(0));
        }
        T a = T(1) / (T(1) + sv.z);
        T r = sqrtT(u0);
        T phi = kPi() * ((u1 < a) ? u1 / a : T(1) + (u1 - a) / (T(1) - a));
        T sinPhi, cosPhi;
        wr_sincos(phi, &sinPhi, &cosPhi);
        T P1 = r * cosPhi;
        T P2 = r * sinPhi * ((u1 < a) ? T(1) : sv.z);
        T pp = P1 * P1 + P2 * P2;
        T sq = sqrtT(fmaxT(T(0), T(1) - pp));
        V3<T> mr = mk<T>(P1 * T1.x + P2 * T2.x + sq * sv.x,
                         P1 * T1.y + P2 * T2.y + sq * sv.y,
                         P1 * T1.z + P2 * T2.z + sq * sv.z);
        mr = normalize(mk<T>(alpha * mr.x, alpha * mr.y, mr.z));
        V3<T> h = mk<T>(tg.x * mr.x + bt.x * mr.y + n.x * mr.z,
                        tg.y * mr.x + bt.y * mr.y + n.y * mr.z,
                        tg.z * mr.x + bt.z * mr.y + n.z * mr.z);
        T VoH = dot(wo, h);
        wiSet = false;
        if (VoH <= T(0)) { pdf = T(0); return; }
        wi = mk<T>(T(2) * VoH * h.x - wo.x, T(2) * VoH * h.y - wo.y, T(2) * VoH * h.z - wo.z);
        wiSet = true;
        T NoL = dot(n, wi);
        if (NoL <= T(0)) { pdf = T(0); return; }
        T NoV = fmaxT(T(0), dot(n, wo));
        T NoH = fmaxT(T(0), dot(n, h));
        T G1 = ggxG1(NoV, alpha);
        T D = ggxD(NoH, alpha);
        pdf = G1 * D / (T(4) * NoV);
    }

    // NEE light sample shared by Lambertian and GGX
    // (lambertian.cuh:17-50, ggx_material.cuh:33-94).
    struct LightSample { V3<T> toLight; T distSq, dist, cosTheta, cosLight, lightPdf; V3<T> Le; };
    bool sampleLight(Ray<T>& ray, const HitInfo<T>& hit, LightSample& ls) {
        T lightU = rnd_dim<T>(ray.seed, ray.rngDimension++);
        uint32_t li = sampleEmissive(lightU);
        uint32_t tri = s.emissiveTris[li];
        uint32_t i0 = s.indices[tri * 3 + 0], i1 = s.indices[tri * 3 + 1], i2 = s.indices[tri * 3 + 2];
        V3<T> v0 = vert(i0), v1 = vert(i1), v2 = vert(i2);
        T u1 = rnd_dim<T>(ray.seed, ray.rngDimension);
        T u2 = rnd_dim<T>(ray.seed, ray.rngDimension + 1);
        ray.rngDimension += 2;
        V3<T> ln; T area;
        V3<T> lp = sampleTriangle(u1, u2, v0, v1, v2, ln, area);
        V3<T> tl = mk<T>(lp.x - hit.position.x, lp.y - hit.position.y, lp.z - hit.position.z);
        ls.distSq = dot(tl, tl);
        ls.dist = sqrtT(ls.distSq);
        ls.toLight = mk<T>(tl.x / ls.dist, tl.y / ls.dist, tl.z / ls.dist);
        ls.cosTheta = dot(hit.normal, ls.toLight);
        ls.cosLight = -dot(ln, ls.toLight);
        if (!(ls.cosTheta > T(0) && ls.cosLight > T(0))) return false;
        const wrk_material& lm = s.materials[s.triMat[tri]];
        T pdfArea = T(1) / (T(float(s.emissiveTris.size())) * area);
        ls.lightPdf = pdfArea * ls.distSq / fmaxT(ls.cosLight, T(1e-8f));
        ls.Le = mk<T>(T(lm.emission[0]), T(lm.emission[1]), T(lm.emission[2]));
        return true;
    }
    static bool validContrib(V3<T> c) {
        return !std::isnan(c.x) && !std::isnan(c.y) && !std::isnan(c.z) &&
               !std::isinf(c.x) && !std::isinf(c.y) && !std::isinf(c.z) &&
               c.x >= T(0) && c.y >= T(0) && c.z >= T(0);
    }

    // materials/lambertian.cuh:8-147
    void shadeLambertian(Ray<T>& ray, const HitInfo<T>& hit, const wrk_material& mat) {
        if (p.useNEE && !s.emissiveTris.empty() && !ray.neeDone) {
            LightSample ls;
            if (sampleLight(ray, hit, ls)) {
                T bsdfPdf = fmaxT(T(0), ls.cosTheta / kPi());
                T w = powerHeuristic(ls.lightPdf, bsdfPdf);
                V3<T> albedo = materialAlbedo(mat, hit.u, hit.v);
                V3<T> bsdf = mk<T>(albedo.x / kPi() * ls.cosTheta, albedo.y / kPi() * ls.cosTheta,
                                   albedo.z / kPi() * ls.cosTheta);
                V3<T> c = mk<T>(ray.throughput.x * bsdf.x * ls.Le.x * w / ls.lightPdf,
                                ray.throughput.y * bsdf.y * ls.Le.y * w / ls.lightPdf,
                                ray.throughput.z * bsdf.z * ls.Le.z * w / ls.lightPdf);
                if (validContrib(c)) {
                    ray.shadowContribution = c;
                    ray.savedDirection = ray.direction;
                    ray.direction = ls.toLight;
                    ray.origin = offsetOrigin(hit.position, hit.normal, ls.toLight);
                    ray.tMin = T(1e-5f);
                    ray.tMax = ls.dist - T(1e-5f);
                    ray.neeDone = 1;
                    ray.stage = Shadow;
                    return;
                }
            }
        }
        ray.neeDone = 0;
        V3<T> tg, bt;
        frame(hit.normal, tg, bt);
        T u1 = rnd_dim<T>(ray.seed, ray.rngDimension);
        T u2 = rnd_dim<T>(ray.seed, ray.rngDimension + 1);
        ray.rngDimension += 2;
        T pdf;
        V3<T> local = sampleCosine(u1, u2, pdf);
        V3<T> wd = toWorld(local, hit.normal, tg, bt);
        V3<T> albedo = materialAlbedo(mat, hit.u, hit.v);
        ray.throughput = ray.throughput * albedo;
        ray.prevPdf = pdf;
        ray.prevWasDelta = false;
        if (!applyRR(ray, false)) return;
        setupNextBounce(ray, hit, wd);
    }

    // materials/glass.cuh:10-140 and materials/ggx_transmission.cuh:12-107
    // (identical arithmetic and RNG use).
    void shadeGlass(Ray<T>& ray, const HitInfo<T>& hit, const wrk_material& mat) {
        V3<T> wo = -ray.direction;
        T cosI = dot(wo, hit.normal);
        T etaI = T(1), etaT = T(mat.ior);
        V3<T> n = hit.normal;
        if (cosI < T(0)) { cosI = -cosI; n = -n; etaI = T(mat.ior); etaT = T(1); }
        T F = fresnelDielectric(cosI, etaI, etaT);
        T r = rnd_dim<T>(ray.seed, ray.rngDimension++);
        if (r < F) {
            T dn = T(2) * dot(ray.direction, n);
            ray.direction = mk<T>(ray.direction.x - n.x * dn, ray.direction.y - n.y * dn, ray.direction.z - n.z * dn);
            ray.throughput = ray.throughput * materialAlbedo(mat, hit.u, hit.v);
        } else {
            T eta = etaI / etaT;
            T sinT2 = eta * eta * (T(1) - cosI * cosI);
            if (sinT2 >= T(1)) {
                T dn = T(2) * dot(ray.direction, n);
                ray.direction = mk<T>(ray.direction.x - n.x * dn, ray.direction.y - n.y * dn, ray.direction.z - n.z * dn);
                ray.throughput = ray.throughput * materialAlbedo(mat, hit.u, hit.v);
            } else {
                T cosT = sqrtT(fmaxT(T(0), T(1) - sinT2));
                T k = eta * cosI - cosT;
                V3<T> rf = mk<T>(eta * ray.direction.x + n.x * k, eta * ray.direction.y + n.y * k,
                                 eta * ray.direction.z + n.z * k);
                ray.direction = rf;
                ray.throughput = ray.throughput * materialAlbedo(mat, hit.u, hit.v);
                T corr = asymCorrection(rf, hit.normal, hit.geometricNormal);
                ray.throughput = ray.throughput * corr;
            }
        }
        ray.prevWasDelta = true;
        if (!applyRR(ray, true)) return;
        setupNextBounce(ray, hit, ray.direction);
    }

    // materials/ggx_material.cuh:10-213
    void shadeGGX(Ray<T>& ray, const HitInfo<T>& hit, const wrk_material& mat) {
        V3<T> wo = -ray.direction;
        V3<T> albedo = mk<T>(T(mat.albedo[0]), T(mat.albedo[1]), T(mat.albedo[2]));
        T rough = T(mat.roughness), metal = T(mat.metallic), ior = T(mat.ior);
        if (p.useNEE && !s.emissiveTris.empty()) {
            LightSample ls;
            if (sampleLight(ray, hit, ls)) {
                V3<T> brdf = evalGGX(wo, ls.toLight, hit.normal, albedo, rough, metal, ior);
                T bsdfPdf = ggxPdf(wo, ls.toLight, hit.normal, rough);
                T w = powerHeuristic(ls.lightPdf, bsdfPdf);
                V3<T> c = mk<T>(ray.throughput.x * brdf.x * ls.Le.x * ls.cosTheta * w / ls.lightPdf,
                                ray.throughput.y * brdf.y * ls.Le.y * ls.cosTheta * w / ls.lightPdf,
                                ray.throughput.z * brdf.z * ls.Le.z * ls.cosTheta * w / ls.lightPdf);
                if (validContrib(c)) {
                    ray.radiance = mk<T>(ray.radiance.x + c.x, ray.radiance.y + c.y, ray.radiance.z + c.z);
                }
            }
        }
        V3<T> tg, bt;
        frame(hit.normal, tg, bt);
        T u0 = rnd_dim<T>(ray.seed, ray.rngDimension);
        T u1 = rnd_dim<T>(ray.seed, ray.rngDimension + 1);
        ray.rngDimension += 2;
        V3<T> wi = mk<T>(T(0), T(0), T(0));
        T pdf; bool wiSet;
        sampleGGX(wo, hit.normal, tg, bt, rough, u0, u1, wi, pdf, wiSet);
        T cosTheta = dot(hit.normal, wi);
        if (pdf < T(1e-5f) || cosTheta <= T(0)) { terminateRay(ray); return; }
        V3<T> brdf = evalGGX(wo, wi, hit.normal, albedo, rough, metal, ior);
        if (std::isnan(brdf.x) || std::isnan(brdf.y) || std::isnan(brdf.z) ||
            std::isinf(brdf.x) || std::isinf(brdf.y) || std::isinf(brdf.z)) {
            terminateRay(ray);
            return;
        }
        ray.throughput = mk<T>(ray.throughput.x * brdf.x * cosTheta / pdf,
                               ray.throughput.y * brdf.y * cosTheta / pdf,
                               ray.throughput.z * brdf.z * cosTheta / pdf);
        ray.prevPdf = pdf;
        ray.prevWasDelta = false;
        if (!applyRR(ray, false)) return;
        setupNextBounce(ray, hit, wi);
    }

    // trace.cu:53-125 (camera ray, depth 0)
    void cameraRay(Ray<T>& ray, uint32_t pixel) {
        uint32_t px = pixel % p.width;
        uint32_t py = pixel / p.width;
        uint32_t seed = camera_seed(pixel, sample);
        uint32_t js = seed;
        T r1 = randf<T>(js);
        T r2 = randf<T>(js);
        T ndcX = (T(2) * (T(float(px)) + r1) / T(float(p.width)) - T(1)) * T(cam.aspect);
        T ndcY = T(1) - T(2) * (T(float(py)) + r2) / T(float(p.height));
        V3<T> P = mk<T>(T(cam.position[0]), T(cam.position[1]), T(cam.position[2]));
        V3<T> Fw = mk<T>(T(cam.forward[0]), T(cam.forward[1]), T(cam.forward[2]));
        V3<T> R = mk<T>(T(cam.right[0]), T(cam.right[1]), T(cam.right[2]));
        V3<T> U = mk<T>(T(cam.up[0]), T(cam.up[1]), T(cam.up[2]));
        T th = T(cam.tanHalfFovY);
        V3<T> d = Fw + R * ndcX * th + U * ndcY * th;
        d = normalize(d);
        if (T(cam.lensRadius) > T(0) && T(cam.focalDistance) > T(0)) {
            T r3 = randf<T>(js);
            T r4 = randf<T>(js);
            T theta = (T(2.0f) * kPi()) * r3;
            T radius = sqrtT(r4) * T(cam.lensRadius);
            T sn, cs;
            wr_sincos(theta, &sn, &cs);
            V3<T> lo = mk<T>(R.x * radius * cs + U.x * radius * sn,
                             R.y * radius * cs + U.y * radius * sn,
                             R.z * radius * cs + U.z * radius * sn);
            T fd = T(cam.focalDistance);
            V3<T> fp = mk<T>(P.x + d.x * fd, P.y + d.y * fd, P.z + d.z * fd);
            ray.origin = mk<T>(P.x + lo.x, P.y + lo.y, P.z + lo.z);
            ray.direction = normalize(mk<T>(fp.x - ray.origin.x, fp.y - ray.origin.y, fp.z - ray.origin.z));
        } else {
            ray.origin = P;
            ray.direction = d;
        }
        ray.throughput = mk<T>(T(1), T(1), T(1));
        ray.radiance = mk<T>(T(0), T(0), T(0));
        ray.pixelIndex = pixel;
        ray.depth = 0;
        ray.stage = Trace;
        ray.seed = seed;
        ray.tMin = T(0.0001f);
        ray.tMax = T(1e20f);
        ray.isFirstHit = true;
        ray.prevPdf = T(1);
        ray.prevWasDelta = true;
        ray.rngDimension = 0;
        ray.neeDone = 0;
    }

    Hit traceClosest(const Ray<T>& ray) const {
        float o[3] = {float(ray.origin.x), float(ray.origin.y), float(ray.origin.z)};
        float d[3] = {float(ray.direction.x), float(ray.direction.y), float(ray.direction.z)};
        return useBvh ? closest_bvh(s, o, d, float(ray.tMin), float(ray.tMax))
                      : closest_brute(s, o, d, float(ray.tMin), float(ray.tMax));
    }
    bool traceAny(const Ray<T>& ray) const {
        float o[3] = {float(ray.origin.x), float(ray.origin.y), float(ray.origin.z)};
        float d[3] = {float(ray.direction.x), float(ray.direction.y), float(ray.direction.z)};
        return useBvh ? any_bvh(s, o, d, float(ray.tMin), float(ray.tMax))
                      : any_brute(s, o, d, float(ray.tMin), float(ray.tMax));
    }

    // trace.cu:155-319 (closest-hit program)
    void closestHit(Ray<T>& ray, HitInfo<T>& hit, T t, uint32_t prim) {
        V3<T> o = ray.origin, d = ray.direction;
        V3<T> hitPos = o + t * d;
        uint32_t i0 = s.indices[prim * 3 + 0], i1 = s.indices[prim * 3 + 1], i2 = s.indices[prim * 3 + 2];
        V3<T> v0 = vert(i0), v1 = vert(i1), v2 = vert(i2);
        V3<T> e1 = v1 - v0, e2 = v2 - v0;
        V3<T> gN = normalize(cross(e1, e2));
        V3<T> v0p = hitPos - v0;
        T d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
        T d20 = dot(v0p, e1), d21 = dot(v0p, e2);
        T denom = d00 * d11 - d01 * d01;
        T vb = (d11 * d20 - d01 * d21) / fmaxT(denom, T(1e-8f));
        T wb = (d00 * d21 - d01 * d20) / fmaxT(denom, T(1e-8f));
        T ub = T(1) - vb - wb;
        V3<T> sN = gN;
        if (s.hasNormals) {
            V3<T> n0 = nrm(i0), n1 = nrm(i1), n2 = nrm(i2);
            T l0 = length(n0), l1 = length(n1), l2 = length(n2);
            if (l0 > T(1e-6f) && l1 > T(1e-6f) && l2 > T(1e-6f)) {
                sN = normalize(ub * n0 + vb * n1 + wb * n2);
                if (dot(sN, gN) < T(0)) sN = -sN;
            }
        }
        V3<T> normal = dot(gN, d) > T(0) ? -gN : gN;
        V3<T> sNf = dot(sN, d) > T(0) ? -sN : sN;
        hit.position = hitPos;
        hit.normal = sNf;
        hit.geometricNormal = normal;
        hit.materialId = s.triMat[prim];
        hit.primIndex = prim;
        if (s.hasUVs) {
            T a0 = T(s.uvs[i0 * 2 + 0]), b0 = T(s.uvs[i0 * 2 + 1]);
            T a1 = T(s.uvs[i1 * 2 + 0]), b1 = T(s.uvs[i1 * 2 + 1]);
            T a2 = T(s.uvs[i2 * 2 + 0]), b2 = T(s.uvs[i2 * 2 + 1]);
            hit.u = ub * a0 + vb * a1 + wb * a2;
            hit.v = ub * b0 + vb * b1 + wb * b2;
        } else {
            hit.u = T(0); hit.v = T(0);
        }
        // first-hit denoiser guides (trace.cu:267-286) do not affect radiance
        ray.isFirstHit = false;
        if (hit.materialId < s.materials.size()) {
            const wrk_material& mat = s.materials[hit.materialId];
            if (mat.type == WRK_MAT_EMISSIVE) {
                T w = T(1);
                if (p.useNEE && ray.depth > 0 && !ray.prevWasDelta && !s.emissiveTris.empty()) {
                    T area = T(0.5f) * length(cross(e1, e2));
                    T distSq = t * t;
                    T cosL = std::fabs(dot(normal, d));
                    T pdfArea = T(1) / (T(float(s.emissiveTris.size())) * area);
                    T lightPdf = pdfArea * distSq / fmaxT(cosL, T(1e-8f));
                    w = powerHeuristic(ray.prevPdf, lightPdf);
                }
                V3<T> E = mk<T>(T(mat.emission[0]), T(mat.emission[1]), T(mat.emission[2]));
                ray.radiance = ray.radiance + ray.throughput * E * w;
            }
        }
    }

    // utils/normal_map.cuh:12-60
    V3<T> effectiveNormal(const wrk_material& mat, const HitInfo<T>& hit) const {
        if (mat.normalTextureId > 0 && mat.normalTextureId <= s.textures.size()) {
            const Texture& tex = s.textures[mat.normalTextureId - 1];
            if (tex.w == 0) return hit.normal;
            V3<T> tn = sampleTex(tex, hit.u, hit.v);
            V3<T> tg, bt;
            frame(hit.normal, tg, bt);
            V3<T> n = mk<T>(tn.x * T(2) - T(1), tn.y * T(2) - T(1), tn.z * T(2) - T(1));
            V3<T> w = mk<T>(tg.x * n.x + bt.x * n.y + hit.normal.x * n.z,
                            tg.y * n.x + bt.y * n.y + hit.normal.y * n.z,
                            tg.z * n.x + bt.z * n.y + hit.normal.z * n.z);
            return normalize(w);
        }
        return hit.normal;
    }

    // shade.cu:24-85
    void shade(Ray<T>& ray, const HitInfo<T>& hit) {
        if (hit.materialId >= s.materials.size()) { ray.stage = Terminated; return; }
        const wrk_material& mat = s.materials[hit.materialId];
        HitInfo<T> eh = hit;
        if (mat.normalTextureId > 0) eh.normal = effectiveNormal(mat, hit);
        switch (mat.type) {
            case WRK_MAT_EMISSIVE: terminateRay(ray); break;
            case WRK_MAT_LAMBERTIAN: shadeLambertian(ray, eh, mat); break;
            case WRK_MAT_GLASS:
            case WRK_MAT_GGX_TRANSMISSION: shadeGlass(ray, eh, mat); break;
            case WRK_MAT_GGX_REFLECTION: shadeGGX(ray, eh, mat); break;
            default: terminateRay(ray); break;
        }
    }
};

// One (pixel, sample) path through the reference host loop
// (renderer.cpp:190-284): every iteration runs raygen/trace (trace.cu:17-153)
// then shade (shade.cu:24-85) on the still-active path; the loop stops after
// iterationCap iterations, dropping paths that are still alive.
template <typename T>
void trace_path(const Scene& s, const Camera& cam, const Params& p, uint32_t pixel,
                uint32_t sample, bool useBvh, PathOut* out, double* outL64) {
    Tracer<T> tr(s, cam, p, useBvh, sample);
    Ray<T> ray;
    std::memset(&ray, 0, sizeof(ray));
    HitInfo<T> hit;
    std::memset(&hit, 0, sizeof(hit));
    const uint32_t cap = p.iterationCap ? p.iterationCap : 0xFFFFFFFFu;
    uint32_t it = 0;
    for (; it < cap && ray.stage != Terminated; ++it) {
        if (ray.stage == Shadow) {
            if (!tr.traceAny(ray)) ray.radiance = ray.radiance + ray.shadowContribution;
            ray.direction = ray.savedDirection;
            ray.stage = Shade;
        } else if (ray.stage == Trace) {
            if (ray.depth == 0) tr.cameraRay(ray, pixel);
            Hit h = tr.traceClosest(ray);
            if (h.prim == 0xFFFFFFFFu) {
                V3<T> env = mk<T>(T(s.env[0]), T(s.env[1]), T(s.env[2]));
                ray.radiance = ray.radiance + ray.throughput * env;
                tr.terminateRay(ray);
            } else {
                tr.closestHit(ray, hit, T(h.t), h.prim);
                ray.stage = Shade;
            }
        }
        if (ray.stage == Shade) tr.shade(ray, hit);
    }
    out->L[0] = tr.flushed ? float(tr.flushL.x) : 0.0f;
    out->L[1] = tr.flushed ? float(tr.flushL.y) : 0.0f;
    out->L[2] = tr.flushed ? float(tr.flushL.z) : 0.0f;
    if (outL64) {
        outL64[0] = tr.flushed ? double(tr.flushL.x) : 0.0;
        outL64[1] = tr.flushed ? double(tr.flushL.y) : 0.0;
        outL64[2] = tr.flushed ? double(tr.flushL.z) : 0.0;
    }
    out->iterations = it;
    out->flushed = tr.flushed ? 1u : 0u;
    out->depth = ray.depth;
    out->rngDim = ray.rngDimension;
}

template void trace_path<float>(const Scene&, const Camera&, const Params&, uint32_t, uint32_t, bool, PathOut*, double*);

void first_hit_guides(const Scene& s, const Camera& cam, const Params& p, uint32_t pixel, uint32_t sample,
                      bool useBvh, float albedo[3], float normal[3]) {
    Tracer<float> tr(s, cam, p, useBvh, sample);
    Ray<float> ray;
    std::memset(&ray, 0, sizeof(ray));
    HitInfo<float> hit;
    std::memset(&hit, 0, sizeof(hit));
    albedo[0] = albedo[1] = albedo[2] = 0.0f;
    normal[0] = normal[1] = normal[2] = 0.0f;
    tr.cameraRay(ray, pixel);
    Hit h = tr.traceClosest(ray);
    if (h.prim == 0xFFFFFFFFu) return;
    tr.closestHit(ray, hit, h.t, h.prim);
    if (hit.materialId >= s.materials.size()) return;
    V3<float> a = tr.materialAlbedo(s.materials[hit.materialId], hit.u, hit.v);
    albedo[0] = a.x; albedo[1] = a.y; albedo[2] = a.z;
    normal[0] = hit.geometricNormal.x; normal[1] = hit.geometricNormal.y; normal[2] = hit.geometricNormal.z;
}
template void trace_path<double>(const Scene&, const Camera&, const Params&, uint32_t, uint32_t, bool, PathOut*, double*);

// ------------------------------------------------ single-step test hooks
static V3<float> v3(const float* a) { return mk<float>(a[0], a[1], a[2]); }
static void put(float* o, V3<float> v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

void shade_event(const Scene& s, const Params& p, EventIO* e) {
    static const Camera noCam{};
    Tracer<float> tr(s, noCam, p, false, 0);
    Ray<float> ray;
    std::memset(&ray, 0, sizeof(ray));
    ray.origin = v3(e->origin); ray.direction = v3(e->direction);
    ray.throughput = v3(e->throughput); ray.radiance = v3(e->radiance);
    ray.tMin = e->tMin; ray.tMax = e->tMax; ray.prevPdf = e->prevPdf;
    ray.seed = e->seed; ray.rngDimension = e->rngDimension; ray.depth = e->depth;
    ray.stage = Shade; ray.neeDone = e->neeDone; ray.prevWasDelta = e->prevWasDelta != 0;
    ray.shadowContribution = v3(e->shadowContribution); ray.savedDirection = v3(e->savedDirection);
    HitInfo<float> hit;
    std::memset(&hit, 0, sizeof(hit));
    hit.position = v3(e->position); hit.normal = v3(e->normal); hit.geometricNormal = v3(e->geometricNormal);
    hit.u = e->u; hit.v = e->v; hit.materialId = e->materialId;
    tr.shade(ray, hit);
    put(e->origin, ray.origin); put(e->direction, ray.direction);
    put(e->throughput, ray.throughput); put(e->radiance, ray.radiance);
    e->tMin = ray.tMin; e->tMax = ray.tMax; e->prevPdf = ray.prevPdf;
    e->seed = ray.seed; e->rngDimension = ray.rngDimension; e->depth = ray.depth; e->stage = ray.stage;
    e->neeDone = ray.neeDone; e->prevWasDelta = ray.prevWasDelta ? 1u : 0u;
    put(e->shadowContribution, ray.shadowContribution); put(e->savedDirection, ray.savedDirection);
    e->flushed = tr.flushed ? 1u : 0u;
    put(e->flushL, tr.flushed ? tr.flushL : mk<float>(0.0f, 0.0f, 0.0f));
}

void hit_attributes(const Scene& s, const float o[3], const float d[3], float t, uint32_t prim, float out[11]) {
    static const Camera noCam{};
    static const Params noParams{1, 1, 0, 1, 0.0f, 0.0f, 0};
    Tracer<float> tr(s, noCam, noParams, false, 0);
    Ray<float> ray;
    std::memset(&ray, 0, sizeof(ray));
    ray.origin = v3(o); ray.direction = v3(d);
    HitInfo<float> hit;
    std::memset(&hit, 0, sizeof(hit));
    tr.closestHit(ray, hit, t, prim);
    put(out, hit.position); put(out + 3, hit.normal); put(out + 6, hit.geometricNormal);
    out[9] = hit.u; out[10] = hit.v;
}

void camera_ray(const Camera& cam, const Params& p, uint32_t pixel, uint32_t sample, float origin[3],
                float dir[3]) {
    static const Scene noScene{};
    Tracer<float> tr(noScene, cam, p, false, sample);
    Ray<float> ray;
    std::memset(&ray, 0, sizeof(ray));
    tr.cameraRay(ray, pixel);
    put(origin, ray.origin); put(dir, ray.direction);
}

void sample_texture(const Scene& s, uint32_t texId, float u, float v, float out[3]) {
    static const Camera noCam{};
    static const Params noParams{1, 1, 0, 1, 0.0f, 0.0f, 0};
    Tracer<float> tr(s, noCam, noParams, false, 0);
    wrk_material m{};
    m.albedo[0] = m.albedo[1] = m.albedo[2] = -1.0f;  // marks "constant albedo taken"
    m.albedoTextureId = texId;
    put(out, tr.materialAlbedo(m, u, v));
}

void effective_normal(const Scene& s, uint32_t mat, const float n[3], float u, float v, float out[3]) {
    static const Camera noCam{};
    static const Params noParams{1, 1, 0, 1, 0.0f, 0.0f, 0};
    Tracer<float> tr(s, noCam, noParams, false, 0);
    HitInfo<float> hit;
    std::memset(&hit, 0, sizeof(hit));
    hit.normal = v3(n); hit.u = u; hit.v = v;
    put(out, tr.effectiveNormal(s.materials[mat], hit));
}

}  // namespace orc

// ------------------------------------------------ math exports for tests
// Thin wrappers over the FP32 restatement so tests/ can pin each function
// against closed forms and invariants.
extern "C" {
using orc::V3;
using TF = orc::Tracer<float>;
float orc_fresnel(float cosI, float etaI, float etaT) { return TF::fresnelDielectric(cosI, etaI, etaT); }
float orc_ggx_D(float NoH, float alpha) { return TF::ggxD(NoH, alpha); }
float orc_ggx_G1(float NoV, float alpha) { return TF::ggxG1(NoV, alpha); }
void orc_ggx_eval(const float* wo, const float* wi, const float* n, const float* albedo, float rough, float metal,
                  float ior, float* out) {
    V3<float> r = TF::evalGGX(orc::mk(wo[0], wo[1], wo[2]), orc::mk(wi[0], wi[1], wi[2]), orc::mk(n[0], n[1], n[2]),
                              orc::mk(albedo[0], albedo[1], albedo[2]), rough, metal, ior);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float orc_ggx_pdf(const float* wo, const float* wi, const float* n, float rough) {
    return TF::ggxPdf(orc::mk(wo[0], wo[1], wo[2]), orc::mk(wi[0], wi[1], wi[2]), orc::mk(n[0], n[1], n[2]), rough);
}
// wi (3 floats) and pdf; returns 1 when wi was written (reference leaves it unset on failure)
int orc_ggx_sample(const float* wo, const float* n, float rough, float u0, float u1, float* wi, float* pdf) {
    V3<float> N = orc::mk(n[0], n[1], n[2]), tg, bt;
    TF::frame(N, tg, bt);
    V3<float> w = orc::mk(0.0f, 0.0f, 0.0f);
    bool set = false;
    TF::sampleGGX(orc::mk(wo[0], wo[1], wo[2]), N, tg, bt, rough, u0, u1, w, *pdf, set);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    return set ? 1 : 0;
}
void orc_sample_cosine(float u1, float u2, float* dir, float* pdf) {
    V3<float> d = TF::sampleCosine(u1, u2, *pdf);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}
void orc_frame(const float* n, float* t, float* b) {
    V3<float> T, B;
    TF::frame(orc::mk(n[0], n[1], n[2]), T, B);
    t[0] = T.x; t[1] = T.y; t[2] = T.z; b[0] = B.x; b[1] = B.y; b[2] = B.z;
}
float orc_power_heuristic(float a, float b) { return TF::powerHeuristic(a, b); }
uint32_t orc_cdf_search(const float* cdf, uint32_t n, float u) { return orc::cdf_search<float>(cdf, n, u); }
}  // extern "C"

// ------------------------------------------- more single-function exports
extern "C" {
using orc::V3;
// sampling.cuh:85-110: point (3), normal (3), area (1)
void orc_sample_triangle(float u1, float u2, const float* v0, const float* v1, const float* v2, float* pt,
                         float* normal, float* area) {
    V3<float> n;
    V3<float> p = orc::Tracer<float>::sampleTriangle(u1, u2, orc::mk(v0[0], v0[1], v0[2]),
                                                      orc::mk(v1[0], v1[1], v1[2]),
                                                      orc::mk(v2[0], v2[1], v2[2]), n, *area);
    pt[0] = p.x; pt[1] = p.y; pt[2] = p.z;
    normal[0] = n.x; normal[1] = n.y; normal[2] = n.z;
}
// utils/ray_offset.cuh:9-21
void orc_offset_origin(const float* pos, const float* n, const float* dir, float* out) {
    V3<float> o = orc::Tracer<float>::offsetOrigin(orc::mk(pos[0], pos[1], pos[2]), orc::mk(n[0], n[1], n[2]),
                                                   orc::mk(dir[0], dir[1], dir[2]));
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
// material_common.cuh:14-29
float orc_asym_correction(const float* wi, const float* sn, const float* gn) {
    return orc::Tracer<float>::asymCorrection(orc::mk(wi[0], wi[1], wi[2]), orc::mk(sn[0], sn[1], sn[2]),
                                              orc::mk(gn[0], gn[1], gn[2]));
}
// Watertight test of one triangle; returns 1 and t on a hit in [tmin, tmax].
int orc_intersect_triangle(const float* o, const float* d, const float* v0, const float* v1, const float* v2,
                           float tmin, float tmax, float* t) {
    orc::RayPre r = orc::ray_pre(o, d);
    return orc::intersect_tri(r, v0, v1, v2, tmin, tmax, t) ? 1 : 0;
}
}  // extern "C"
