"""Seeded synthetic scenes for the benchmark configurations.

A scene is a list of builder operations (materials, textures, meshes,
environment) replayed in order onto any object with the wr::Scene method
names — the product Scene or the oracle's — so both see identical inputs.

  cornell()          Cornell box: 5 walls (10 tris), 2 boxes (24 tris), a
                     0.5x0.5 ceiling light (2 emissive tris, radiance
                     (17,12,4)); palette and walls of
                     libVLRM/test/cornell_box_test.cpp:35-102 with the light
                     wound to face the room as in
                     libWR/test/cornell_box_var_test.cpp:65-70. 36 triangles.
  material_scene()   ground plane + 25 subdivision-4 icospheres on a seeded
                     5x5 grid, materials cycling over textured Lambertian,
                     glass, GGX reflection, GGX transmission and normal-mapped
                     Lambertian; 4 emissive quads, environment 0.1.
  large_scene()      seeded displaced icospheres (5,120 tris each) in a
                     100^3 box, uniform material kinds, 64 emissive
                     triangles with power-law radiance.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np


@dataclass
class SceneDesc:
    ops: list = field(default_factory=list)
    camera: dict = field(default_factory=dict)
    width: int = 512
    height: int = 512
    spp: int = 1
    maxBounces: int = 8
    rrDepth: float = 3.0
    numTris: int = 0

    def build(self, scene):
        """Replay onto a wr::Scene-like builder; returns the material ids."""
        mats = []
        for op in self.ops:
            kind = op[0]
            if kind == "lambert":
                mats.append(scene.addLambertianMaterial(op[1], op[2]))
            elif kind == "emissive":
                mats.append(scene.addEmissiveMaterial(op[1]))
            elif kind == "glass":
                mats.append(scene.addGlassMaterial(op[1], op[2]))
            elif kind == "ggx":
                mats.append(scene.addGGXReflectionMaterial(op[1], op[2], op[3]))
            elif kind == "ggxt":
                mats.append(scene.addGGXTransmissionMaterial(op[1], op[2], op[3]))
            elif kind == "normalmap":
                scene.setMaterialNormalMap(op[1], op[2])
            elif kind == "texture":
                scene.addTextureRGBA8(op[1])
            elif kind == "mesh":
                scene.addTriangleMesh(op[1], op[2], op[3], op[4], op[5])
            elif kind == "env":
                scene.setEnvironmentRadiance(op[1])
            else:
                raise ValueError(kind)
        return mats


# ----------------------------------------------------------------- geometry
def quad(p0, p1, p2, p3):
    v = np.array([p0, p1, p2, p3], dtype=np.float32).ravel()
    i = np.array([0, 1, 2, 0, 2, 3], dtype=np.uint32)
    return v, i


def box(center, half, angle_y):
    """Axis box rotated about y, 12 outward-facing triangles."""
    c, s = math.cos(angle_y), math.sin(angle_y)
    corners = []
    for x in (-1, 1):
        for y in (-1, 1):
            for z in (-1, 1):
                px, py, pz = x * half[0], y * half[1], z * half[2]
                corners.append([center[0] + c * px + s * pz, center[1] + py, center[2] - s * px + c * pz])
    v = np.array(corners, dtype=np.float32)
    # corner index = (x>0)*4 + (y>0)*2 + (z>0)
    faces = [
        (0, 1, 3, 2),  # -x
        (4, 6, 7, 5),  # +x
        (0, 4, 5, 1),  # -y
        (2, 3, 7, 6),  # +y
        (0, 2, 6, 4),  # -z
        (1, 5, 7, 3),  # +z
    ]
    idx = []
    for a, b, cc, d in faces:
        idx += [a, b, cc, a, cc, d]
    return v.ravel(), np.array(idx, dtype=np.uint32)


@lru_cache(maxsize=4)
def icosphere(subdiv: int = 4):
    """Unit icosphere: 10*4^s + 2 vertices, 20*4^s faces (s=4: 2,562 / 5,120)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts, dtype=np.float64)
    F = np.array(faces, dtype=np.uint32)
    return V, F


def sphere_uv(n):
    u = 0.5 + np.arctan2(n[:, 2], n[:, 0]) / (2 * np.pi)
    v = 0.5 - np.arcsin(np.clip(n[:, 1], -1, 1)) / np.pi
    return np.stack([u, v], axis=1)


def value_noise3(p: np.ndarray, seed: int, freq: float) -> np.ndarray:
    """Seeded lattice value noise in [0,1], trilinear with smoothstep."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    f = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 0x5BD1E995
        x = x ^ (x >> 15)
        return (x & 0xFFFF).astype(np.float64) / 65535.0

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out = out + w * h(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
    return out


def checker_texture(size: int = 1024, cells: int = 16) -> np.ndarray:
    y, x = np.mgrid[0:size, 0:size]
    c = ((x * cells // size) + (y * cells // size)) % 2
    img = np.empty((size, size, 4), dtype=np.uint8)
    img[..., 0] = np.where(c, 230, 40)
    img[..., 1] = np.where(c, 220, 60)
    img[..., 2] = np.where(c, 200, 90)
    img[..., 3] = 255
    return img


def noise_normal_map(size: int, seed: int) -> np.ndarray:
    """Tangent-space normal map from the gradient of seeded 2-D value noise."""
    y, x = np.mgrid[0:size, 0:size].astype(np.float64)
    p = np.stack([x.ravel() / size, y.ravel() / size, np.zeros(size * size)], axis=1)
    hgt = value_noise3(p, seed, 24.0).reshape(size, size)
    gx = (np.roll(hgt, -1, axis=1) - np.roll(hgt, 1, axis=1)) * 4.0
    gy = (np.roll(hgt, -1, axis=0) - np.roll(hgt, 1, axis=0)) * 4.0
    n = np.stack([-gx, -gy, np.ones_like(gx)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.empty((size, size, 4), dtype=np.uint8)
    img[..., :3] = np.clip((n * 0.5 + 0.5) * 255.0 + 0.5, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    return img


# ------------------------------------------------------------------ scenes
def cornell(width: int = 512, height: int = 512, spp: int = 64) -> SceneDesc:
    d = SceneDesc(width=width, height=height, spp=spp, maxBounces=8, rrDepth=3.0)
    white, red, green, light = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (17.0, 12.0, 4.0)
    d.ops += [("lambert", white, 0), ("lambert", red, 0), ("lambert", green, 0), ("emissive", light)]
    W, R, G, L = 0, 1, 2, 3
    walls = [
        (((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1)), W),   # floor
        (((-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)), W),       # ceiling
        (((-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1)), W),   # back
        (((-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)), R),   # left
        (((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)), G),       # right
    ]
    for pts, m in walls:
        v, i = quad(*pts)
        d.ops.append(("mesh", v, i, None, None, m))
    v, i = box((0.35, -0.7, 0.3), (0.3, 0.3, 0.3), math.radians(-18.0))
    d.ops.append(("mesh", v, i, None, None, W))
    v, i = box((-0.35, -0.4, -0.35), (0.3, 0.6, 0.3), math.radians(15.0))
    d.ops.append(("mesh", v, i, None, None, W))
    v, i = quad((-0.25, 0.99, -0.25), (0.25, 0.99, -0.25), (0.25, 0.99, 0.25), (-0.25, 0.99, 0.25))
    d.ops.append(("mesh", v, i, None, None, L))
    d.ops.append(("env", (0.0, 0.0, 0.0)))
    d.camera = dict(position=(0.0, 0.0, 3.0), target=(0.0, 0.0, 2.0), up=(0.0, 1.0, 0.0),
                    fovY=math.radians(45.0), aspect=width / height)
    d.numTris = 36
    return d


def material_scene(width: int = 1920, height: int = 1080, spp: int = 256, seed: int = 1,
                   tex_size: int = 1024) -> SceneDesc:
    rng = np.random.default_rng(seed)
    d = SceneDesc(width=width, height=height, spp=spp, maxBounces=16, rrDepth=3.0)
    d.ops.append(("texture", checker_texture(tex_size)))            # id 1
    d.ops.append(("texture", noise_normal_map(256, seed + 7)))       # id 2
    mats = []
    ground = len(mats); mats.append(("lambert", (0.8, 0.8, 0.8), 1))
    kinds = []
    for k in range(25):
        kind = k % 5
        if kind == 0:
            kinds.append(("lambert", (0.9, 0.9, 0.9), 1))
        elif kind == 1:
            kinds.append(("glass", (1.0, 1.0, 1.0), 1.5))
        elif kind == 2:
            # GGX reflection, roughness U(0.1, 0.6), metallic at libWR's default 0 (scene.h:45)
            kinds.append(("ggx", tuple(float(x) for x in rng.uniform(0.5, 1.0, 3)), float(rng.uniform(0.1, 0.6)),
                          0.0))
        elif kind == 3:
            kinds.append(("ggxt", (1.0, 1.0, 1.0), 0.2, 1.5))
        else:
            kinds.append(("lambert", tuple(float(x) for x in rng.uniform(0.3, 0.9, 3)), 0))
    mats += kinds
    emis = []
    for q in range(4):
        emis.append(("emissive", (8.0 + 4 * q, 8.0 + 2 * q, 6.0 + q)))
    for m in mats + emis:
        d.ops.append(m)
    mat_ids = list(range(len(mats) + len(emis)))
    # normal map on the 5th-kind lambertians
    for k in range(25):
        if k % 5 == 4:
            d.ops.append(("normalmap", 1 + k, 2))
    ntris = 0
    v, i = quad((-30, 0, -30), (-30, 0, 30), (30, 0, 30), (30, 0, -30))
    uv = np.array([0, 0, 0, 8, 8, 8, 8, 0], dtype=np.float32)
    d.ops.append(("mesh", v, i, None, uv, ground))
    ntris += 2
    V, F = icosphere(4)
    for k in range(25):
        gx, gz = k % 5, k // 5
        cx = (gx - 2) * 2.6 + rng.uniform(-0.3, 0.3)
        cz = (gz - 2) * 2.6 + rng.uniform(-0.3, 0.3)
        r = 0.9
        verts = (V * r + np.array([cx, r, cz])).astype(np.float32)
        d.ops.append(("mesh", verts.ravel(), F.ravel(), V.astype(np.float32).ravel(),
                      sphere_uv(V).astype(np.float32).ravel(), 1 + k))
        ntris += F.shape[0]
    for q in range(4):
        x0 = -6.0 + 4.0 * q
        y = 9.0
        v, i = quad((x0, y, -2.0), (x0 + 2.0, y, -2.0), (x0 + 2.0, y, 0.0), (x0, y, 0.0))
        d.ops.append(("mesh", v, i, None, None, len(mats) + q))
        ntris += 2
    d.ops.append(("env", (0.1, 0.1, 0.1)))
    d.camera = dict(position=(0.0, 6.5, 15.0), target=(0.0, 0.8, 0.0), up=(0.0, 1.0, 0.0),
                    fovY=math.radians(40.0), aspect=width / height)
    d.numTris = ntris
    assert mat_ids  # silence linters
    return d


def large_scene(width: int = 3840, height: int = 2160, spp: int = 16, seed: int = 2,
                num_spheres: int = 2000, num_emitters: int = 64) -> SceneDesc:
    rng = np.random.default_rng(seed)
    d = SceneDesc(width=width, height=height, spp=spp, maxBounces=8, rrDepth=3.0)
    kinds = [("lambert", (0.7, 0.7, 0.7), 0), ("glass", (1.0, 1.0, 1.0), 1.5), ("ggx", (0.9, 0.8, 0.6), 0.3, 1.0),
             ("ggxt", (1.0, 1.0, 1.0), 0.2, 1.5), ("lambert", (0.4, 0.6, 0.8), 0)]
    for k in kinds:
        d.ops.append(k)
    # power-law radiance on [1, 100], alpha = 1.5: x = (1 - u (1 - 100^-0.5))^-2
    u = rng.uniform(size=num_emitters)
    rad = (1.0 - u * (1.0 - 100.0 ** -0.5)) ** -2.0
    for e in range(num_emitters):
        d.ops.append(("emissive", (float(rad[e]), float(rad[e]) * 0.9, float(rad[e]) * 0.8)))
    V, F = icosphere(4)
    centers = rng.uniform(0.0, 100.0, size=(num_spheres, 3))
    scales = np.exp(rng.uniform(math.log(0.2), math.log(3.0), size=num_spheres))
    mtypes = rng.integers(0, 5, size=num_spheres)
    disp = value_noise3(V * 4.0, seed, 1.0)
    ntris = 0
    # merge spheres of one material into a single mesh per chunk to keep the
    # op list short
    chunk = 100
    for c0 in range(0, num_spheres, chunk):
        for m in range(5):
            sel = [s for s in range(c0, min(c0 + chunk, num_spheres)) if mtypes[s] == m]
            if not sel:
                continue
            vs, fs = [], []
            for j, s in enumerate(sel):
                r = scales[s] * (0.85 + 0.3 * np.roll(disp, s * 37))
                vs.append(V * r[:, None] + centers[s])
                fs.append(F + j * V.shape[0])
            d.ops.append(("mesh", np.concatenate(vs).astype(np.float32).ravel(),
                          np.concatenate(fs).astype(np.uint32).ravel(), None, None, m))
            ntris += F.shape[0] * len(sel)
    for e in range(num_emitters):
        c = rng.uniform(0.0, 100.0, size=3)
        a = rng.normal(size=(3, 3))
        tri = c + a * 1.5
        d.ops.append(("mesh", tri.astype(np.float32).ravel(), np.array([0, 1, 2], dtype=np.uint32), None, None,
                      5 + e))
        ntris += 1
    d.ops.append(("env", (0.05, 0.05, 0.05)))
    d.camera = dict(position=(50.0, 50.0, -70.0), target=(50.0, 50.0, 50.0), up=(0.0, 1.0, 0.0),
                    fovY=math.radians(55.0), aspect=width / height)
    d.numTris = ntris
    return d


# ------------------------------------------------------------ serving load
SERVING_RESOLUTIONS = ((256, 256), (512, 512), (1024, 768), (1920, 1080))
SERVING_SPP = (1, 4, 16, 64)
SERVING_SCENES = ("cornell", "material", "large")


def serving_requests(n: int, seed: int = 7, depth: int = 8) -> list[dict]:
    """Seeded request stream over the three resident scenes: per request a
    scene, a random camera around it, a resolution from
    SERVING_RESOLUTIONS and a sample count from SERVING_SPP, depth 8."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        scene = SERVING_SCENES[int(rng.integers(0, len(SERVING_SCENES)))]
        w, h = SERVING_RESOLUTIONS[int(rng.integers(0, len(SERVING_RESOLUTIONS)))]
        spp = SERVING_SPP[int(rng.integers(0, len(SERVING_SPP)))]
        if scene == "cornell":
            pos = (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(2.2, 3.2))
            tgt = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 0.0)
            fov = 45.0
        elif scene == "material":
            th = rng.uniform(0.0, 2.0 * math.pi)
            rad = rng.uniform(10.0, 18.0)
            pos = (rad * math.sin(th), rng.uniform(3.0, 9.0), rad * math.cos(th))
            tgt = (rng.uniform(-2.0, 2.0), 0.8, rng.uniform(-2.0, 2.0))
            fov = 40.0
        else:
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            pos = tuple(50.0 + v * rng.uniform(60.0, 90.0))
            tgt = tuple(50.0 + rng.normal(size=3) * 10.0)
            fov = 55.0
        out.append(dict(scene=scene, width=w, height=h, spp=spp, maxBounces=depth, rrDepth=3.0,
                        camera=dict(position=tuple(float(x) for x in pos), target=tuple(float(x) for x in tgt),
                                    up=(0.0, 1.0, 0.0), fovY=math.radians(fov), aspect=w / h)))
    return out
