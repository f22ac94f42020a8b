"""ctypes bridge to the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never imports it.

OracleScene exposes the wr::Scene builder method names so a SceneDesc can be
replayed onto it exactly as onto the product Scene.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# WR_ORACLE_LIB selects another build of the same sources (the sanitized one)
LIB = Path(os.environ.get("WR_ORACLE_LIB", HERE / "build" / "liboracle.so"))
_lib = None

WRK_MAT_LAMBERTIAN, WRK_MAT_EMISSIVE, WRK_MAT_GLASS, WRK_MAT_GGX_REFLECTION, WRK_MAT_GGX_TRANSMISSION = range(5)


class wrk_material(C.Structure):
    _fields_ = [("albedo", C.c_float * 4), ("emission", C.c_float * 4), ("ior", C.c_float), ("type", C.c_uint32),
                ("roughness", C.c_float), ("metallic", C.c_float), ("albedoTextureId", C.c_uint32),
                ("normalTextureId", C.c_uint32), ("_pad", C.c_uint32 * 2)]


class orc_camera_in(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fovY", C.c_float),
                ("aspect", C.c_float), ("focalDistance", C.c_float), ("lensRadius", C.c_float)]


class orc_params_in(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("useNEE", C.c_uint32), ("maxBounces", C.c_uint32),
                ("rrStartDepth", C.c_float), ("fireflyClamp", C.c_float), ("iterationCap", C.c_uint32)]


class EventIO(C.Structure):
    """orc::EventIO (wr_oracle.hpp): one shade() visit's ray state and hit."""
    _fields_ = [("origin", C.c_float * 3), ("direction", C.c_float * 3), ("throughput", C.c_float * 3),
                ("radiance", C.c_float * 3), ("tMin", C.c_float), ("tMax", C.c_float), ("prevPdf", C.c_float),
                ("seed", C.c_uint32), ("rngDimension", C.c_uint32), ("depth", C.c_uint32), ("stage", C.c_uint32),
                ("neeDone", C.c_uint32), ("prevWasDelta", C.c_uint32), ("shadowContribution", C.c_float * 3),
                ("savedDirection", C.c_float * 3), ("position", C.c_float * 3), ("normal", C.c_float * 3),
                ("geometricNormal", C.c_float * 3), ("u", C.c_float), ("v", C.c_float), ("materialId", C.c_uint32),
                ("flushed", C.c_uint32), ("flushL", C.c_float * 3)]


STAGE_TRACE, STAGE_SHADE, STAGE_SHADOW, STAGE_TERMINATED = range(4)

_P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "orc_scene_new": (_P, []),
            "orc_scene_free": (None, [_P]),
            "orc_scene_add_material": (C.c_uint32, [_P, C.POINTER(wrk_material)]),
            "orc_scene_add_texture": (C.c_uint32, [_P, _P, C.c_uint32, C.c_uint32]),
            "orc_scene_set_normal_map": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
            "orc_scene_add_mesh": (None, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32,
                                          C.c_uint32]),
            "orc_scene_set_env": (None, [_P, C.c_float, C.c_float, C.c_float]),
            "orc_scene_finalize": (C.c_int, [_P, C.c_int]),
            "orc_scene_num_tris": (C.c_uint32, [_P]),
            "orc_scene_num_emissive": (C.c_uint32, [_P]),
            "orc_scene_emissive": (None, [_P, _P, _P]),
            "orc_pcg_hash": (C.c_uint32, [C.c_uint32]),
            "orc_pcg_hash2": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_rnd_dim": (C.c_float, [C.c_uint32, C.c_uint32]),
            "orc_camera_seed": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_sincos": (None, [_P, C.c_uint32, _P, _P]),
            "orc_trace_closest": (None, [_P, _P, _P, C.c_uint32, C.c_int, C.c_int, _P, _P]),
            "orc_trace_any": (None, [_P, _P, _P, C.c_uint32, C.c_int, C.c_int, _P]),
            "orc_camera_basis": (None, [C.POINTER(orc_camera_in), _P]),
            "orc_render_paths": (None, [_P, C.POINTER(orc_camera_in), C.POINTER(orc_params_in), _P, _P, C.c_uint32,
                                        C.c_int, C.c_int, _P, _P]),
            "orc_render_guides": (None, [_P, C.POINTER(orc_camera_in), C.POINTER(orc_params_in), C.c_uint32,
                                         C.c_int, C.c_int, _P, _P]),
            "orc_render_image": (None, [_P, C.POINTER(orc_camera_in), C.POINTER(orc_params_in), C.c_uint32,
                                        C.c_uint32, C.c_int, C.c_int, C.c_int, _P, _P]),
            "orc_compact": (C.c_uint32, [_P, _P, C.c_uint32, _P]),
            "orc_aces": (None, [_P, C.c_uint32, _P]),
            "orc_fresnel": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "orc_ggx_D": (C.c_float, [C.c_float, C.c_float]),
            "orc_ggx_G1": (C.c_float, [C.c_float, C.c_float]),
            "orc_ggx_eval": (None, [_P, _P, _P, _P, C.c_float, C.c_float, C.c_float, _P]),
            "orc_ggx_pdf": (C.c_float, [_P, _P, _P, C.c_float]),
            "orc_ggx_sample": (C.c_int, [_P, _P, C.c_float, C.c_float, C.c_float, _P, _P]),
            "orc_sample_cosine": (None, [C.c_float, C.c_float, _P, _P]),
            "orc_frame": (None, [_P, _P, _P]),
            "orc_power_heuristic": (C.c_float, [C.c_float, C.c_float]),
            "orc_cdf_search": (C.c_uint32, [_P, C.c_uint32, C.c_float]),
            "orc_shade_event": (None, [_P, C.POINTER(orc_params_in), C.POINTER(EventIO)]),
            "orc_hit_attributes": (None, [_P, _P, _P, C.c_float, C.c_uint32, _P]),
            "orc_camera_ray": (None, [C.POINTER(orc_camera_in), C.POINTER(orc_params_in), C.c_uint32, C.c_uint32,
                                      _P, _P]),
            "orc_sample_texture": (None, [_P, C.c_uint32, C.c_float, C.c_float, _P]),
            "orc_effective_normal": (C.c_int, [_P, C.c_uint32, _P, C.c_float, C.c_float, _P]),
            "orc_sample_triangle": (None, [C.c_float, C.c_float, _P, _P, _P, _P, _P, _P]),
            "orc_offset_origin": (None, [_P, _P, _P, _P]),
            "orc_asym_correction": (C.c_float, [_P, _P, _P]),
            "orc_intersect_triangle": (C.c_int, [_P, _P, _P, _P, _P, C.c_float, C.c_float, _P]),
            "orc_wall_seconds": (C.c_double, []),
            "orc_max_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _mat(kind, albedo=(0, 0, 0), emission=(0, 0, 0), ior=1.0, rough=0.0, metal=0.0, tex=0):
    m = wrk_material()
    m.albedo[:] = [float(albedo[0]), float(albedo[1]), float(albedo[2]), 0.0]
    m.emission[:] = [float(emission[0]), float(emission[1]), float(emission[2]), 0.0]
    m.ior, m.type, m.roughness, m.metallic = float(ior), kind, float(rough), float(metal)
    m.albedoTextureId, m.normalTextureId = int(tex), 0
    return m


class OracleScene:
    """wr::Scene builder semantics (libWR/src/internal/scene.cpp:91-244) on the host."""

    def __init__(self):
        self.h = lib().orc_scene_new()
        self._mats = []
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    def _add(self, m):
        self._mats.append(m)
        return lib().orc_scene_add_material(self.h, C.byref(m))

    # scene.cpp:91-168 material records
    def addLambertianMaterial(self, albedo, tex=0):
        return self._add(_mat(WRK_MAT_LAMBERTIAN, albedo=albedo, ior=1.0, rough=1.0, tex=tex))

    def addEmissiveMaterial(self, emission):
        return self._add(_mat(WRK_MAT_EMISSIVE, emission=emission))

    def addGlassMaterial(self, albedo, ior):
        return self._add(_mat(WRK_MAT_GLASS, albedo=albedo, ior=ior))

    def addGGXReflectionMaterial(self, albedo, rough, metal=0.0):
        return self._add(_mat(WRK_MAT_GGX_REFLECTION, albedo=albedo, ior=1.5, rough=rough, metal=metal))

    def addGGXTransmissionMaterial(self, albedo, rough, ior):
        return self._add(_mat(WRK_MAT_GGX_TRANSMISSION, albedo=albedo, ior=ior, rough=rough))

    def setMaterialNormalMap(self, mat, tex):
        if lib().orc_scene_set_normal_map(self.h, int(mat), int(tex)) != 0:
            raise RuntimeError("Invalid material ID")

    def addTextureRGBA8(self, rgba):
        a = np.ascontiguousarray(rgba, dtype=np.uint8)
        self._keep.append(a)
        return lib().orc_scene_add_texture(self.h, _p(a), a.shape[1], a.shape[0])

    def addTriangleMesh(self, v, i, n=None, uv=None, mat=0):
        v = np.ascontiguousarray(v, dtype=np.float32).ravel()
        i = np.ascontiguousarray(i, dtype=np.uint32).ravel()
        n = None if n is None else np.ascontiguousarray(n, dtype=np.float32).ravel()
        uv = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32).ravel()
        lib().orc_scene_add_mesh(self.h, _p(v), v.size, _p(i), i.size, _p(n), 0 if n is None else n.size,
                                 _p(uv), 0 if uv is None else uv.size, int(mat))

    def setEnvironmentRadiance(self, rgb):
        lib().orc_scene_set_env(self.h, *map(float, rgb))

    def finalize(self, bvh=True):
        if lib().orc_scene_finalize(self.h, 1 if bvh else 0) != 0:
            raise RuntimeError("Scene has no geometry")

    def emissive(self):
        n = lib().orc_scene_num_emissive(self.h)
        tris = np.zeros(n, dtype=np.uint32)
        cdf = np.zeros(n, dtype=np.float32)
        if n:
            lib().orc_scene_emissive(self.h, _p(tris), _p(cdf))
        return tris, cdf


def build_oracle_scene(desc, bvh=True) -> OracleScene:
    s = OracleScene()
    desc.build(s)
    s.finalize(bvh)
    return s


def camera_in(cam: dict) -> orc_camera_in:
    c = orc_camera_in()
    c.position[:] = [float(x) for x in cam["position"]]
    c.target[:] = [float(x) for x in cam["target"]]
    c.up[:] = [float(x) for x in cam["up"]]
    c.fovY = float(cam["fovY"])
    c.aspect = float(cam.get("aspect", 1.0))
    c.focalDistance = float(cam.get("focalDistance", 0.0))
    c.lensRadius = float(cam.get("lensRadius", 0.0))
    return c


def params_in(width, height, maxBounces=8, rrDepth=3.0, useNEE=True, fireflyClamp=0.0, iterationCap=8):
    p = orc_params_in()
    p.width, p.height, p.useNEE, p.maxBounces = width, height, 1 if useNEE else 0, maxBounces
    p.rrStartDepth, p.fireflyClamp, p.iterationCap = float(rrDepth), float(fireflyClamp), iterationCap
    return p


def trace_closest(scene: OracleScene, org4, dir4, use_bvh=True, threads=0):
    org4 = np.ascontiguousarray(org4, dtype=np.float32)
    dir4 = np.ascontiguousarray(dir4, dtype=np.float32)
    n = org4.shape[0]
    t = np.zeros(n, dtype=np.float32)
    prim = np.zeros(n, dtype=np.uint32)
    lib().orc_trace_closest(scene.h, _p(org4), _p(dir4), n, 1 if use_bvh else 0, threads, _p(t), _p(prim))
    return t, prim


def trace_any(scene: OracleScene, org4, dir4, use_bvh=True, threads=0):
    org4 = np.ascontiguousarray(org4, dtype=np.float32)
    dir4 = np.ascontiguousarray(dir4, dtype=np.float32)
    n = org4.shape[0]
    occ = np.zeros(n, dtype=np.uint8)
    lib().orc_trace_any(scene.h, _p(org4), _p(dir4), n, 1 if use_bvh else 0, threads, _p(occ))
    return occ


def render_paths(scene: OracleScene, cam: dict, prm: orc_params_in, pixels, samples, use_bvh=True, threads=0):
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    n = pixels.size
    L = np.zeros((n, 3), dtype=np.float32)
    info = np.zeros((n, 4), dtype=np.uint32)
    ci = camera_in(cam)
    lib().orc_render_paths(scene.h, C.byref(ci), C.byref(prm), _p(pixels), _p(samples), n, 1 if use_bvh else 0,
                           threads, _p(L), _p(info))
    return L, info


def render_image(scene: OracleScene, cam: dict, prm: orc_params_in, spp: int, s0: int = 0, use_bvh=True, threads=0,
                 fp64=False):
    npix = prm.width * prm.height
    ci = camera_in(cam)
    if fp64:
        acc = np.zeros((npix, 3), dtype=np.float64)
        lib().orc_render_image(scene.h, C.byref(ci), C.byref(prm), s0, spp, 1 if use_bvh else 0, threads, 1, None,
                               _p(acc))
    else:
        acc = np.zeros((npix, 3), dtype=np.float32)
        lib().orc_render_image(scene.h, C.byref(ci), C.byref(prm), s0, spp, 1 if use_bvh else 0, threads, 0,
                               _p(acc), None)
    return acc.reshape(prm.height, prm.width, 3)


def render_guides(scene: OracleScene, cam: dict, prm: orc_params_in, spp: int, use_bvh=True, threads=0):
    """First-hit albedo and normal guide layers averaged over samples 0..spp-1,
    each (height*width, 3) float32."""
    n = prm.width * prm.height
    alb = np.zeros((n, 3), dtype=np.float32)
    nrm = np.zeros((n, 3), dtype=np.float32)
    ci = camera_in(cam)
    lib().orc_render_guides(scene.h, C.byref(ci), C.byref(prm), spp, 1 if use_bvh else 0, threads, _p(alb), _p(nrm))
    return alb, nrm


def compact(keep, items=None):
    keep = np.ascontiguousarray(keep, dtype=np.uint8)
    n = keep.size
    out = np.zeros(n, dtype=np.uint32)
    it = None if items is None else np.ascontiguousarray(items, dtype=np.uint32)
    c = lib().orc_compact(_p(keep), _p(it), n, _p(out))
    return out[:c]


def aces(img):
    a = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros(a.shape, dtype=np.uint8)
    lib().orc_aces(_p(a), a.size // 3, _p(out))
    return out


# ------------------------------------------------------- single-step hooks
def _f32(v, n=3):
    a = np.ascontiguousarray(v, dtype=np.float32).ravel()
    assert a.size == n
    return a


def shade_event(scene: OracleScene, prm: orc_params_in, *, direction, position, normal, geometric_normal=None,
                material=0, seed=0, dim=0, depth=0, throughput=(1, 1, 1), radiance=(0, 0, 0), nee_done=0,
                prev_was_delta=1, prev_pdf=1.0, u=0.0, v=0.0) -> EventIO:
    """Run one shade() visit (shade.cu:24-85) and return the resulting state."""
    e = EventIO()
    e.direction[:] = [float(x) for x in direction]
    e.position[:] = [float(x) for x in position]
    e.origin[:] = [float(x) for x in position]
    e.normal[:] = [float(x) for x in normal]
    e.geometricNormal[:] = [float(x) for x in (normal if geometric_normal is None else geometric_normal)]
    e.throughput[:] = [float(x) for x in throughput]
    e.radiance[:] = [float(x) for x in radiance]
    e.tMin, e.tMax, e.prevPdf = 0.0, 1e20, float(prev_pdf)
    e.seed, e.rngDimension, e.depth = int(seed), int(dim), int(depth)
    e.stage, e.neeDone, e.prevWasDelta = STAGE_SHADE, int(nee_done), int(prev_was_delta)
    e.u, e.v, e.materialId = float(u), float(v), int(material)
    lib().orc_shade_event(scene.h, C.byref(prm), C.byref(e))
    return e


def hit_attributes(scene: OracleScene, o, d, t, prim):
    """(position, shading normal, geometric normal, u, v) of trace.cu:167-264."""
    out = np.zeros(11, np.float32)
    o, d = _f32(o), _f32(d)
    lib().orc_hit_attributes(scene.h, _p(o), _p(d), float(t), int(prim), _p(out))
    return out[0:3], out[3:6], out[6:9], float(out[9]), float(out[10])


def camera_ray(cam: dict, prm: orc_params_in, pixel, sample):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    ci = camera_in(cam)
    lib().orc_camera_ray(C.byref(ci), C.byref(prm), int(pixel), int(sample), _p(o), _p(d))
    return o, d


def sample_texture(scene: OracleScene, tex_id, u, v):
    out = np.zeros(3, np.float32)
    lib().orc_sample_texture(scene.h, int(tex_id), float(u), float(v), _p(out))
    return out


def effective_normal(scene: OracleScene, mat, n, u=0.0, v=0.0):
    out = np.zeros(3, np.float32)
    n = _f32(n)
    if lib().orc_effective_normal(scene.h, int(mat), _p(n), float(u), float(v), _p(out)) != 0:
        raise ValueError("material id out of range")
    return out


def sample_triangle(u1, u2, v0, v1, v2):
    pt, nrm, area = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
    a, b, c = _f32(v0), _f32(v1), _f32(v2)
    lib().orc_sample_triangle(float(u1), float(u2), _p(a), _p(b), _p(c), _p(pt), _p(nrm), _p(area))
    return pt, nrm, float(area[0])


def offset_origin(pos, n, d):
    out = np.zeros(3, np.float32)
    a, b, c = _f32(pos), _f32(n), _f32(d)
    lib().orc_offset_origin(_p(a), _p(b), _p(c), _p(out))
    return out


def asym_correction(wi, sn, gn):
    a, b, c = _f32(wi), _f32(sn), _f32(gn)
    return float(lib().orc_asym_correction(_p(a), _p(b), _p(c)))


def intersect_triangle(o, d, v0, v1, v2, tmin=0.0, tmax=1e20):
    t = np.zeros(1, np.float32)
    args = [_f32(x) for x in (o, d, v0, v1, v2)]
    hit = lib().orc_intersect_triangle(*[_p(a) for a in args], float(tmin), float(tmax), _p(t))
    return bool(hit), float(t[0])
