"""MI355X wavefront path tracer behind libWR's wr:: API.

Python mirror of the reference's public C++ interface
(libWR/include/wr/{context,scene,renderer,types}.h): Context, Scene,
Texture2D, Renderer, Camera, RenderParams — same names, argument meaning and
error behaviour (RuntimeError carrying the reference's messages). Every call
goes through the C ABI of the gfx950 library lib/libwr_amd.so
(include/wr_c.h over the wr:: C++ layer, which drives the wrk_* HIP kernels of
include/wrk.h). There is no CPU fallback: importing the renderer without the
built library raises.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libwr_amd.so"

_lib = None


class wrk_bvh(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("tris", C.c_uint64), ("numNodes", C.c_uint32), ("numTris", C.c_uint32),
                ("maxDepth", C.c_uint32), ("nodeBytes", C.c_uint32), ("sceneLo", C.c_float * 4),
                ("sceneHi", C.c_float * 4)]


class wrk_scene_view(C.Structure):
    _fields_ = [("vertices", C.c_uint64), ("normals", C.c_uint64), ("uvs", C.c_uint64), ("indices", C.c_uint64),
                ("triMaterialIds", C.c_uint64), ("materials", C.c_uint64), ("textures", C.c_uint64),
                ("emissiveTris", C.c_uint64), ("emissiveCdf", C.c_uint64), ("numTris", C.c_uint32),
                ("numMaterials", C.c_uint32), ("numTextures", C.c_uint32), ("numEmissive", C.c_uint32),
                ("envRadiance", C.c_float * 4), ("emissiveGuide", C.c_uint64), ("emissiveGuideSize", C.c_uint32),
                ("_pad2", C.c_uint32 * 3)]


class wrk_material(C.Structure):
    _fields_ = [("albedo", C.c_float * 4), ("emission", C.c_float * 4), ("ior", C.c_float), ("type", C.c_uint32),
                ("roughness", C.c_float), ("metallic", C.c_float), ("albedoTextureId", C.c_uint32),
                ("normalTextureId", C.c_uint32), ("_pad", C.c_uint32 * 2)]


class wr_c_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fovY", C.c_float),
                ("aspect", C.c_float), ("focalDistance", C.c_float), ("lensRadius", C.c_float)]


class wr_c_render_params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("useNEE", C.c_uint32),
                ("maxBounces", C.c_uint32), ("russianRouletteDepth", C.c_float), ("fireflyClamp", C.c_float),
                ("denoiser", C.c_uint32), ("firstSample", C.c_uint32)]


class wr_c_stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("paths", C.c_uint64), ("closestRays", C.c_uint64),
                ("shadowRays", C.c_uint64), ("iterations", C.c_uint32), ("batches", C.c_uint32),
                ("msGenerate", C.c_double), ("msTraceClosest", C.c_double), ("msTraceAny", C.c_double),
                ("msClassify", C.c_double), ("msPartition", C.c_double), ("msShade", C.c_double),
                ("msAccumulate", C.c_double), ("launchesTraceClosest", C.c_uint32),
                ("launchesTraceAny", C.c_uint32), ("closestNodeFetches", C.c_uint64),
                ("closestTriFetches", C.c_uint64), ("anyNodeFetches", C.c_uint64), ("anyTriFetches", C.c_uint64),
                ("classifyItems", C.c_uint64), ("partitionItems", C.c_uint64), ("shadeItems", C.c_uint64),
                ("closestStackSpills", C.c_uint64), ("anyStackSpills", C.c_uint64), ("fusedPaths", C.c_uint64),
                ("msFused", C.c_double)]


class wrk_paths(C.Structure):
    """include/wrk.h wrk_paths: device addresses of the path pool arrays."""
    _fields_ = [(name, C.c_uint64) for name in ("org", "dir", "thr", "rad", "state", "contrib", "savedOrg",
                                                 "savedDir", "hit", "occluded", "result", "guides")] + \
               [("capacity", C.c_uint32), ("_pad", C.c_uint32)]


class wr_c_server_config(C.Structure):
    _fields_ = [("maxPathsInFlight", C.c_uint32), ("maxChunkPaths", C.c_uint32), ("memoryBudgetBytes", C.c_uint64),
                ("iterationCap", C.c_uint32), ("flags", C.c_uint32), ("expressMaxPaths", C.c_uint32),
                ("expressPathsInFlight", C.c_uint32), ("expressIterations", C.c_uint32), ("fusedMaxPaths", C.c_uint32),
                ("backfillPaths", C.c_uint32)]


class wr_c_request_status(C.Structure):
    _fields_ = [("done", C.c_uint32), ("admitted", C.c_uint32), ("submitTime", C.c_double),
                ("admitTime", C.c_double), ("doneTime", C.c_double), ("paths", C.c_uint64)]


_P = C.c_void_p
_SIGS = {
    # wr_c_*
    "wr_c_last_error": (C.c_char_p, []),
    "wr_c_context_create": (C.c_int, [C.c_int, C.c_int]),
    "wr_c_scene_create": (_P, []),
    "wr_c_scene_destroy": (None, [_P]),
    "wr_c_scene_add_lambertian": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float, C.c_uint32]),
    "wr_c_scene_add_emissive": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float]),
    "wr_c_scene_add_glass": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "wr_c_scene_add_ggx_reflection": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "wr_c_scene_add_ggx_transmission": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "wr_c_scene_set_normal_map": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "wr_c_scene_add_mesh": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, _P, C.c_size_t, _P, C.c_size_t,
                                      C.c_uint32]),
    "wr_c_texture_create": (C.c_uint32, [_P, C.c_char_p]),
    "wr_c_texture_destroy": (None, [_P, C.c_uint32]),
    "wr_c_scene_add_texture_rgba8": (C.c_uint32, [_P, _P, C.c_uint32, C.c_uint32]),
    "wr_c_scene_set_environment_radiance": (C.c_int, [_P, C.c_float, C.c_float, C.c_float]),
    "wr_c_scene_set_environment_map": (C.c_int, [_P, C.c_char_p]),
    "wr_c_scene_finalize": (C.c_int, [_P]),
    "wr_c_scene_get_bvh": (C.c_int, [_P, C.POINTER(wrk_bvh)]),
    "wr_c_scene_get_view": (C.c_int, [_P, C.POINTER(wrk_scene_view)]),
    "wr_c_scene_build_seconds": (C.c_double, [_P]),
    "wr_c_renderer_create": (_P, []),
    "wr_c_renderer_destroy": (None, [_P]),
    "wr_c_renderer_set_wavefront": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "wr_c_renderer_set_fused": (C.c_int, [_P, C.c_uint64]),
    "wr_c_render": (C.c_int, [_P, _P, C.POINTER(wr_c_camera), C.POINTER(wr_c_render_params), _P]),
    "wr_c_render_paths": (C.c_int, [_P, _P, C.POINTER(wr_c_camera), C.POINTER(wr_c_render_params), C.c_uint64,
                                    C.c_uint32, _P]),
    "wr_c_renderer_stats": (C.c_int, [_P, C.POINTER(wr_c_stats)]),
    "wr_c_renderer_load_denoiser": (C.c_int, [_P, C.c_char_p]),
    "wr_c_render_guides": (C.c_int, [_P, _P, C.POINTER(wr_c_camera), C.POINTER(wr_c_render_params), _P, _P, _P,
                                     _P]),
    "wr_c_server_create": (_P, [C.POINTER(wr_c_server_config)]),
    "wr_c_server_destroy": (None, [_P]),
    "wr_c_server_submit": (C.c_uint64, [_P, _P, C.POINTER(wr_c_camera), C.POINTER(wr_c_render_params)]),
    "wr_c_server_step": (C.c_int, [_P]),
    "wr_c_server_status": (C.c_int, [_P, C.c_uint64, C.POINTER(wr_c_request_status)]),
    "wr_c_server_fetch": (C.c_int, [_P, C.c_uint64, _P]),
    "wr_c_server_take_completed": (C.c_int, [_P, _P, C.c_uint32]),
    "wr_c_server_stats": (C.c_int, [_P, C.POINTER(wr_c_stats)]),
    "wr_c_server_pending": (C.c_uint32, [_P]),
    "wr_c_server_in_flight": (C.c_uint32, [_P]),
    "wr_c_server_clock": (C.c_double, []),
    # wrk_unet_* / denoiser kernels (include/wrk_unet.h)
    "wrk_conv3x3_fwd": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_conv3x3_fwd_first": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_conv3x3_fwd_up": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_conv3x3_fwd_pool": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P,
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_conv3x3_fwd_unpack": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32,
                                         C.c_uint32, C.c_uint32, _P]),
    "wrk_conv3x3_wgrad_scratch_bytes": (C.c_uint64, [C.c_uint32] * 5),
    "wrk_conv3x3_wgrad": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, _P, _P, _P, C.c_uint64, _P]),
    "wrk_weights_bf16": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "wrk_relu_bwd": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P]),
    "wrk_pool2_fwd": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_pool2_bwd": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_uint32, _P]),
    "wrk_up2_fwd": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_up2_bwd": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "wrk_unet_pack": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32, _P, C.c_uint32,
                                C.c_uint32, C.c_uint32, _P]),
    "wrk_unet_unpack": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint64, _P]),
    "wrk_unet_loss_scratch_bytes": (C.c_uint64, [C.c_uint32] * 3),
    "wrk_unet_loss": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, _P,
                                C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P]),
    "wrk_kpn_taps": (C.c_uint32, []),
    "wrk_kpn_apply": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32, C.c_uint32,
                                C.c_uint32, _P]),
    "wrk_kpn_loss": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_uint32, _P, _P]),
    "wrk_adam": (C.c_int, [_P, _P, _P, _P, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, _P]),
    "wrk_unet_last_error": (C.c_char_p, []),
    "wrk_unet_create": (_P, []),
    "wrk_unet_create_head": (_P, [C.c_uint32]),
    "wrk_unet_head": (C.c_uint32, [_P]),
    "wrk_unet_params_for_head": (C.c_uint64, [C.c_uint32]),
    "wrk_unet_destroy": (None, [_P]),
    "wrk_unet_num_params": (C.c_uint64, [_P]),
    "wrk_unet_init": (C.c_int, [_P, C.c_uint64]),
    "wrk_unet_get_params": (C.c_int, [_P, _P]),
    "wrk_unet_set_params": (C.c_int, [_P, _P]),
    "wrk_unet_get_grads": (C.c_int, [_P, _P]),
    "wrk_unet_get_state": (C.c_int, [_P, _P, _P, _P]),
    "wrk_unet_set_state": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "wrk_unet_set_sample_counts": (C.c_int, [_P, _P, C.c_uint32]),
    "wrk_unet_set_noise": (C.c_int, [_P, _P]),
    "wrk_unet_forward": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                   _P, _P]),
    "wrk_unet_train_step": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, _P, _P]),
    # wrk_*
    "wrk_last_error": (C.c_char_p, []),
    "wrk_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "wrk_bvh_query_sizes": (C.c_int, [C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]),
    "wrk_bvh_build": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P, C.c_uint64, C.POINTER(wrk_bvh), _P]),
    "wrk_trace_closest": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "wrk_trace_any": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "wrk_trace_closest_counted": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, _P, _P, _P, _P]),
    "wrk_trace_any_counted": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, _P, _P, _P, _P]),
    "wrk_partition_scratch_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32]),
    "wrk_compact_scratch_bytes": (C.c_uint64, [C.c_uint32]),
    "wrk_compact": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P, C.c_uint64, _P]),
    "wrk_partition": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint64, _P]),
    "wrk_partition_noalloc": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "wrk_radix_sort_scratch_bytes": (C.c_uint64, [C.c_uint32]),
    "wrk_radix_sort_pairs_u64": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint64,
                                           _P]),
    "wrk_exclusive_scan_scratch_bytes": (C.c_uint64, [C.c_uint32]),
    "wrk_exclusive_scan_u32": (C.c_int, [_P, _P, C.c_uint32, _P, _P, C.c_uint64, _P]),
    "wrk_partition_noalloc_max_items": (C.c_uint32, []),
    "wrk_generate": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32,
                               _P]),
    "wrk_classify": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P]),
    "wrk_trace_any_resolve": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "wrk_trace_closest_resolve": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "wrk_bvh_tag_categories": (C.c_int, [C.POINTER(wrk_bvh), _P, _P]),
    "wrk_trace_closest_categorized": (C.c_int, [C.POINTER(wrk_bvh), _P, _P, _P, C.c_uint32, _P, _P, _P, _P, _P]),
    "wrk_shade": (C.c_int, [_P, _P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P]),
    "wrk_render_fused": (C.c_int, [_P, _P, _P, _P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint32, _P]),
    "wrk_slot_finish": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P]),
    "wrk_slot_finish_map": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, _P, _P]),
    "wrk_accumulate": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P]),
    "wrk_noise_stderr": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "wrk_average": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "wrk_tonemap_aces": (C.c_int, [_P, C.c_uint32, _P, _P]),
}


def lib():
    """Load lib/libwr_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m offlinerenderer_amd.build`")
        L = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _err():
    return lib().wr_c_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {_err()}")


def _wrk_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().wrk_last_error().decode()}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- types
@dataclass
class Camera:
    """libWR/include/wr/types.h:16-51."""
    position: tuple = (0.0, 0.0, 0.0)
    target: tuple = (0.0, 0.0, -1.0)
    up: tuple = (0.0, 1.0, 0.0)
    fovY: float = math.radians(45.0)
    aspect: float = 1.0
    focalDistance: float = 0.0
    lensRadius: float = 0.0

    def to_c(self):
        c = wr_c_camera()
        c.position[:] = [float(x) for x in self.position]
        c.target[:] = [float(x) for x in self.target]
        c.up[:] = [float(x) for x in self.up]
        c.fovY, c.aspect = float(self.fovY), float(self.aspect)
        c.focalDistance, c.lensRadius = float(self.focalDistance), float(self.lensRadius)
        return c


@dataclass
class DenoiserConfig:
    enabled: bool = False
    useAlbedo: bool = False
    useNormal: bool = False
    hdrIntensity: float = 1.0


@dataclass
class RenderParams:
    """libWR/include/wr/renderer.h:25-36."""
    width: int = 0
    height: int = 0
    spp: int = 1
    denoiser: DenoiserConfig = field(default_factory=DenoiserConfig)
    useNEE: bool = True
    maxBounces: int = 8
    russianRouletteDepth: float = 3.0
    fireflyClamp: float = 0.0
    firstSample: int = 0  # additive: samples [firstSample, firstSample + spp)

    def to_c(self):
        p = wr_c_render_params()
        p.width, p.height, p.spp = int(self.width), int(self.height), int(self.spp)
        p.useNEE = 1 if self.useNEE else 0
        p.maxBounces = int(self.maxBounces)
        p.russianRouletteDepth = float(self.russianRouletteDepth)
        p.fireflyClamp = float(self.fireflyClamp)
        p.denoiser = 1 if self.denoiser.enabled else 0
        p.firstSample = int(self.firstSample)
        return p


@dataclass
class WavefrontConfig:
    """iterationCap 8 = libWR's host loop (renderer.cpp:213); 0 = run to completion.
    fusedMaxPaths: renders of at most that many paths run as one fused
    launch per chunk (same per-path results; 0 = bounce iterations only)."""
    iterationCap: int = 8
    maxPathsInFlight: int = 1 << 26
    maxChunkPaths: int = 0
    timeKernels: bool = False
    countTraversal: bool = False
    fusedMaxPaths: int = 0


@dataclass
class ContextConfig:
    deviceId: int = 0
    logLevel: int = 0
    enableValidation: bool = False


class Context:
    """libWR/include/wr/context.h: process-wide device context."""

    def __init__(self, config: ContextConfig | None = None):
        cfg = config or ContextConfig()
        self._device = cfg.deviceId
        _check(lib().wr_c_context_create(cfg.deviceId, cfg.logLevel), "Context")

    def createScene(self) -> "Scene":
        return Scene()

    def createRenderer(self) -> "Renderer":
        return Renderer()

    def getDeviceId(self) -> int:
        return self._device


class Scene:
    """libWR/include/wr/scene.h:35-103."""

    def __init__(self):
        self._h = lib().wr_c_scene_create()
        if not self._h:
            raise RuntimeError(f"Scene: {_err()}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.wr_c_scene_destroy(h)
            self._h = None

    def _id(self, v, what):
        if v == 0xFFFFFFFF:
            raise RuntimeError(f"{what}: {_err()}")
        return v

    def addLambertianMaterial(self, albedo, albedoTextureId: int = 0) -> int:
        return self._id(lib().wr_c_scene_add_lambertian(self._h, *map(float, albedo), int(albedoTextureId)),
                        "addLambertianMaterial")

    def addEmissiveMaterial(self, emission) -> int:
        return self._id(lib().wr_c_scene_add_emissive(self._h, *map(float, emission)), "addEmissiveMaterial")

    def addGlassMaterial(self, albedo, ior: float) -> int:
        return self._id(lib().wr_c_scene_add_glass(self._h, *map(float, albedo), float(ior)), "addGlassMaterial")

    def addGGXReflectionMaterial(self, albedo, roughness: float, metallic: float = 0.0) -> int:
        return self._id(lib().wr_c_scene_add_ggx_reflection(self._h, *map(float, albedo), float(roughness),
                                                             float(metallic)), "addGGXReflectionMaterial")

    def addGGXTransmissionMaterial(self, albedo, roughness: float, ior: float) -> int:
        return self._id(lib().wr_c_scene_add_ggx_transmission(self._h, *map(float, albedo), float(roughness),
                                                               float(ior)), "addGGXTransmissionMaterial")

    def setMaterialNormalMap(self, materialId: int, normalTextureId: int):
        _check(lib().wr_c_scene_set_normal_map(self._h, int(materialId), int(normalTextureId)), "setMaterialNormalMap")

    def addTriangleMesh(self, vertices, indices, normals=None, uvs=None, materialId: int = 0):
        v = np.ascontiguousarray(vertices, dtype=np.float32).ravel()
        i = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
        n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).ravel()
        u = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).ravel()
        _check(lib().wr_c_scene_add_mesh(self._h, _ptr(v), v.size, _ptr(i), i.size,
                                         None if n is None else _ptr(n), 0 if n is None else n.size,
                                         None if u is None else _ptr(u), 0 if u is None else u.size,
                                         int(materialId)), "addTriangleMesh")

    def addTextureRGBA8(self, rgba: np.ndarray) -> int:
        a = np.ascontiguousarray(rgba, dtype=np.uint8)
        h, w = a.shape[0], a.shape[1]
        tid = lib().wr_c_scene_add_texture_rgba8(self._h, _ptr(a), w, h)
        if tid == 0:
            raise RuntimeError(f"addTextureRGBA8: {_err()}")
        return tid

    def loadTexture(self, path: str) -> int:
        tid = lib().wr_c_texture_create(self._h, str(path).encode())
        if tid == 0:
            raise RuntimeError(_err())
        return tid

    def destroyTexture(self, textureId: int):
        lib().wr_c_texture_destroy(self._h, int(textureId))

    def setEnvironmentRadiance(self, radiance):
        _check(lib().wr_c_scene_set_environment_radiance(self._h, *map(float, radiance)), "setEnvironmentRadiance")

    def setEnvironmentMap(self, path: str):
        _check(lib().wr_c_scene_set_environment_map(self._h, str(path).encode()), "setEnvironmentMap")

    def finalize(self):
        _check(lib().wr_c_scene_finalize(self._h), "finalize")

    def getBVH(self) -> wrk_bvh:
        b = wrk_bvh()
        _check(lib().wr_c_scene_get_bvh(self._h, C.byref(b)), "getBVH")
        return b

    def getDeviceView(self) -> wrk_scene_view:
        v = wrk_scene_view()
        _check(lib().wr_c_scene_get_view(self._h, C.byref(v)), "getDeviceView")
        return v

    def getBuildSeconds(self) -> float:
        return float(lib().wr_c_scene_build_seconds(self._h))


class Texture2D:
    """libWR/include/wr/scene.h:27-32."""

    @staticmethod
    def create(scene: Scene, path: str) -> int:
        return scene.loadTexture(path)

    @staticmethod
    def destroy(scene: Scene, textureId: int):
        scene.destroyTexture(textureId)


class Renderer:
    """libWR/include/wr/renderer.h:42-65."""

    def __init__(self):
        self._h = lib().wr_c_renderer_create()
        if not self._h:
            raise RuntimeError(f"Renderer: {_err()}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.wr_c_renderer_destroy(h)
            self._h = None

    def setWavefrontConfig(self, cfg: WavefrontConfig):
        flags = (1 if cfg.timeKernels else 0) | (2 if cfg.countTraversal else 0)
        _check(lib().wr_c_renderer_set_wavefront(self._h, int(cfg.iterationCap), int(cfg.maxPathsInFlight),
                                                      int(cfg.maxChunkPaths), flags),
               "setWavefrontConfig")
        _check(lib().wr_c_renderer_set_fused(self._h, int(cfg.fusedMaxPaths)), "setWavefrontConfig")

    def render(self, scene: Scene, camera: Camera, params: RenderParams) -> np.ndarray:
        """Returns the averaged linear image, shape (height, width, 3), float32."""
        out = np.zeros((params.height, params.width, 3), dtype=np.float32)
        cam, prm = camera.to_c(), params.to_c()
        _check(lib().wr_c_render(self._h, scene._h, C.byref(cam), C.byref(prm), _ptr(out)), "render")
        return out

    def loadDenoiser(self, path):
        """Learned denoiser weights (denoiser.UNet.save); render() then denoises
        when params.denoiser.enabled."""
        _check(lib().wr_c_renderer_load_denoiser(self._h, str(path).encode()), "loadDenoiser")

    def renderWithGuides(self, scene: Scene, camera: Camera, params: RenderParams, noise: bool = False):
        """(color, albedo, normal), each (height, width, 3) float32 averaged
        over spp; with noise=True also the (height, width) standard error of
        each pixel's mean (r + g + b) / 3."""
        shape = (params.height, params.width, 3)
        col, alb, nrm = (np.zeros(shape, dtype=np.float32) for _ in range(3))
        nse = np.zeros(shape[:2], dtype=np.float32) if noise else None
        cam, prm = camera.to_c(), params.to_c()
        _check(lib().wr_c_render_guides(self._h, scene._h, C.byref(cam), C.byref(prm), _ptr(col), _ptr(alb),
                                        _ptr(nrm), _ptr(nse) if noise else None), "renderWithGuides")
        return (col, alb, nrm, nse) if noise else (col, alb, nrm)

    def renderPaths(self, scene: Scene, camera: Camera, params: RenderParams, firstPath: int, count: int) -> np.ndarray:
        """Per-path values added to the accumulation buffer, shape (count, 3)."""
        out = np.zeros((count, 4), dtype=np.float32)
        cam, prm = camera.to_c(), params.to_c()
        _check(lib().wr_c_render_paths(self._h, scene._h, C.byref(cam), C.byref(prm), int(firstPath), int(count),
                                       _ptr(out)), "renderPaths")
        return out[:, :3]

    def getStats(self) -> dict:
        s = wr_c_stats()
        _check(lib().wr_c_renderer_stats(self._h, C.byref(s)), "getStats")
        return {name: getattr(s, name) for name, _ in wr_c_stats._fields_}


@dataclass
class ServerConfig:
    """Continuous-batching server (wr::RenderServer): path pool, admission
    chunk, memory budget for accumulation/image buffers (0 = unlimited), and
    the express lane for requests of at most expressMaxPaths paths (0 = off);
    requests of at most fusedMaxPaths paths render as one fused launch per
    chunk in the lane they take (0 = off); with backfillPaths the requests
    behind the oldest unfinished one share at most that many pool entries
    while it has paths in flight (0 = no limit)."""
    maxPathsInFlight: int = 1 << 26
    maxChunkPaths: int = 1 << 24
    memoryBudgetBytes: int = 0
    iterationCap: int = 0
    timeKernels: bool = False
    shortestFirst: bool = False
    roundRobin: bool = False
    expressMaxPaths: int = 0
    expressPathsInFlight: int = 1 << 23
    expressIterations: int = 32
    expressOverlap: bool = False
    fusedMaxPaths: int = 0
    backfillPaths: int = 0


class RenderServer:
    """Requests of different scenes, resolutions and spp share one wavefront;
    submit() queues, step() advances every in-flight request by one bounce
    iteration and admits new ones as path slots and memory free up."""

    def __init__(self, cfg: ServerConfig | None = None):
        c = cfg or ServerConfig()
        cc = wr_c_server_config(int(c.maxPathsInFlight), int(c.maxChunkPaths), int(c.memoryBudgetBytes),
                                int(c.iterationCap), (1 if c.timeKernels else 0) | (2 if c.shortestFirst else 0)
                                | (4 if c.roundRobin else 0) | (8 if c.expressOverlap else 0),
                                int(c.expressMaxPaths), int(c.expressPathsInFlight),
                                int(c.expressIterations), int(c.fusedMaxPaths), int(c.backfillPaths))
        self._h = lib().wr_c_server_create(C.byref(cc))
        if not self._h:
            raise RuntimeError(f"RenderServer: {_err()}")
        self._shapes = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.wr_c_server_destroy(h)
            self._h = None

    def submit(self, scene: Scene, camera: Camera, params: RenderParams) -> int:
        cam, prm = camera.to_c(), params.to_c()
        rid = lib().wr_c_server_submit(self._h, scene._h, C.byref(cam), C.byref(prm))
        if rid == 0:
            raise RuntimeError(f"submit: {_err()}")
        self._shapes[rid] = (params.height, params.width)
        return rid

    def step(self) -> bool:
        r = lib().wr_c_server_step(self._h)
        if r < 0:
            raise RuntimeError(f"step: {_err()}")
        return r == 1

    def status(self, rid: int) -> dict:
        st = wr_c_request_status()
        _check(lib().wr_c_server_status(self._h, rid, C.byref(st)), "status")
        return {name: getattr(st, name) for name, _ in wr_c_request_status._fields_}

    def takeCompleted(self, max_ids: int = 1024) -> list[int]:
        """Ids of requests finished since they were last taken, in completion order."""
        buf = (C.c_uint64 * max_ids)()
        n = lib().wr_c_server_take_completed(self._h, buf, max_ids)
        if n < 0:
            raise RuntimeError(f"takeCompleted: {_err()}")
        return list(buf[:n])

    def fetch(self, rid: int) -> np.ndarray:
        h, w = self._shapes.pop(rid)
        out = np.zeros((h, w, 3), dtype=np.float32)
        _check(lib().wr_c_server_fetch(self._h, rid, _ptr(out)), "fetch")
        return out

    def getStats(self) -> dict:
        s = wr_c_stats()
        _check(lib().wr_c_server_stats(self._h, C.byref(s)), "getStats")
        return {name: getattr(s, name) for name, _ in wr_c_stats._fields_}

    def numPending(self) -> int:
        return int(lib().wr_c_server_pending(self._h))

    def numInFlight(self) -> int:
        return int(lib().wr_c_server_in_flight(self._h))

    @staticmethod
    def clock() -> float:
        return float(lib().wr_c_server_clock())


def aces_tonemap(image: np.ndarray) -> np.ndarray:
    """Tonemap on the GPU (wrk_tonemap_aces) via torch device buffers."""
    import torch

    img = torch.as_tensor(np.ascontiguousarray(image, dtype=np.float32)).cuda()
    out = torch.empty(img.numel(), dtype=torch.uint8, device=img.device)
    _wrk_check(lib().wrk_tonemap_aces(C.c_void_p(img.data_ptr()), img.numel() // 3, C.c_void_p(out.data_ptr()),
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream)), "wrk_tonemap_aces")
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(image.shape)
