"""Helpers for GPU tests: call the wrk_* C ABI on torch device buffers."""
import ctypes as C

import numpy as np
import torch

import offlinerenderer_amd as wr


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def vp(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def check(rc):
    if rc != 0:
        raise RuntimeError(wr.lib().wrk_last_error().decode())


def trace_closest(bvh, org4: np.ndarray, dir4: np.ndarray, queue=None):
    o, d = dev(org4.astype(np.float32)), dev(dir4.astype(np.float32))
    num_rays = org4.shape[0]
    hits = torch.zeros((num_rays, 2), dtype=torch.int32, device="cuda")
    q = None if queue is None else dev(queue.astype(np.uint32))
    # n counts queue entries; every entry must index a ray below num_rays
    n = num_rays if queue is None else int(queue.size)
    assert queue is None or int(queue.max()) < num_rays
    check(wr.lib().wrk_trace_closest(C.byref(bvh), vp(o), vp(d), None if q is None else vp(q), n, None, vp(hits),
                                     stream()))
    torch.cuda.synchronize()
    h = hits.cpu().numpy().view(np.uint32)
    return h[:, 0].view(np.float32), h[:, 1]


def trace_any(bvh, org4: np.ndarray, dir4: np.ndarray):
    o, d = dev(org4.astype(np.float32)), dev(dir4.astype(np.float32))
    n = org4.shape[0]
    occ = torch.zeros(n, dtype=torch.uint8, device="cuda")
    check(wr.lib().wrk_trace_any(C.byref(bvh), vp(o), vp(d), None, n, None, vp(occ), stream()))
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def product_scene(desc):
    s = wr.Scene()
    desc.build(s)
    s.finalize()
    return s


def camera(desc):
    return wr.Camera(**desc.camera)


def params(desc, width=None, height=None, spp=None):
    return wr.RenderParams(width=width or desc.width, height=height or desc.height, spp=spp or desc.spp,
                           maxBounces=desc.maxBounces, russianRouletteDepth=desc.rrDepth)
