"""The C-ABI library loads on a machine without a GPU and exports every entry
point its public headers declare (no compute calls)."""
import re
import subprocess
from pathlib import Path

import offlinerenderer_amd as wr

ROOT = Path(__file__).resolve().parent.parent


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:wrk|wr_c)_[a-z0-9_]+)\s*\(", text)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(wr.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_headers_declare_entry_points():
    assert len(_declared("wrk.h")) >= 18
    assert len(_declared("wr_c.h")) >= 25
    assert len(_declared("wrk_unet.h")) >= 20


def test_library_exports_every_declared_symbol():
    exported = _exported()
    for h in ("wrk.h", "wr_c.h", "wrk_unet.h"):
        missing = [s for s in _declared(h) if s not in exported]
        assert not missing, f"{h}: not exported: {missing}"


def test_ctypes_binds_every_symbol():
    L = wr.lib()
    for name in wr._SIGS:
        assert getattr(L, name) is not None
    # no GPU needed for error plumbing
    assert isinstance(L.wrk_last_error(), bytes)


def test_cpp_api_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", str(wr.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    for sig in ["wr::Context::Context(wr::ContextConfig const&)", "wr::Scene::finalize(wr::SceneBuildConfig const&)",
                "wr::Renderer::render(wr::Scene*, wr::Camera const&, wr::Vec3*, wr::RenderParams const&)",
                "wr::Texture2D::create(wr::Scene*, std::__cxx11::basic_string<char, std::char_traits<char>, "
                "std::allocator<char> > const&)",
                "wr::Scene::addTriangleMesh(std::span<float const, 18446744073709551615ul>, "
                "std::span<unsigned int const, 18446744073709551615ul>, unsigned int)"]:
        assert sig in out, sig


def test_null_handles_are_errors_not_crashes():
    """Every wr_c_* entry point taking a handle or a required pointer rejects
    NULL with an error code and a message (no device work is reached)."""
    import ctypes as C

    L = wr.lib()
    cam, prm = wr.wr_c_camera(), wr.wr_c_render_params()
    bad = 0xFFFFFFFF
    assert L.wr_c_scene_add_lambertian(None, 1.0, 1.0, 1.0, 0) == bad
    assert b"null" in L.wr_c_last_error()
    assert L.wr_c_scene_add_emissive(None, 1.0, 1.0, 1.0) == bad
    assert L.wr_c_scene_finalize(None) == -1
    assert L.wr_c_scene_build_seconds(None) == -1.0
    assert L.wr_c_texture_create(None, b"x.png") == 0
    assert L.wr_c_render(None, None, C.byref(cam), C.byref(prm), None) == -1
    assert L.wr_c_renderer_stats(None, None) == -1
    assert L.wr_c_server_create(None) is None
    assert b"server config" in L.wr_c_last_error()
    assert L.wr_c_server_step(None) == -1
    assert L.wr_c_server_submit(None, None, C.byref(cam), C.byref(prm)) == 0
    assert L.wr_c_server_take_completed(None, None, 4) == -1
    assert L.wr_c_server_pending(None) == 0 and L.wr_c_server_in_flight(None) == 0
    assert L.wrk_unet_init(None, 1) != 0 and b"null" in L.wrk_unet_last_error()
    assert L.wrk_unet_num_params(None) == 0
    # destroying NULL is a no-op, as delete is
    L.wrk_unet_destroy(None)
    L.wr_c_scene_destroy(None)
    L.wr_c_renderer_destroy(None)
    L.wr_c_server_destroy(None)


def test_fused_convolution_arguments_are_checked():
    """The fused convolution entry points reject bad shapes with an error code
    and a message before any device work (so this runs without a GPU)."""
    import ctypes as C

    L = wr.lib()
    p = C.c_void_p(64)  # never dereferenced: every call below fails its checks first

    def err(rc, words):
        assert rc != 0
        msg = L.wrk_last_error()
        assert all(w in msg for w in words), msg

    err(L.wrk_conv3x3_fwd_first(p, 32, 32, p, None, p, 32, 32, 0, 0, 1, 8, 8, None), [b"coutUsed"])
    err(L.wrk_conv3x3_fwd_first(p, 32, 32, p, None, p, 32, 32, 33, 0, 1, 8, 8, None), [b"coutUsed"])
    err(L.wrk_conv3x3_fwd_up(p, 96, 96, None, 64, 64, p, None, p, 32, 32, 1, 1, 8, 8, None), [b"null"])
    err(L.wrk_conv3x3_fwd_up(p, 96, 96, p, 64, 48, p, None, p, 32, 32, 1, 1, 8, 8, None), [b"cUp"])
    err(L.wrk_conv3x3_fwd_up(p, 96, 96, p, 64, 128, p, None, p, 32, 32, 1, 1, 8, 8, None), [b"cUp"])
    err(L.wrk_conv3x3_fwd_up(p, 96, 96, p, 32, 64, p, None, p, 32, 32, 1, 1, 8, 8, None), [b"pitch"])
    err(L.wrk_conv3x3_fwd_up(p, 96, 96, p, 64, 64, p, None, p, 32, 32, 1, 1, 9, 8, None), [b"even"])
    err(L.wrk_conv3x3_fwd_pool(p, 32, 32, p, None, p, 32, 32, 1, p, 16, 1, 8, 8, None), [b"pitch"])
    err(L.wrk_conv3x3_fwd_pool(p, 32, 32, p, None, p, 32, 32, 1, p, 32, 1, 8, 7, None), [b"even"])
    err(L.wrk_conv3x3_fwd_unpack(p, 64, 64, p, None, 32, p, 32, p, 1, 8, 8, None), [b"32 input"])
    err(L.wrk_conv3x3_fwd_unpack(p, 32, 32, p, None, 32, p, 2, p, 1, 8, 8, None), [b"base pitch"])
