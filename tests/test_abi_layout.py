"""The ctypes mirrors of the C-ABI structs have the layout the C compiler
gives the headers: same size and the same offset for every field (a CPU
test: gcc compiles a probe against include/, nothing touches a GPU)."""
import ctypes as C
import shutil
import subprocess

import pytest

import offlinerenderer_amd as wr
from tests.test_abi import ROOT

MIRRORS = [wr.wrk_bvh, wr.wrk_scene_view, wr.wrk_material, wr.wrk_paths, wr.wr_c_camera, wr.wr_c_render_params,
           wr.wr_c_stats, wr.wr_c_server_config, wr.wr_c_request_status]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_c_layout(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "wr_c.h"', "int main(void) {"]
    for s in MIRRORS:
        t = s.__name__
        lines.append(f'    printf("{t} size %zu\\n", sizeof({t}));')
        for name, *_ in s._fields_:
            lines.append(f'    printf("{t} {name} %zu\\n", offsetof({t}, {name}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        t, name, v = line.split()
        got[(t, name)] = int(v)
    for s in MIRRORS:
        t = s.__name__
        assert got[(t, "size")] == C.sizeof(s), f"{t}: C size {got[(t, 'size')]} vs ctypes {C.sizeof(s)}"
        for name, *_ in s._fields_:
            assert got[(t, name)] == getattr(s, name).offset, f"{t}.{name}: offset differs"
