"""C++ programs written against include/wr/wr.h (libWR's API) compile, link
against the library and, on a GPU, render a plausible Cornell box and serve
requests through wr::RenderServer with images equal to stand-alone renders."""
import subprocess
from pathlib import Path

import pytest

import offlinerenderer_amd as wr

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "cornell_drop_in.cpp"
SERVER_SRC = ROOT / "tests" / "cpp" / "server_drop_in.cpp"


def _compile(tmp_path, src=SRC):
    exe = tmp_path / src.stem
    lib_dir = wr.LIB_PATH.parent
    cmd = ["g++", "-std=c++20", "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe), f"-L{lib_dir}",
           "-lwr_amd", f"-Wl,-rpath,{lib_dir}", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_compiles_and_links(tmp_path):
    assert _compile(tmp_path).exists()
    assert _compile(tmp_path, SERVER_SRC).exists()


@pytest.mark.gpu
def test_runs_on_gpu(tmp_path):
    exe = _compile(tmp_path)
    out = subprocess.run([str(exe), "64", "16"], check=True, capture_output=True, text=True, timeout=120).stdout
    vals = [float(x) for x in out.split()[1:4]]
    # warm light (17,12,4) on white/red/green walls: red channel brightest
    assert vals[0] > vals[1] > vals[2] > 0.01


@pytest.mark.gpu
def test_server_loop_runs_on_gpu(tmp_path):
    exe = _compile(tmp_path, SERVER_SRC)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "equal 6/6" in r.stdout, r.stdout + r.stderr
