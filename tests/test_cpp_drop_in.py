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


def test_camera_matrices_and_matrix_types(tmp_path):
    """Camera::getViewMatrix / getProjectionMatrix (libWR/include/wr/types.h:28-35,
    glm::lookAt and glm::perspective: right-handed, depth to [-1, 1]) and the
    Mat3 / Mat4 types (types.h:12-13) in a C++ caller, checked against what
    those matrices must do: the eye maps to the origin, the camera basis to
    the axes (forward to -z), near and far planes to depth -1 and +1, the
    corner of the field of view to (1, 1), the target to the image centre."""
    src = ROOT / "tests" / "cpp" / "camera_matrices.cpp"
    exe = tmp_path / "camera_matrices"
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = {ln.split()[0]: [float(x) for x in ln.split()[1:] if x[0].isdigit() or x[0] == "-"]
            for ln in out.splitlines()}
    import math

    import numpy as np

    dist = math.dist((1.0, 2.0, 3.0), (0.5, 1.0, -1.0))
    assert np.allclose(rows["eye"], [0, 0, 0, 1], atol=1e-6)
    assert np.allclose(rows["target"], [0, 0, -dist, 1], atol=1e-5)
    assert np.allclose(rows["right"], [1, 0, 0, 0], atol=1e-6)
    assert np.allclose(rows["up"], [0, 1, 0, 0], atol=1e-6)
    assert np.allclose(rows["forward"], [0, 0, -1, 0], atol=1e-6)
    ndc = lambda v: np.array(v[:3]) / v[3]  # noqa: E731
    assert np.allclose(ndc(rows["near"]), [0, 0, -1], atol=1e-6)
    assert np.allclose(ndc(rows["far"]), [0, 0, 1], atol=1e-6)
    assert np.allclose(ndc(rows["corner"])[:2], [1, 1], atol=1e-6)
    assert np.allclose(ndc(rows["pv"])[:2], [0, 0], atol=1e-6)
    assert rows["mat3"][:3] == [1, 2, 3] and rows["mat3"][3:] == [36, 64, 16]
