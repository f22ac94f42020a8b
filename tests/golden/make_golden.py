"""Writes tests/golden/*.npz: per-path oracle outcomes for fixed (pixel,
sample) sets, so later changes to the oracle or the kernels are caught
against a frozen record.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle as po  # noqa: E402
from offlinerenderer_amd import scenes  # noqa: E402

OUT = Path(__file__).resolve().parent


def cornell_paths():
    d = scenes.cornell(64, 64, 4)
    s = po.build_oracle_scene(d, bvh=False)
    rng = np.random.default_rng(2024)
    g = np.sort(rng.choice(64 * 64 * 4, 4096, replace=False)).astype(np.uint64)
    prm = po.params_in(64, 64, 8, 3.0, iterationCap=8)
    L, info = po.render_paths(s, d.camera, prm, (g % 4096).astype(np.uint32), (g // 4096).astype(np.uint32),
                              use_bvh=False)
    np.savez_compressed(OUT / "cornell64_paths.npz", path=g, L=L, info=info)


def material_paths():
    d = scenes.material_scene(96, 54, 2, tex_size=64)
    s = po.build_oracle_scene(d, bvh=True)
    npix = 96 * 54
    rng = np.random.default_rng(2025)
    g = np.sort(rng.choice(npix * 2, 2048, replace=False)).astype(np.uint64)
    prm = po.params_in(96, 54, 16, 3.0, iterationCap=0)
    L, info = po.render_paths(s, d.camera, prm, (g % npix).astype(np.uint32), (g // npix).astype(np.uint32),
                              use_bvh=True)
    np.savez_compressed(OUT / "material96_paths.npz", path=g, L=L, info=info)


if __name__ == "__main__":
    cornell_paths()
    material_paths()
    print("golden fixtures written")
