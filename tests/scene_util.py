"""Flatten a SceneDesc into the arrays Scene::finalize builds (scene.cpp:190-244)."""
import numpy as np


def flatten(desc):
    verts, idx, mats = [], [], []
    off = 0
    for op in desc.ops:
        if op[0] != "mesh":
            continue
        v = np.asarray(op[1], dtype=np.float32).reshape(-1, 3)
        i = np.asarray(op[2], dtype=np.uint32).reshape(-1, 3)
        verts.append(v)
        idx.append(i + off)
        mats.append(np.full(i.shape[0], op[5], dtype=np.uint32))
        off += v.shape[0]
    return np.concatenate(verts), np.concatenate(idx), np.concatenate(mats)


def make_rays(desc, n, seed, tmin=0.0):
    """Random rays around the scene plus rays aimed exactly at vertices, edge
    midpoints and centroids (shared-edge ties) and axis-aligned rays."""
    rng = np.random.default_rng(seed)
    V, F, _ = flatten(desc)
    lo, hi = V.min(0), V.max(0)
    ext = hi - lo
    n1 = n // 2
    o1 = rng.uniform(lo - 0.2 * ext, hi + 0.2 * ext, size=(n1, 3))
    d1 = rng.normal(size=(n1, 3))
    n2 = n - n1
    o2 = rng.uniform(lo - 0.2 * ext, hi + 0.2 * ext, size=(n2, 3))
    tri = rng.integers(0, F.shape[0], size=n2)
    kind = rng.integers(0, 4, size=n2)
    a, b, c = V[F[tri, 0]], V[F[tri, 1]], V[F[tri, 2]]
    target = np.where(kind[:, None] == 0, a, np.where(kind[:, None] == 1, (a + b) * 0.5,
                                                       np.where(kind[:, None] == 2, (a + b + c) / 3.0, (b + c) * 0.5)))
    d2 = target - o2
    # a slice of axis-aligned directions (zero components)
    ax = rng.integers(0, 3, size=n2 // 8)
    sgn = rng.choice([-1.0, 1.0], size=n2 // 8)
    d2[: n2 // 8] = 0.0
    d2[np.arange(n2 // 8), ax] = sgn
    o = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    org4 = np.concatenate([o, np.full((n, 1), tmin, np.float32)], axis=1)
    tmax = np.where(rng.uniform(size=n) < 0.2, rng.uniform(0.1, 3.0, size=n) * np.linalg.norm(ext), 1e20)
    dir4 = np.concatenate([d, tmax[:, None].astype(np.float32)], axis=1)
    return np.ascontiguousarray(org4), np.ascontiguousarray(dir4)
