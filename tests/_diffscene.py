"""Small differentiable scenes for the adjoint tests (closed blobs of cubics
whose consecutive segments share their endpoint) and an f64 reference render
for finite differences."""
import numpy as np

from rasterize_amd import DiffScene, Transform


def blob_scene(seed, n_paths, n_cubics, size):
    rng = np.random.default_rng(seed)
    uniq = []
    colors = []
    for _ in range(n_paths):
        c = rng.uniform(0.3 * size, 0.7 * size, 2)
        R = rng.uniform(0.12, 0.3) * size
        th = np.sort(rng.uniform(0, 2 * np.pi, 3 * n_cubics))
        r = R * rng.uniform(0.7, 1.3, 3 * n_cubics)
        uniq.append(np.stack([c[0] + r * np.cos(th), c[1] + r * np.sin(th)], 1))
        a = rng.uniform(0.4, 0.9)
        colors.append(np.array([rng.uniform(0, a), rng.uniform(0, a), rng.uniform(0, a), a]))
    return DiffScene(np.concatenate(uniq), np.stack(colors), n_cubics)


def render_f64(O, ds, points, colors, W, H, rule=0, flatness=0.05):
    """Scene::render in f64: full-canvas mask_iter coverage of each path (the
    oracle), then source-over in f64 (finite-difference reference)."""
    dst = np.zeros((H, W, 4))
    for p, c in zip(ds.paths(points), np.asarray(colors, np.float64).reshape(-1, 4)):
        a = O.sd_mask_iter(p, Transform.identity(), W, H, rule, flatness)
        s = c[None, None, :] * a[..., None]
        dst = s + dst * (1.0 - s[..., 3:4])
    return dst


def path_of(ds, points, k):
    """Path k of a DiffScene from unique points `points` (only that path)."""
    from rasterize_amd import Path
    nc = ds.n_cubics
    stored = np.asarray(points, np.float64).reshape(-1, 2)[ds.index[4 * nc * k: 4 * nc * (k + 1)]]
    return Path(np.full(nc, 2, np.uint8), np.arange(nc + 1, dtype=np.uint32) * 4,
                stored.reshape(-1), np.array([0, nc], np.uint32), np.array([1], np.uint8))


def local_index(ds, k):
    """For path k's stored points, the index of their unique point within the
    path's own unique points."""
    nc = ds.n_cubics
    return ds.index[4 * nc * k: 4 * nc * (k + 1)] - 3 * nc * k
