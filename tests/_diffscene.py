"""Small differentiable scenes for the adjoint tests: closed blobs of cubic
Beziers whose consecutive segments share their endpoint. Parameters are the
unique points (3 per cubic); the stored path points gather from them."""
import numpy as np

from rasterize_amd import LinColor, Path, Scene


def blob_scene(seed, n_paths, n_cubics, size):
    rng = np.random.default_rng(seed)
    uniq = []     # unique points per path, (3 * n_cubics, 2)
    colors = []
    for _ in range(n_paths):
        c = rng.uniform(0.3 * size, 0.7 * size, 2)
        R = rng.uniform(0.12, 0.3) * size
        th = np.sort(rng.uniform(0, 2 * np.pi, 3 * n_cubics))
        r = R * rng.uniform(0.7, 1.3, 3 * n_cubics)
        uniq.append(np.stack([c[0] + r * np.cos(th), c[1] + r * np.sin(th)], 1))
        a = rng.uniform(0.4, 0.9)
        colors.append(np.array([rng.uniform(0, a), rng.uniform(0, a), rng.uniform(0, a), a]))
    return np.stack(uniq), np.stack(colors)


def gather_index(n_paths, n_cubics):
    """Index of the unique point behind each stored point (4 per cubic)."""
    m = 3 * n_cubics
    idx = []
    for k in range(n_paths):
        for i in range(n_cubics):
            idx += [k * m + 3 * i, k * m + 3 * i + 1, k * m + 3 * i + 2, k * m + (3 * i + 3) % m]
    return np.array(idx)


def paths_from(uniq):
    """Path objects (one closed subpath of cubics each) from unique points."""
    n_paths, m, _ = uniq.shape
    nc = m // 3
    out = []
    for k in range(n_paths):
        u = uniq[k]
        pts = []
        for i in range(nc):
            pts += [u[3 * i], u[3 * i + 1], u[3 * i + 2], u[(3 * i + 3) % m]]
        pts = np.array(pts, np.float64).reshape(-1)
        kinds = np.full(nc, 2, np.uint8)
        offs = np.arange(nc + 1, dtype=np.uint32) * 4
        out.append(Path(kinds, offs, pts, np.array([0, nc], np.uint32), np.array([1], np.uint8)))
    return out


def scene_from(uniq, colors, rule=0):
    return Scene.group([Scene.fill(p, LinColor(*c), rule) for p, c in zip(paths_from(uniq), colors)])


def render_f64(O, uniq, colors, W, H, rule=0, flatness=0.05):
    """Scene::render in f64 (full-canvas mask_iter coverage of each path, then
    source-over in f64), for finite differences."""
    from rasterize_amd import Transform
    dst = np.zeros((H, W, 4))
    for p, c in zip(paths_from(uniq), colors):
        a = O.sd_mask_iter(p, Transform.identity(), W, H, rule, flatness)
        s = np.asarray(c, np.float64)[None, None, :] * a[..., None]
        dst = s + dst * (1.0 - s[..., 3:4])
    return dst
