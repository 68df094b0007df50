"""Glyph atlas on the GPU: Rasterizer::mask of many independent glyph paths,
each into its own view of one atlas (src/image.rs:300-311), in one pipeline.
Every view is checked against the oracle's SignedDifferenceRasterizer::mask
of that glyph alone; pixels outside the views must be left untouched."""
import sys

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import FillRule, Transform, atlas_slot, mask_atlas, workload_glyphs

pytestmark = pytest.mark.gpu

MAX_ABS = 2e-3
MEAN_ABS = 1e-5


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def glyphs():
    return workload_glyphs(4, 300, 1024)


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_atlas_matches_per_glyph_masks(glyphs, rule):
    W, H = glyphs.atlas_size
    atlas = np.full((H, W), -1.0, np.float32)
    _say(f"atlas {len(glyphs.paths)} glyphs rule={rule} {W}x{H}")
    mask_atlas(glyphs.paths, glyphs.slots, atlas, rule)
    inside = np.zeros((H, W), bool)
    worst, total, count = 0.0, 0.0, 0
    for path, sl in zip(glyphs.paths, glyphs.slots):
        ref = O.sd_mask(path, Transform.identity(), sl.width, sl.height, rule)
        got = atlas[sl.y:sl.y + sl.height, sl.x:sl.x + sl.width]
        d = np.abs(got.astype(np.float64) - ref)
        worst = max(worst, float(d.max()))
        total += float(d.sum())
        count += d.size
        inside[sl.y:sl.y + sl.height, sl.x:sl.x + sl.width] = True
    _say(f"atlas rule={rule}: max|d|={worst:.3e} mean|d|={total / count:.3e}")
    assert worst <= MAX_ABS and total / count <= MEAN_ABS
    assert np.all(atlas[~inside] == -1.0), "pixels outside every glyph view changed"


def test_atlas_transforms_and_edges():
    """Per-glyph transforms that push glyphs across their views' edges (left
    split, right-edge column, rows above and below)."""
    ws = workload_glyphs(9, 40, 512)
    rng = np.random.default_rng(3)
    slots = []
    for sl in ws.slots:
        tr = Transform.new_translate(float(rng.uniform(-12, 12)), float(rng.uniform(-12, 12))) * \
            Transform.new_rotate(float(rng.uniform(0, 6.28)))
        slots.append(atlas_slot(sl.x, sl.y, sl.width, sl.height, tr))
    W, H = ws.atlas_size
    atlas = np.zeros((H, W), np.float32)
    mask_atlas(ws.paths, slots, atlas, FillRule.NonZero)
    for path, sl in zip(ws.paths, slots):
        tr = Transform(*sl.tr.m)
        ref = O.sd_mask(path, tr, sl.width, sl.height, FillRule.NonZero)
        got = atlas[sl.y:sl.y + sl.height, sl.x:sl.x + sl.width]
        d = np.abs(got.astype(np.float64) - ref)
        assert d.max() <= MAX_ABS, f"glyph at {sl.x},{sl.y}: max {d.max()}"


def test_device_glyphs_equal_blocking(glyphs):
    import torch
    W, H = glyphs.atlas_size
    ref = np.zeros((H, W), np.float32)
    mask_atlas(glyphs.paths, glyphs.slots, ref, FillRule.NonZero)
    dg = rz.DeviceGlyphs(glyphs.paths, glyphs.slots)
    out = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    for _ in range(4):
        dg.mask_device(out.data_ptr(), W)
        s = rz._abi.lib().rz_sync(dg.ctx.h)
        if s != rz._abi.RZ_E_CAPACITY:
            break
    dg.ctx.check(s)
    got = out.cpu().numpy()
    dg.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
