"""CPU checks of the C ABI library and the host-side logic (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import rasterize_amd as rz
from rasterize_amd import PathBuilder, _abi


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    declared = _abi.declared_symbols()
    assert {"rz_mask", "rz_flatten", "rz_bin", "rz_mask_device"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = sorted(declared - exported)
    assert not missing, f"declared but not exported: {missing}"
    for name in declared:
        getattr(lib, name)


def test_prototypes_cover_declarations():
    bound = {name for name, _, _ in _abi.PROTOTYPES}
    assert _abi.declared_symbols() <= bound


def _rnd_python(seed, n):
    """benches/scene_bench.rs:51-112, restated from its definition."""
    state = seed
    out = []

    def step():
        nonlocal state
        state = (state * 214013 + 2531011) & 0x7FFFFFFF
        return state >> 16

    def u32():
        hi = step() & 0xFFFF
        lo = step() & 0xFFFF
        return (hi << 16) | lo

    for _ in range(n):
        u64 = (u32() << 32) | u32()
        out.append((u64 >> 10) * 2.0 ** -53)
    return np.array(out)


def test_lcg_matches_its_definition():
    for seed in (0, 1, 2, 12345):
        assert np.array_equal(rz.rnd_uniform(seed, 500), _rnd_python(seed, 500))
    u = rz.rnd_uniform(3, 20000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_workloads_are_deterministic_and_shaped():
    a = rz.workload_cubics64(1, 512.0).paths[0]
    b = rz.workload_cubics64(1, 512.0).paths[0]
    assert np.array_equal(a.points, b.points)
    assert a.segments_count() == 64 and np.all(a.kinds == _abi.RZ_CUBIC) and len(a) == 1
    assert a.closed[0] == 1
    pts = a.points.reshape(-1, 2)
    assert pts.min() >= 0.0 and pts.max() < 512.0
    assert np.array_equal(pts[0], pts[-1])  # last endpoint is the first point
    w = rz.workload_random_walk(2, 20_000, 4096.0)
    p = w.paths[0]
    assert len(p) == 40  # a new subpath every 500 commands
    kinds = np.bincount(p.kinds, minlength=3)
    assert kinds[0] > 0.35 * 20_000 and kinds[1] > 0.25 * 20_000 and kinds[2] > 0.25 * 20_000
    m = w.transform.m
    scale = np.hypot(m[0], m[3])
    assert 0.5 <= scale <= 2.0


def test_path_builder_semantics():
    # src/path.rs:895-903: line_to a point equal to the position adds nothing;
    # close() sets the position to the subpath start
    p = (PathBuilder().move_to((1.0, 1.0)).line_to((1.0, 1.0)).line_to((3.0, 1.0))
         .line_to((3.0, 4.0)).close().line_to((0.0, 0.0)).build())
    assert p.segments_count() == 3
    assert list(p.subpaths) == [0, 2, 3] and list(p.closed) == [1, 0]
    assert tuple(p.points[8:10]) == (1.0, 1.0)  # third segment starts at the subpath start
    c = PathBuilder().move_to((10.0, 10.0)).circle(5.0).build()
    assert c.segments_count() == 4 and list(c.closed) == [1]


def test_sin_cos_is_glibc_sincos():
    libm = C.CDLL("libm.so.6")
    libm.sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = np.random.default_rng(0)
    for a in rng.uniform(-10, 10, 1000):
        s, c = C.c_double(), C.c_double()
        libm.sincos(a, C.byref(s), C.byref(c))
        assert rz.sin_cos(a) == (s.value, c.value)


def test_lincolor_from_rgba8_matches_oracle():
    """LinColor::from(RGBA) (src/color.rs:394-406) on the host, bit-exact."""
    import ctypes as C

    import _oracle as O
    from rasterize_amd import LinColor
    rng = np.random.default_rng(0)
    for rgba in [(0x89, 0x15, 0x51, 0x80), (255, 255, 255, 255), (0, 0, 0, 0), (10, 11, 12, 13)] + \
            [tuple(int(v) for v in rng.integers(0, 256, 4)) for _ in range(50)]:
        out = (C.c_float * 4)()
        O.lib().orc_rgba_to_lin((C.c_uint8 * 4)(*rgba), out)
        assert list(LinColor.from_rgba8(*rgba)) == list(out), rgba


def test_paint_abi_layout_matches_oracle():
    """rz_paint and the oracle's paint struct share one layout, so a paint
    built by the package evaluates identically in the oracle."""
    import ctypes as C

    import _oracle as O
    from rasterize_amd import GradRadial, LinColor, _abi
    assert C.sizeof(_abi.rz_paint) == C.sizeof(O.orc_paint)
    for (n1, t1), (n2, t2) in zip(_abi.rz_paint._fields_, O.orc_paint._fields_):
        assert n1 == n2 and getattr(_abi.rz_paint, n1).offset == getattr(O.orc_paint, n2).offset
    g = GradRadial([(0.0, LinColor(1, 0, 0, 1)), (1.0, LinColor(0, 0, 1, 1))], center=(0.5, 0.0),
                   radius=0.5, fcenter=(0.25, 0.0), fradius=0.1, linear_colors=True)
    assert np.allclose(O.paint_at(g.abi(), 0.675, 0.0), [0.5, 0.0, 0.5, 1.0], atol=1e-6)
