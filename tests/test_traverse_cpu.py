"""The product's segment traversal (rz_geom.hpp: virtual-subtree walk aside,
the DFS with its register stack and recompute-from-root, PathFlattenIter,
src/path.rs:758-796) built for the host and compared bit for bit with the
oracle's flatten of the same single segments: every quad and cubic of a
sample of the 200k-segment random walk, at three tolerances."""
import os
import struct
import subprocess

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import PathBuilder

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("trav") / "cpu_traverse")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpu_traverse.cpp")], check=True)
    return exe


def _segments(n_curves):
    ws = rz.workload_random_walk(2, 20_000, 4096.0)
    path, tr = ws.paths[0], ws.transform
    kinds = np.asarray(path.kinds)
    offs = np.asarray(path.point_offsets)
    pts = np.asarray(path.points, np.float64).reshape(-1, 2)
    segs = []
    for k in range(len(kinds)):
        if kinds[k] in (1, 2):
            segs.append((int(kinds[k]), pts[offs[k]:offs[k] + kinds[k] + 2]))
            if len(segs) == n_curves:
                break
    return segs, tr


@pytest.mark.parametrize("tol", [0.05, 0.01, 0.25])
def test_traversal_matches_oracle_flatten(driver, tol):
    segs, tr = _segments(600)
    m = [tr.abi().m[i] for i in range(6)]
    blob = struct.pack("<I6dd", len(segs), *m, tol)
    for kind, q in segs:
        blob += struct.pack("<B", kind) + q.astype("<f8").tobytes()
    out = subprocess.run([driver], input=blob, stdout=subprocess.PIPE, check=True).stdout
    pos, leaves = 0, 0
    for kind, q in segs:
        st, cnt = struct.unpack_from("<II", out, pos)
        pos += 8
        got = np.frombuffer(out, "<f8", 4 * cnt, pos).reshape(cnt, 4)
        pos += 32 * cnt
        b = PathBuilder().move_to(tuple(q[0]))
        b = b.quad_to(tuple(q[1]), tuple(q[2])) if kind == 1 else b.cubic_to(tuple(q[1]), tuple(q[2]), tuple(q[3]))
        ref = O.flatten(b.build(), tr, tol, close=False)
        assert st == 0
        assert got.shape == ref.shape and np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (kind, q)
        leaves += cnt
    assert pos == len(out) and leaves > 10 * len(segs)
