"""Output encoders (the step after the fine stage): RGBA::from(LinColor)
(src/color.rs:164-175) on the GPU against the oracle bit for bit, and the BMP
(src/image.rs:164-213) and PNG (src/image.rs:215-235) containers. The
containers are host code (CPU tests); the pixel conversion is a kernel (GPU
tests)."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import _abi


def _png_decode(data):
    """A minimal PNG reader (8-bit RGBA, filter 0 rows) written from the PNG
    specification, checking every chunk's CRC."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, w, h = 8, b"", 0, 0
    while p < len(data):
        n, = struct.unpack(">I", data[p:p + 4])
        kind, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        crc, = struct.unpack(">I", data[p + 8 + n:p + 12 + n])
        assert zlib.crc32(kind + body) & 0xFFFFFFFF == crc, kind
        if kind == b"IHDR":
            w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
        p += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (300, 250), (0, 5)])
def test_png_container_round_trip(w, h):
    rng = np.random.default_rng(w * 1000 + h)
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    L = _abi.lib()
    n = C.c_size_t()
    src = px.ctypes.data_as(C.POINTER(C.c_uint8))
    assert L.rz_encode_png(src, w, h, 4 * w, None, 0, C.byref(n)) == 0
    buf = (C.c_uint8 * n.value)()
    assert L.rz_encode_png(src, w, h, 4 * w, buf, n.value - 1, C.byref(n)) == _abi.RZ_E_CAPACITY
    assert L.rz_encode_png(src, w, h, 4 * w, buf, n.value, C.byref(n)) == 0
    got = _png_decode(bytes(buf))
    assert got.shape == (h, w, 4) and np.array_equal(got, px)


def test_png_many_stored_blocks():
    """Rows longer than one stored block (65535 bytes)."""
    px = np.arange(200 * 180 * 4, dtype=np.uint32).astype(np.uint8).reshape(180, 200, 4)
    L = _abi.lib()
    n = C.c_size_t()
    src = px.ctypes.data_as(C.POINTER(C.c_uint8))
    L.rz_encode_png(src, 200, 180, 800, None, 0, C.byref(n))
    buf = (C.c_uint8 * n.value)()
    assert L.rz_encode_png(src, 200, 180, 800, buf, n.value, C.byref(n)) == 0
    assert np.array_equal(_png_decode(bytes(buf)), px)


def test_bmp_header_fields():
    """The fields ImageRead::write_bmp writes (src/image.rs:173-205)."""
    hdr = (C.c_uint8 * 122)()
    assert _abi.lib().rz_bmp_header(5, 3, hdr) == 0
    b = bytes(hdr)
    assert b[:2] == b"BM"
    size, _, off, dib, w, h, planes, bpp, comp, dsize, xr, yr = struct.unpack("<IIIIiiHHIIii", b[2:46])
    assert (size, off, dib, w, h, planes, bpp, comp, dsize) == (122 + 60, 122, 108, 5, 3, 1, 32, 3, 60)
    assert (xr, yr) == (2835, 2835)
    assert b[46:54] == bytes(8)
    masks = struct.unpack("<IIII", b[54:70])
    assert masks == (0x000000FF, 0x0000FF00, 0x00FF0000, 0xFF000000)  # R, G, B, A byte order
    assert b[70:74] == bytes([0x42, 0x47, 0x52, 0x73]) and b[74:] == bytes(48)


def _special_colours(rng, n):
    c = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    c[:, :3] *= c[:, 3:4]  # premultiplied
    sp = np.array([[0, 0, 0, 0], [1, 1, 1, 1], [0.5, 0.2, 0.1, 1e-6], [0.5, 0.2, 0.1, 2e-6],
                   [2.0, 0.5, -0.1, 1.0], [0.3, 0.3, 0.3, 1.5], [1e-9, 1e-9, 1e-9, 1e-8],
                   [np.nan, 0.2, 0.2, 0.5], [0.0031308, 0.0031309, 0.002, 1.0],
                   [np.inf, 0.1, 0.1, 1.0]], np.float32)
    return np.concatenate([sp, c])


@pytest.mark.gpu
def test_encode_rgba8_matches_oracle():
    rng = np.random.default_rng(3)
    cols = _special_colours(rng, 300 * 200 - 10)
    img = cols.reshape(200, 300, 4)
    got = rz.encode_rgba8(img)
    ref = np.zeros((len(cols), 4), np.uint8)
    L = O.lib()
    for i, c in enumerate(cols):
        o = (C.c_uint8 * 4)()
        L.orc_lin_to_rgba((C.c_float * 4)(*[float(v) for v in c]), o)
        ref[i] = list(o)
    assert np.array_equal(got.reshape(-1, 4), ref)
    flipped = rz.encode_rgba8(img, bottom_up=True)
    assert np.array_equal(flipped, got[::-1])


@pytest.mark.gpu
def test_rendered_layer_to_bmp_and_png(tmp_path):
    ws = rz.workload_scene(3, 40, 512.0)
    lay = ws.scene().render(rz.SignedDifferenceRasterizer(), rz.Transform.new_scale(0.25, 0.25),
                            (0.0, 0.0, 128.0, 96.0))
    px = lay.rgba8()
    bmp = rz.bmp_bytes(lay.image)
    assert len(bmp) == 122 + 128 * 96 * 4
    body = np.frombuffer(bmp[122:], np.uint8).reshape(96, 128, 4)
    assert np.array_equal(body, px[::-1])
    lay.write_png(tmp_path / "layer.png")
    assert np.array_equal(_png_decode((tmp_path / "layer.png").read_bytes()), px)
    assert px[..., 3].max() > 0
