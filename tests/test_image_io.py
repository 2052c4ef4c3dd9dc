"""Image files of the accumulation target (nb_image_write_pfm / nb_image_write_exr), host-only:
each file is parsed back by an independent reader written from the published formats and
compared bit for bit with the image written."""
import struct

import numpy as np
import pytest

import nbounce_amd as nb
from nbounce_amd._lib import NbError


def _image(h=13, w=21, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.normal(scale=3.0, size=(h, w, 4)).astype(np.float32)
    img[0, 0, :3] = [np.inf, -0.0, 1e-42]        # infinity, negative zero, a subnormal
    img[1, 1, 0] = 70000.0                        # above the largest half
    return img


def read_pfm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    scale = float(parts[2])
    assert scale < 0                               # little-endian
    px = np.frombuffer(parts[3], dtype="<f4").reshape(h, w, 3)
    return px                                      # PFM rows run bottom to top


def read_exr(path):
    """Minimal OpenEXR 2 reader: single-part scanline, NO_COMPRESSION, HALF or FLOAT channels."""
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == 20000630 and (version & 0xFF) == 2 and not (version & 0x1E00)
    off, attrs = 8, {}
    while b[off] != 0:
        name_end = b.index(b"\0", off)
        name = b[off:name_end].decode()
        type_end = b.index(b"\0", name_end + 1)
        typ = b[name_end + 1:type_end].decode()
        (size,) = struct.unpack_from("<i", b, type_end + 1)
        attrs[name] = (typ, b[type_end + 5:type_end + 5 + size])
        off = type_end + 5 + size
    off += 1
    chans, raw = [], attrs["channels"][1]
    p = 0
    while raw[p] != 0:
        e = raw.index(b"\0", p)
        cname = raw[p:e].decode()
        ptype, _plin, xs, ys = struct.unpack_from("<iB3xii", raw, e + 1)
        assert xs == 1 and ys == 1
        chans.append((cname, ptype))
        p = e + 1 + 16
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    table = struct.unpack_from(f"<{h}Q", b, off)
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for y in range(h):
        o = table[y]
        line_y, nbytes = struct.unpack_from("<ii", b, o)
        o += 8
        for c, ptype in chans:
            dt = {1: "<f2", 2: "<f4"}[ptype]
            size = np.dtype(dt).itemsize * w
            out[c][line_y - y0] = np.frombuffer(b, dtype=dt, count=w, offset=o).astype(np.float32)
            o += size
        assert o - table[y] - 8 == nbytes
    return out, chans


def test_pfm_round_trip(tmp_path):
    img = _image()
    path = tmp_path / "frame.pfm"
    nb.write_pfm(path, img)
    got = read_pfm(path)
    np.testing.assert_array_equal(got.view(np.uint32), img[..., :3].view(np.uint32))


@pytest.mark.parametrize("half", [False, True])
def test_exr_round_trip(tmp_path, half):
    img = _image(seed=1 if half else 2)
    path = tmp_path / "frame.exr"
    nb.write_exr(path, img, half=half)
    chans, names = read_exr(path)
    assert [c for c, _ in names] == ["A", "B", "G", "R"]
    assert all(t == (1 if half else 2) for _, t in names)
    want = img[::-1]                                   # scanline 0 is the top: the last target row
    if half:
        with np.errstate(over="ignore"):
            want = want.astype(np.float16).astype(np.float32)   # numpy rounds to nearest even
    for k, c in enumerate("RGBA"):
        np.testing.assert_array_equal(chans[c].view(np.uint32), np.ascontiguousarray(want[..., k]).view(np.uint32))


def test_half_conversion_of_every_edge_case(tmp_path):
    """The writer's binary32 -> binary16 rounding equals numpy's on values chosen around every
    boundary: subnormal halves, the rounding tie at each exponent, overflow, NaN."""
    v = np.array([0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25,
                  3 * 2.0 ** -26, 2.0 ** -26, 2.0 ** -14, 2.0 ** -14 - 2.0 ** -25, 1 + 2.0 ** -11,
                  1 + 3 * 2.0 ** -11, 5.96e-8, 6.1e-5, np.inf, -np.inf], dtype=np.float32)
    rng = np.random.default_rng(7)
    r = (rng.normal(size=4000) * np.exp2(rng.uniform(-30, 18, 4000))).astype(np.float32)
    vals = np.concatenate([v, r, np.nextafter(v, np.float32(np.inf)), np.nextafter(v, np.float32(-np.inf))])
    n = (len(vals) + 3) // 4 * 4
    img = np.zeros(n, np.float32)
    img[:len(vals)] = vals
    img = img.reshape(1, n // 4, 4)
    path = tmp_path / "edge.exr"
    nb.write_exr(path, img, half=True)
    chans, _ = read_exr(path)
    got = np.stack([chans[c][0] for c in "RGBA"], -1).reshape(-1)
    with np.errstate(over="ignore"):
        want = img.reshape(-1).astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # NaN stays NaN
    img2 = np.full((1, 1, 4), np.nan, np.float32)
    nb.write_exr(path, img2, half=True)
    chans, _ = read_exr(path)
    assert all(np.isnan(chans[c][0, 0]) for c in "RGBA")


def test_writer_errors(tmp_path):
    img = _image()
    with pytest.raises(NbError):
        nb.write_exr(tmp_path / "no_such_dir" / "x.exr", img)
    with pytest.raises(NbError):
        nb.write_pfm(tmp_path / "no_such_dir" / "x.pfm", img)
    with pytest.raises(ValueError):
        nb.write_pfm(tmp_path / "x.pfm", img[..., :3])
    from nbounce_amd._lib import lib, ptr
    assert lib.nb_image_write_exr(str(tmp_path / "y.exr").encode(), ptr(img), img.shape[1], img.shape[0], 6) != 0
    assert lib.nb_image_write_pfm(str(tmp_path / "y.pfm").encode(), ptr(img), 0, 5) != 0
    assert lib.nb_image_write_pfm(None, ptr(img), 4, 4) != 0
