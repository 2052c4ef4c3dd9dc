"""DDS cube environment maps and the BC6H decoder (EnvMap::load, src/pathtracing/envmap.rs:12 ->
Texture::create_cubemap, src/common/texture.rs:89-141).  The reference's own DDS files are not in
its repository (.MISSING_LARGE_BLOBS), so the decoder is pinned by:
  * hand-built blocks whose texels follow from the published BC6H rules by hand (unquantisation,
    6-bit interpolation weights, the 31/64 and 31/32 scaling, sign extension, reserved modes);
  * Pillow's independent BC6H decoder over random blocks of all fourteen modes (Pillow returns
    8-bit RGB, floor(255 x value) clamped, computed at a finer precision than the format's
    integer pipeline: texels agree within one 8-bit step, and exactly for all but a few);
  * DDS headers packed here byte by byte (DX10 and legacy FourCC codes, mip chains, the array
    size the reference forces to 6) and round trips through the writer.
"""
import ctypes
import os
import struct

import numpy as np
import pytest

from nbounce_amd._lib import check, lib, ptr

W4 = [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]   # published 4-bit weights


def pack_bits(fields):
    """(value, nbits) fields, lowest block bit first -> 16 bytes."""
    v, pos = 0, 0
    for val, n in fields:
        v |= (val & ((1 << n) - 1)) << pos
        pos += n
    assert pos <= 128
    return v.to_bytes(16, "little")


def decode(block, signed=False):
    out = np.zeros(48, dtype=np.uint16)
    b = np.frombuffer(block, dtype=np.uint8).copy()
    check(lib.nb_bc6h_decode_block(ptr(b), 1 if signed else 0, ptr(out)))
    return out.reshape(16, 3)


def half(bits):
    return np.array(bits, dtype=np.uint16).view(np.float16).astype(np.float64)


def test_mode11_endpoints_unquantise_and_scale():
    """Mode 11 (bits 00011: one region, 10-bit endpoints): a texel with index 0 is endpoint A.
    Unsigned unquantisation maps 0 -> 0 and 1023 -> 0xFFFF, otherwise ((c << 16) + 0x8000) >> 10;
    the result is scaled by 31/64 (truncating) into half bits: 0x3FF gives 0x7BFF = 65504, the
    largest half."""
    blk = pack_bits([(0b00011, 5), (0, 10), (512, 10), (1023, 10), (0, 10), (512, 10), (1023, 10), (0, 63)])
    t = decode(blk)
    g = ((512 << 16) + 0x8000) >> 10
    assert g == 0x8020
    assert (t == [0, (g * 31) >> 6, 0x7BFF]).all()
    assert half([(g * 31) >> 6])[0] == pytest.approx(1.5146484375)


def test_mode11_interpolation_weights():
    """Texel t takes index t (texel 0: index 0 in 3 bits): value = ((A(64 - w) + Bw + 32) >> 6)
    scaled by 31/64, with A = 0, B = 0xFFFF and the published 4-bit weights."""
    fields = [(0b00011, 5)] + [(0, 10)] * 3 + [(1023, 10)] * 3 + [(0, 3)] + [(t, 4) for t in range(1, 16)]
    t = decode(pack_bits(fields))
    want = [((0xFFFF * W4[i] + 32) >> 6) * 31 >> 6 for i in range(16)]
    assert (t == np.array(want)[:, None]).all()
    assert t[15, 0] == 0x7BFF and t[0, 0] == 0


def test_mode1_two_regions_deltas_and_partition():
    """Mode 1 (bits 00: two regions, 10-bit base, 5-bit signed deltas): partition 13 (0xFF00)
    puts rows 0-1 in region 0 and rows 2-3 in region 1; with every index 0 each texel is its
    region's first endpoint: region 0 the base W, region 1 Y = W + sign-extended delta (mod 2^10)."""
    rw, gw, bw = 300, 600, 20
    ry, gy, by = 0b01010, 0b10110, 0b11111            # +10, -10, -1
    f = {"rw": rw, "gw": gw, "bw": bw, "rx": 0, "gx": 0, "bx": 0, "ry": ry, "gy": gy, "by": by, "rz": 0, "gz": 0,
         "bz": 0, "d": 13}
    # mode 1 layout (lowest bit first): m[1:0] gy4 by4 bz4 rw[9:0] gw[9:0] bw[9:0] rx[4:0] gz4 gy[3:0] gx[4:0] bz0
    # gz[3:0] bx[4:0] bz1 by[3:0] ry[4:0] bz2 rz[4:0] bz3 d[4:0]
    bit = lambda k, i: (f[k] >> i) & 1
    fields = [(0, 2), (bit("gy", 4), 1), (bit("by", 4), 1), (bit("bz", 4), 1), (rw, 10), (gw, 10), (bw, 10),
              (f["rx"], 5), (bit("gz", 4), 1), (gy & 15, 4), (f["gx"], 5), (bit("bz", 0), 1), (f["gz"] & 15, 4),
              (f["bx"], 5), (bit("bz", 1), 1), (by & 15, 4), (ry, 5), (bit("bz", 2), 1), (f["rz"], 5),
              (bit("bz", 3), 1), (13, 5), (0, 46)]
    t = decode(pack_bits(fields))

    def unq(c):
        return 0 if c == 0 else (0xFFFF if c == 1023 else ((c << 16) + 0x8000) >> 10)

    r0 = [(unq(c) * 31) >> 6 for c in (rw, gw, bw)]
    r1 = [(unq(c) * 31) >> 6 for c in ((rw + 10) & 1023, (gw - 10) & 1023, (bw - 1) & 1023)]
    assert (t[:8] == r0).all() and (t[8:] == r1).all()


def test_signed_mode11_sign_extends_and_scales_by_31_32():
    """BC6H_SF16: endpoint 0x3FF is -1 (10-bit two's complement); signed unquantisation gives
    -(((1 << 15) + 0x4000) >> 9) = -96 and the 31/32 scaling -(96 * 31 >> 5) = -93: half bits
    0x805D.  Endpoint 511 (the largest positive) unquantises to 0x7FFF -> 0x7BFF."""
    blk = pack_bits([(0b00011, 5), (0x3FF, 10), (511, 10), (0, 10), (0x3FF, 10), (511, 10), (0, 10), (0, 63)])
    t = decode(blk, signed=True)
    assert (t[0] == [0x805D, 0x7BFF, 0]).all()
    assert half([0x805D])[0] < 0


@pytest.mark.parametrize("mode", [0b10011, 0b10111, 0b11011, 0b11111])
def test_reserved_modes_decode_to_zero(mode):
    blk = (int.from_bytes(os.urandom(16), "little") & ~31 | mode).to_bytes(16, "little")
    assert (decode(blk) == 0).all()


MODES = [0x00, 0x01, 0x02, 0x06, 0x0A, 0x0E, 0x12, 0x16, 0x1A, 0x1E, 0x03, 0x07, 0x0B, 0x0F]


@pytest.mark.parametrize("mode", MODES)
def test_bc6h_matches_pillow(mode):
    """Random blocks of one mode against Pillow's BC6H decoder: every texel within one 8-bit step
    (Pillow's finer-precision arithmetic), at most 0.5 % of them off by that step, and a good share
    of texels unsaturated (0 < value < 255) so that the comparison sees the endpoint bits."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(mode)
    nb = 2 if mode < 2 else 5
    off = total = unsat = 0
    for _ in range(300):
        v = int.from_bytes(rng.bytes(16), "little")
        blk = ((v & ~((1 << nb) - 1)) | mode).to_bytes(16, "little")
        pil = np.asarray(Image.frombytes("RGB", (4, 4), blk, "bcn", (6, "BC6H"))).reshape(16, 3).astype(np.int64)
        f = half(decode(blk))
        mine = np.clip(np.floor(np.clip(f, 0, None) * 255), 0, 255).astype(np.int64)
        d = np.abs(pil - mine)
        assert d.max() <= 1, (mode, blk.hex())
        off += int((d > 0).sum())
        total += d.size
        unsat += int(((pil > 0) & (pil < 255)).sum())
    assert off <= 0.005 * total
    assert unsat > 0.15 * total


# ---- DDS files -------------------------------------------------------------------------------

def dds_header(n, fourcc, dxgi=None, mips=1, array_size=1, cube_flag=True):
    """A DDS header packed from the format's definition: magic, 124-byte header, DX10 extension."""
    h = struct.pack("<4sIIIIIII44s", b"DDS ", 124, 0x1 | 0x2 | 0x4 | 0x1000 | 0x20000, n, n, 0, 0, mips, b"\0" * 44)
    h += struct.pack("<II4sIIIII", 32, 0x4, fourcc, 0, 0, 0, 0, 0)
    h += struct.pack("<IIIII", 0x1000 | 0x8 | 0x400000, 0x200 | 0xFC00 if cube_flag else 0, 0, 0, 0)
    assert len(h) == 128
    if dxgi is not None:
        h += struct.pack("<IIIII", dxgi, 3, 0x4 if cube_flag else 0, array_size, 0)
    return h


def cube(n, seed=0):
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.0, 4.0, size=(6, n, n, 4)).astype(np.float32)
    f[..., 3] = 1.0
    return f


def read_mem(data):
    n = ctypes.c_uint32()
    buf = np.frombuffer(data, dtype=np.uint8).copy()
    check(lib.nb_dds_read_cube_memory(ptr(buf), len(data), None, 0, ctypes.byref(n)))
    out = np.zeros((6, n.value, n.value, 4), dtype=np.float32)
    check(lib.nb_dds_read_cube_memory(ptr(buf), len(data), ptr(out), out.size, ctypes.byref(n)))
    return out


def test_rgba32f_dx10_and_legacy_codes():
    f = cube(8)
    for head in (dds_header(8, b"DX10", dxgi=2), dds_header(8, struct.pack("<I", 116))):
        assert np.array_equal(read_mem(head + f.tobytes()), f)
    h16 = f.astype(np.float16)
    for head in (dds_header(8, b"DX10", dxgi=10), dds_header(8, struct.pack("<I", 113))):
        assert np.array_equal(read_mem(head + h16.tobytes()), h16.astype(np.float32))


def test_mip_chains_and_forced_array_size():
    """Faces are stored face-major, each with its full mip chain; mip 0 of each face is read.  The
    reference forces the array size to 6 (texture.rs:95), so a header saying 6 (without the cube
    flag) reads the same six faces."""
    n, mips = 16, 3
    f = cube(n, 1)
    body = b""
    for k in range(6):
        body += f[k].tobytes()
        for m in range(1, mips):
            body += np.full((n >> m, n >> m, 4), 99.0, dtype=np.float32).tobytes()   # other levels
    for head in (dds_header(n, b"DX10", dxgi=2, mips=mips),
                 dds_header(n, b"DX10", dxgi=2, mips=mips, array_size=6, cube_flag=False)):
        assert np.array_equal(read_mem(head + body), f)


def test_bc6h_file_faces_decode_blockwise():
    """A BC6H_UF16 cube: texel (x, y) of face k is texel (x % 4, y % 4) of block (x // 4, y // 4)."""
    n = 8
    rng = np.random.default_rng(5)
    blocks = []
    for _ in range(6 * 4):
        v = int.from_bytes(rng.bytes(16), "little")
        blocks.append(((v & ~31) | MODES[int(rng.integers(2, 14))]).to_bytes(16, "little"))
    got = read_mem(dds_header(n, b"DX10", dxgi=95) + b"".join(blocks))
    for k in range(6):
        for b in range(4):
            t = half(decode(blocks[4 * k + b])).reshape(4, 4, 3)
            by, bx = divmod(b, 2)
            assert np.array_equal(got[k, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4, :3], t.astype(np.float32))
    assert (got[..., 3] == 1).all()


def test_rejected_files():
    f = cube(8)
    bad = [
        (b"XXXX" + dds_header(8, b"DX10", dxgi=2)[4:] + f.tobytes(), "magic"),
        (dds_header(8, b"DX10", dxgi=98) + f.tobytes(), "DXGI format 98"),          # BC7
        (dds_header(8, b"DXT1") + f.tobytes(), "FourCC"),
        (dds_header(8, b"DX10", dxgi=2) + f.tobytes()[:1000], "six faces need"),   # truncated
    ]
    n = ctypes.c_uint32()
    for data, what in bad:
        buf = np.frombuffer(data, dtype=np.uint8).copy()
        assert lib.nb_dds_read_cube_memory(ptr(buf), len(data), None, 0, ctypes.byref(n)) == 7    # NB_ERR_FORMAT
        assert what in lib.nb_last_error().decode()
    assert lib.nb_dds_read_cube(b"/nonexistent/sky.dds", None, 0, ctypes.byref(n)) == 8           # NB_ERR_IO


def test_writer_round_trips(tmp_path):
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    sky = scenegen.sky_cube(32)
    p32, p16, pbc = (str(tmp_path / f"sky{k}.dds") for k in ("32", "16", "bc"))
    nb.write_dds_cube(p32, sky, nb.DDS_RGBA32F)
    nb.write_dds_cube(p16, sky, nb.DDS_RGBA16F)
    nb.write_dds_cube(pbc, sky, nb.DDS_BC6H_UF16)
    assert np.array_equal(nb.read_dds_cube(p32), sky)
    assert np.array_equal(nb.read_dds_cube(p16), sky.astype(np.float16).astype(np.float32))
    assert os.path.getsize(pbc) == 148 + 6 * 64 * 16
    bc = nb.read_dds_cube(pbc)
    rel = np.abs(bc[..., :3] - sky[..., :3]) / np.maximum(sky[..., :3], 1e-3)
    print(f"BC6H mode-11 round trip of the sky: mean rel err {rel.mean():.4f}, max {rel.max():.4f}")
    assert rel.mean() < 0.02 and np.percentile(rel, 99) < 0.1
    # the writer's blocks decode through the same decoder as any file (header pinned above)
    with open(pbc, "rb") as fp:
        data = fp.read()
    assert data[:4] == b"DDS " and struct.unpack("<I", data[128:132])[0] == 95
    blk0 = data[148:164]
    assert np.array_equal(half(decode(blk0)).reshape(4, 4, 3).astype(np.float32), bc[0, :4, :4, :3])
