"""PNG diffuse maps (vct_png_load) as GLHelper::createTextureFromImage uploads
them (src/Graphics/GLHelper.cpp:165-210): stb_image's decode with the file's
own channel count, glTextureSubImage2D at the default unpack alignment of 4,
GL_R8 / GL_RGB8 / GL_RGBA8 (2 channels: no storage), and a box mip chain.

The files are written here by a PNG encoder of the test's own (zlib + the
five row filters, chosen per row, and Adam7 interlacing) from known pixel
arrays, so every decoded texel has a known answer: the array itself, widened
as the format says (gray d-bit samples times 255 / (2^d - 1), 16-bit samples
to their high byte, palette entries, tRNS alpha).  The mip levels are checked
against the 2 x 2 box mean rounded half up, computed here in integers.
"""
import struct
import zlib

import numpy as np
import pytest

import vct_amd


def chunk(tag, body):
    return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xffffffff)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def filter_rows(rows, bpp, kinds):
    """rows: list of bytes (unfiltered); kinds: a filter type per row."""
    out = b""
    prev = bytes(len(rows[0])) if rows else b""
    for r, f in zip(rows, kinds):
        enc = bytearray()
        for i in range(len(r)):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, paeth(a, b, c)][f]
            enc.append((r[i] - pred) & 0xFF)
        out += bytes([f]) + bytes(enc)
        prev = r
    return out


def pack_row(samples, depth):
    """samples: 1-D int array of one row's samples."""
    if depth == 8:
        return bytes(np.asarray(samples, np.uint8))
    if depth == 16:
        return np.asarray(samples, ">u2").tobytes()
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def write_png(path, img, ctype, depth, plte=None, trns=None, interlace=False, seed=0):
    """img: (h, w, samples) integer array of the file's samples."""
    h, w, ns = img.shape
    rng = np.random.default_rng(seed)
    bpp = max(1, ns * depth // 8)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        raw += filter_rows(rows, bpp, rng.integers(0, 5, len(rows)))
    body = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", body)
    if plte is not None:
        data += chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        data += chunk(b"tRNS", bytes(trns))
    comp = zlib.compress(raw, 6)
    data += chunk(b"IDAT", comp[:len(comp) // 2]) + chunk(b"IDAT", comp[len(comp) // 2:]) + chunk(b"IEND", b"")
    path.write_bytes(data)
    return path


def rgba_of(tex, w, h):
    t = np.asarray(tex[:w * h], np.uint32).reshape(h, w)
    return np.stack([(t >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.int64)


def as_gl(rgba8, channels):
    """The level-0 texels GL samples from a tightly packed stb_image buffer of
    `channels` channels read with rows every round_up(w * channels, 4) bytes."""
    h, w, _ = rgba8.shape
    buf = rgba8[..., :channels].astype(np.uint8).reshape(-1)
    stride = (w * channels + 3) & ~3
    out = np.zeros((h, w, 4), np.int64)
    out[..., 3] = 255
    for y in range(h):
        for x in range(w):
            for k in range(channels):
                o = y * stride + x * channels + k
                out[y, x, k] = buf[o] if o < buf.size else 0
    if channels == 1:
        out[..., 1:3] = 0
    return out


def box_levels(level0):
    levels = [level0]
    q = int(np.floor(np.log2(max(level0.shape[:2]))))
    for _ in range(q):
        p = levels[-1]
        ph, pw = p.shape[:2]
        nh, nw = max(1, ph // 2), max(1, pw // 2)
        ys = np.minimum(np.arange(nh)[:, None] * 2 + np.array([0, 1]), ph - 1)
        xs = np.minimum(np.arange(nw)[:, None] * 2 + np.array([0, 1]), pw - 1)
        s = sum(p[ys[:, j][:, None], xs[:, i][None, :]] for j in range(2) for i in range(2))
        levels.append((s + 2) // 4)
    return levels


def check(tmp_path, name, img, ctype, depth, want_rgba, channels, **kw):
    p = write_png(tmp_path / name, img, ctype, depth, **kw)
    w, h, levels, tex = vct_amd.load_png(p)
    H, W = img.shape[:2]
    assert (w, h) == (W, H)
    want0 = as_gl(want_rgba, channels)
    lv = box_levels(want0)
    assert levels == len(lv)
    off = 0
    for L in lv:
        lh, lw = L.shape[:2]
        got = rgba_of(tex[off:off + lw * lh], lw, lh)
        assert np.array_equal(got, L), (name, lw, lh)
        off += lw * lh
    assert off == len(tex)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("w,h", [(16, 16), (13, 7), (1, 5), (33, 2)])
def test_rgb_and_rgba_8_bit(tmp_path, interlace, w, h):
    rng = np.random.default_rng(w * 100 + h)
    img = rng.integers(0, 256, (h, w, 4))
    check(tmp_path, "rgba.png", img, 6, 8, img, 4, interlace=interlace)
    rgb = img[..., :3]
    want = np.concatenate([rgb, np.full((h, w, 1), 255)], -1)
    check(tmp_path, "rgb.png", rgb, 2, 8, want, 3, interlace=interlace)


@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
def test_gray_depths_are_one_red_channel(tmp_path, depth):
    rng = np.random.default_rng(depth)
    h, w = 9, 10
    img = rng.integers(0, 1 << depth, (h, w, 1))
    v = img[..., 0] >> 8 if depth == 16 else img[..., 0] * (255 // ((1 << depth) - 1))
    want = np.zeros((h, w, 4), np.int64)
    want[..., 0] = v
    check(tmp_path, "g.png", img, 0, depth, want, 1)


def test_sixteen_bit_rgba_keeps_the_high_byte(tmp_path):
    rng = np.random.default_rng(16)
    img = rng.integers(0, 65536, (6, 8, 4))
    check(tmp_path, "c16.png", img, 6, 16, img >> 8, 4)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_palette_with_and_without_transparency(tmp_path, depth):
    rng = np.random.default_rng(depth + 40)
    h, w = 7, 12
    n = 1 << depth
    plte = rng.integers(0, 256, (n, 3))
    img = rng.integers(0, n, (h, w, 1))
    want = np.concatenate([plte[img[..., 0]], np.full((h, w, 1), 255)], -1)
    check(tmp_path, "p.png", img, 3, depth, want, 3, plte=plte)
    alpha = rng.integers(0, 256, n - 1)  # the last entry keeps 255
    want_a = want.copy()
    want_a[..., 3] = np.concatenate([alpha, [255]])[img[..., 0]]
    check(tmp_path, "pa.png", img, 3, depth, want_a, 4, plte=plte, trns=bytes(alpha.astype(np.uint8)))


def test_rgb_key_colour_becomes_alpha(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 4, (8, 8, 3)) * 60
    key = (60, 0, 120)
    img[2, 3] = key
    trns = b"".join(struct.pack(">H", k) for k in key)
    want = np.concatenate([img, np.full((8, 8, 1), 255)], -1)
    want[..., 3] = np.where((img == key).all(-1), 0, 255)
    check(tmp_path, "key.png", img, 2, 8, want, 4, trns=trns)


def test_two_channels_get_no_storage(tmp_path):
    img = np.random.default_rng(2).integers(0, 256, (4, 4, 2))
    p = write_png(tmp_path / "ga.png", img, 4, 8)
    w, h, levels, tex = vct_amd.load_png(p)
    assert (w, h, levels, len(tex)) == (4, 4, 0, 0)


def test_bad_files_are_rejected(tmp_path):
    with pytest.raises(vct_amd.VCTError):
        vct_amd.load_png(tmp_path / "missing.png")
    (tmp_path / "x.png").write_bytes(b"\x89PNG\r\n\x1a\n" + b"\0" * 40)
    with pytest.raises(vct_amd.VCTError):
        vct_amd.load_png(tmp_path / "x.png")
    img = np.zeros((4, 4, 3), np.int64)
    p = write_png(tmp_path / "t.png", img, 2, 8)
    data = p.read_bytes()
    i = data.index(b"IDAT")
    p.write_bytes(data[:i + 8])  # cut inside the image data
    with pytest.raises(vct_amd.VCTError):
        vct_amd.load_png(p)
    # a 30000 x 30000 RGBA16 header over 20 bytes of data: deflate cannot
    # expand that far, so it is refused without allocating the 7 GB it names
    body = struct.pack(">IIBBBBB", 30000, 30000, 16, 6, 0, 0, 0)
    p.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", body) + chunk(b"IDAT", zlib.compress(b"\0" * 64))
                  + chunk(b"IEND", b""))
    with pytest.raises(vct_amd.VCTError, match="too short"):
        vct_amd.load_png(p)


def test_obj_material_map_kd_png(tmp_path):
    img = np.random.default_rng(9).integers(0, 256, (8, 8, 4))
    write_png(tmp_path / "m.png", img, 6, 8)
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 1 1 1\nmap_Kd m.PNG\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\n"
                                    "usemtl a\nf 1/1 2/2 3/3\n")
    (tmp_path / "m.PNG").write_bytes((tmp_path / "m.png").read_bytes())
    m = vct_amd.load_obj(tmp_path / "m.obj")
    assert m["texid"].tolist() == [0]
    w, h, levels, tex = m["textures"][0]
    assert (w, h, levels) == (8, 8, 4)
    assert np.array_equal(rgba_of(tex, 8, 8), img)
