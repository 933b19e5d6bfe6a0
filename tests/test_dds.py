"""DDS diffuse maps (vct_dds_load, ResourceLoader::loadDDS at
src/ResourceLoader.h:26-108) decoded by EXT_texture_compression_s3tc.

The reference ships no DDS fixture (its Sponza assets are missing from the
mount), so the files are written here, block by block, and every expected
texel is the extension's closed form evaluated by hand: colours
RGB565 -> byte round(255 c / 31) (or / 63), the interpolated colours
(2 c0 + c1) / 3, (c0 + 2 c1) / 3 or (c0 + c1) / 2, DXT3's 4-bit alpha x 17 and
DXT5's 8-entry alpha tables.  Texel (x, y) of a block takes code bits
2 (4 y + x); levels follow the reference's loop (w >>= 1, h >>= 1 while
level < dwMipMapCount and a dimension is non-zero).
"""
import numpy as np
import pytest

import vct_amd
from dds_fixtures import colour_block, dds_bytes

pytestmark = pytest.mark.filterwarnings("ignore")


def load(tmp_path, data, name="t.dds"):
    p = tmp_path / name
    p.write_bytes(data)
    return vct_amd.load_dds(p)


def rgba(texels):
    t = np.asarray(texels, np.uint32)
    return np.stack([(t >> (8 * k)) & 0xFF for k in range(4)], -1).astype(int)


def test_dxt1_four_colour_block(tmp_path):
    # red (31, 0, 0) and blue (0, 0, 31): color0 0xF800 > color1 0x001F
    codes = [i % 4 for i in range(16)]
    w, h, lv, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT1", [colour_block(0xF800, 0x001F, codes)]))
    assert (w, h, lv) == (4, 4, 1)
    want = {0: (255, 0, 0, 255), 1: (0, 0, 255, 255),
            2: (170, 0, 85, 255),   # (2 * 31 + 0) / 93 * 255 = 170, 31 / 93 * 255 = 85
            3: (85, 0, 170, 255)}
    got = rgba(tex)
    for i in range(16):
        assert tuple(got[i]) == want[codes[i]], i
    # the green channel on 6 bits: 63 and 0
    _, _, _, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT1", [colour_block(0x07E0, 0x0000, codes)]))
    got = rgba(tex)
    assert [tuple(got[i][:3]) for i in range(4)] == [(0, 255, 0), (0, 0, 0), (0, 170, 0), (0, 85, 0)]


def test_dxt1_three_colour_block_and_opaque_black(tmp_path):
    # color0 0x001F <= color1 0xF800: RGB2 = (RGB0 + RGB1) / 2, RGB3 = black,
    # opaque (COMPRESSED_RGB_S3TC_DXT1); 255 * 31 / 62 = 127.5 rounds up
    codes = [i % 4 for i in range(16)]
    _, _, _, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT1", [colour_block(0x001F, 0xF800, codes)]))
    got = rgba(tex)
    assert [tuple(got[i]) for i in range(4)] == [(0, 0, 255, 255), (255, 0, 0, 255), (128, 0, 128, 255),
                                                (0, 0, 0, 255)]


def test_dxt3_alpha_and_always_four_colour(tmp_path):
    alpha = b"".join(bytes([(2 * i) | ((2 * i + 1) << 4)]) for i in range(8))  # texel i has alpha i
    codes = [i % 4 for i in range(16)]
    block = alpha + colour_block(0x001F, 0xF800, codes)  # color0 < color1, still four colours
    _, _, _, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT3", [block]))
    got = rgba(tex)
    assert [int(got[i][3]) for i in range(16)] == [17 * i for i in range(16)]
    assert [tuple(got[i][:3]) for i in range(4)] == [(0, 0, 255), (255, 0, 0), (85, 0, 170), (170, 0, 85)]


@pytest.mark.parametrize("a0,a1,table", [
    (255, 0, [255, 0, 219, 182, 146, 109, 73, 36]),   # ((8 - c) a0 + (c - 1) a1) / 7
    (0, 255, [0, 255, 51, 102, 153, 204, 0, 255]),     # ((6 - c) a0 + (c - 1) a1) / 5, then 0, 255
])
def test_dxt5_alpha_tables(tmp_path, a0, a1, table):
    codes = [i % 8 for i in range(16)]
    bits = 0
    for i, c in enumerate(codes):
        bits |= c << (3 * i)
    block = bytes([a0, a1]) + bits.to_bytes(6, "little") + colour_block(0xFFFF, 0x0000, [0] * 16)
    _, _, _, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT5", [block]))
    got = rgba(tex)
    assert [int(got[i][3]) for i in range(16)] == [table[c] for c in codes]
    assert all(tuple(got[i][:3]) == (255, 255, 255) for i in range(16))


def test_partial_blocks_and_mip_chain(tmp_path):
    # 6 x 5, three levels (6x5, 3x2, 1x1): 2x2 blocks, then one, then one.
    # Each block is one colour, code 0, colour0 = the block's index (blue bits)
    def solid(k):
        return colour_block(0xF800 | k, 0x0000, [0] * 16)
    lv0 = [solid(k) for k in range(4)]
    data = dds_bytes(6, 5, 3, b"DXT1", lv0 + [solid(8), solid(9)])
    w, h, levels, tex = load(tmp_path, data)
    assert (w, h, levels) == (6, 5, 3)
    assert tex.size == 6 * 5 + 3 * 2 + 1
    l0 = rgba(tex[:30]).reshape(5, 6, 4)
    for y in range(5):
        for x in range(6):
            k = (y // 4) * 2 + (x // 4)
            assert l0[y, x, 2] == round(255 * k / 31), (x, y)
    assert np.all(rgba(tex[30:36])[:, 2] == round(255 * 8 / 31))
    assert rgba(tex[36:])[0, 2] == round(255 * 9 / 31)


def test_texel_order_inside_a_block(tmp_path):
    # code 1 at exactly one texel (x, y): bits 2 (4 y + x)
    for x, y in [(0, 0), (3, 0), (0, 3), (2, 1)]:
        codes = [0] * 16
        codes[4 * y + x] = 1
        _, _, _, tex = load(tmp_path, dds_bytes(4, 4, 1, b"DXT1", [colour_block(0xF800, 0x001F, codes)]))
        img = rgba(tex).reshape(4, 4, 4)
        hit = np.argwhere(img[:, :, 2] == 255)
        assert hit.tolist() == [[y, x]]


def test_non_square_chain_stops_at_an_empty_level(tmp_path):
    # 4 x 2 with three levels: 4x2, 2x1, then 1x0 -- the reference uploads an
    # empty image there, leaving the texture incomplete; the list stops at 2
    blk = colour_block(0xF800, 0x001F, [0] * 16)
    w, h, levels, tex = load(tmp_path, dds_bytes(4, 2, 3, b"DXT1", [blk, blk, blk]))
    assert (w, h, levels) == (4, 2, 2)
    assert tex.size == 4 * 2 + 2 * 1


@pytest.mark.parametrize("fourcc", [b"DXT2", b"DXT4", b"ATI2", b"\0\0\0\0"])
def test_other_formats_are_rejected(tmp_path, fourcc):
    with pytest.raises(vct_amd.VCTError):
        load(tmp_path, dds_bytes(4, 4, 1, fourcc, [b"\0" * 16]))


def test_truncated_and_foreign_files_are_rejected(tmp_path):
    blk = colour_block(0xF800, 0x001F, [0] * 16)
    with pytest.raises(vct_amd.VCTError):
        load(tmp_path, dds_bytes(8, 8, 1, b"DXT1", [blk]))  # 4 blocks announced, 1 present
    with pytest.raises(vct_amd.VCTError):
        load(tmp_path, b"PNG " + b"\0" * 200)
    with pytest.raises(vct_amd.VCTError):
        vct_amd.load_dds(tmp_path / "missing.dds")


def test_obj_material_map_kd_loads_its_dds(tmp_path):
    blk = colour_block(0x07E0, 0x0000, [0] * 16)  # green
    (tmp_path / "tex").mkdir()
    (tmp_path / "tex" / "g.dds").write_bytes(dds_bytes(4, 4, 3, b"DXT1", [blk, blk, blk]))
    (tmp_path / "m.mtl").write_text("newmtl green\nKd 0.5 0.5 0.5\nmap_Kd tex\\g.dds\n"
                                    "newmtl plain\nKd 1 0 0\nmap_Kd photo.png\n")
    (tmp_path / "s.obj").write_text(
        "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvt 1 1\n"
        "usemtl green\nf 1/1 2/2 3/3\nusemtl plain\nf 2/2 4/4 3/3\nusemtl green\nf 1 2 4\n")
    m = vct_amd.load_obj(tmp_path / "s.obj")
    assert len(m["textures"]) == 1
    w, h, levels, tex = m["textures"][0]
    assert (w, h, levels) == (4, 4, 3)
    assert np.all(rgba(tex)[:, 1] == 255)
    assert m["texid"].tolist() == [0, -1, 0]
    # texture v flipped (Mesh.cpp:162); a face without vt has uv (0, 0)
    assert np.allclose(m["uv"][0], [[0, 1], [1, 1], [0, 0]])
    assert np.allclose(m["uv"][2], 0.0)
    # a map in another image format is not decoded: the face keeps Kd
    assert m["albedo"][1] == 0xFF0000FF
