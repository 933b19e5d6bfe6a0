"""Writes DDS files (the layout ResourceLoader::loadDDS reads,
src/ResourceLoader.h:26-108) block by block for the texture tests."""
import struct

import numpy as np


def dds_bytes(w, h, mips, fourcc, levels_blocks):
    """"DDS " + the 124-byte header (dwHeight, dwWidth, dwMipMapCount, ddspf
    with the four-character code) + the blocks of every level."""
    hdr = struct.pack("<7I", 124, 0x1 | 0x2 | 0x4 | 0x1000 | 0x20000, h, w, 0, 0, mips)
    hdr += b"\0" * 44
    hdr += struct.pack("<2I", 32, 0x4) + fourcc + b"\0" * 20
    hdr += struct.pack("<5I", 0x1000, 0, 0, 0, 0)
    assert len(hdr) == 124
    return b"DDS " + hdr + b"".join(levels_blocks)


def colour_block(c0, c1, codes):
    """An S3TC colour block: two RGB565 colours and 16 two-bit codes."""
    bits = 0
    for i, c in enumerate(codes):
        bits |= (c & 3) << (2 * i)
    return struct.pack("<HHI", c0, c1, bits)


def random_dds(rng, w, h, fourcc=b"DXT5"):
    """A full mip chain of random blocks (w, h powers of two, square)."""
    size = 8 if fourcc == b"DXT1" else 16
    levels = []
    lw, lh, n = w, h, 0
    while lw or lh:
        nb = ((max(lw, 1) + 3) // 4) * ((max(lh, 1) + 3) // 4)
        levels.append(rng.integers(0, 256, nb * size, dtype=np.uint16).astype(np.uint8).tobytes())
        lw, lh, n = lw // 2, lh // 2, n + 1
    return dds_bytes(w, h, n, fourcc, levels)
