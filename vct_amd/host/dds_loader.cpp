// dds_loader.cpp -- DDS (DXT1 / DXT3 / DXT5) diffuse maps decoded to RGBA8
// levels, for vct_set_textures.  Follows ResourceLoader::loadDDS
// (src/ResourceLoader.h:26-108): the "DDS " code, the 124-byte header, the
// four-character code in ddspf.dwFourCC, then the levels back to back, level l
// of (width >> l) x (height >> l) texels in ((w + 3) / 4) * ((h + 3) / 4)
// blocks of 8 (DXT1) or 16 bytes, for l < dwMipMapCount while a dimension is
// non-zero.  Where the reference hands the blocks to glCompressedTexImage2D,
// the blocks are decoded here by EXT_texture_compression_s3tc:
//   colour block: color0, color1 (RGB 5:6:5, little endian), 32 bits of 2-bit
//     codes, texel (x, y) at bits 2 (4 y + x); RGB0/RGB1 the two colours as
//     unsigned normalized values, RGB2 = (2 RGB0 + RGB1) / 3 and RGB3 =
//     (RGB0 + 2 RGB1) / 3 when color0 > color1 (always, in DXT3 and DXT5),
//     else RGB2 = (RGB0 + RGB1) / 2 and RGB3 = black (DXT1, opaque: the
//     reference's format is COMPRESSED_RGB_S3TC_DXT1);
//   DXT3: 64 bits of 4-bit alpha before the colour block, texel i at bits 4 i;
//   DXT5: alpha0, alpha1, 48 bits of 3-bit codes (texel i at bits 3 i): codes
//     0 and 1 the two alphas, 2 .. 7 ((8 - c) alpha0 + (c - 1) alpha1) / 7
//     when alpha0 > alpha1, else 2 .. 5 ((6 - c) alpha0 + (c - 1) alpha1) / 5,
//     6 -> 0 and 7 -> 255.
// Each channel is the exact value rounded to the nearest byte, the byte the
// texture stores (the extension leaves the decoder's precision open).  Only
// the three-colour midpoint can fall halfway -- (RGB0 + RGB1) / 2 with an odd
// 5-bit sum 31 or 6-bit sum 63 -- and rounds up; the other denominators (31,
// 63, 93, 189, 15, 7, 5 in units of the byte) never make an odd multiple of one
// half.  Texels of a level row-major from the first block row (GL's t = 0 row).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "image.h"
#include "vct_app.h"

namespace {

void set_err(char *err, int len, const std::string &msg) {
    if (err && len > 0) std::snprintf(err, (size_t)len, "%s", msg.c_str());
}

uint32_t rd32(const unsigned char *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t rd16(const unsigned char *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// round(255 * num / den), halves up, for 0 <= num <= den, den > 0
uint32_t byte_of(uint32_t num, uint32_t den) { return (510u * num + den) / (2u * den); }

// The four colours of a colour block as RGB bytes; `four`: four-colour mode.
void block_colours(const unsigned char *b, bool dxt1, uint32_t rgb[4][3]) {
    const uint16_t c0 = rd16(b), c1 = rd16(b + 2);
    const uint32_t r0 = c0 >> 11, g0 = (c0 >> 5) & 63u, b0 = c0 & 31u;
    const uint32_t r1 = c1 >> 11, g1 = (c1 >> 5) & 63u, b1 = c1 & 31u;
    rgb[0][0] = byte_of(r0, 31);
    rgb[0][1] = byte_of(g0, 63);
    rgb[0][2] = byte_of(b0, 31);
    rgb[1][0] = byte_of(r1, 31);
    rgb[1][1] = byte_of(g1, 63);
    rgb[1][2] = byte_of(b1, 31);
    if (!dxt1 || c0 > c1) {
        rgb[2][0] = byte_of(2 * r0 + r1, 93);
        rgb[2][1] = byte_of(2 * g0 + g1, 189);
        rgb[2][2] = byte_of(2 * b0 + b1, 93);
        rgb[3][0] = byte_of(r0 + 2 * r1, 93);
        rgb[3][1] = byte_of(g0 + 2 * g1, 189);
        rgb[3][2] = byte_of(b0 + 2 * b1, 93);
    } else {
        rgb[2][0] = byte_of(r0 + r1, 62);
        rgb[2][1] = byte_of(g0 + g1, 126);
        rgb[2][2] = byte_of(b0 + b1, 62);
        rgb[3][0] = rgb[3][1] = rgb[3][2] = 0u;
    }
}

// Decodes one 4x4 block into the 16 texels t[y * 4 + x].
void decode_block(const unsigned char *b, int fmt, uint32_t t[16]) {
    uint32_t alpha[16];
    const unsigned char *cb = b;
    if (fmt == 1) {
        for (int i = 0; i < 16; ++i) alpha[i] = 255u;
    } else if (fmt == 3) {
        for (int i = 0; i < 16; ++i) alpha[i] = byte_of((b[i >> 1] >> (4 * (i & 1))) & 15u, 15);
        cb = b + 8;
    } else {
        const uint32_t a0 = b[0], a1 = b[1];
        uint64_t codes = 0;
        for (int k = 0; k < 6; ++k) codes |= (uint64_t)b[2 + k] << (8 * k);
        for (int i = 0; i < 16; ++i) {
            const uint32_t c = (uint32_t)(codes >> (3 * i)) & 7u;
            uint32_t a;
            if (c == 0) a = a0;
            else if (c == 1) a = a1;
            else if (a0 > a1) a = byte_of((8 - c) * a0 + (c - 1) * a1, 7 * 255) ;
            else if (c <= 5) a = byte_of((6 - c) * a0 + (c - 1) * a1, 5 * 255);
            else a = c == 6 ? 0u : 255u;
            alpha[i] = a;
        }
        cb = b + 8;
    }
    uint32_t rgb[4][3];
    block_colours(cb, fmt == 1, rgb);
    const uint32_t codes = rd32(cb + 4);
    for (int i = 0; i < 16; ++i) {
        const uint32_t c = (codes >> (2 * i)) & 3u;
        t[i] = rgb[c][0] | (rgb[c][1] << 8) | (rgb[c][2] << 16) | (alpha[i] << 24);
    }
}

}  // namespace

extern "C" {

int vct_dds_load(const char *path, vct_image **out, char *err, int errlen) {
    if (!path || !out) return VCT_ERR_INVALID;
    *out = nullptr;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        set_err(err, errlen, std::string("cannot open ") + path);
        return VCT_ERR_INVALID;
    }
    std::vector<unsigned char> data;
    unsigned char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) data.insert(data.end(), buf, buf + n);
    std::fclose(fp);
    if (data.size() < 128 || std::memcmp(data.data(), "DDS ", 4) != 0) {
        set_err(err, errlen, std::string("not a DDS file: ") + path);
        return VCT_ERR_INVALID;
    }
    const unsigned char *h = data.data() + 4;  // dds_header: dwSize at 0
    const uint32_t height = rd32(h + 8), width = rd32(h + 12), mips = rd32(h + 24);
    const unsigned char *fourcc = h + 72 + 8;  // ddspf at 72: dwSize, dwFlags, dwFourCC
    int fmt;
    if (std::memcmp(fourcc, "DXT1", 4) == 0) fmt = 1;
    else if (std::memcmp(fourcc, "DXT3", 4) == 0) fmt = 3;
    else if (std::memcmp(fourcc, "DXT5", 4) == 0) fmt = 5;
    else {
        set_err(err, errlen, std::string("unsupported DDS format (DXT1, DXT3, DXT5 only): ") + path);
        return VCT_ERR_INVALID;
    }
    if (width == 0 || height == 0 || width > 32768 || height > 32768) {
        set_err(err, errlen, std::string("bad DDS size: ") + path);
        return VCT_ERR_INVALID;
    }
    const size_t bsize = fmt == 1 ? 8 : 16;
    vct_image *img = new vct_image();
    img->width = (int)width;
    img->height = (int)height;
    size_t off = 128;
    uint32_t w = width, hgt = height;
    for (uint32_t level = 0; level < mips && (w || hgt); ++level, w /= 2, hgt /= 2) {
        if (w == 0 || hgt == 0) break;  // an empty image: the chain ends incomplete
        const size_t bw = (w + 3) / 4, bh = (hgt + 3) / 4;
        if (off + bw * bh * bsize > data.size()) {
            set_err(err, errlen, std::string("truncated DDS level ") + std::to_string(level) + ": " + path);
            delete img;
            return VCT_ERR_INVALID;
        }
        const size_t base = img->texels.size();
        img->texels.resize(base + (size_t)w * hgt);
        uint32_t *dst = img->texels.data() + base;
        for (size_t by = 0; by < bh; ++by)
            for (size_t bx = 0; bx < bw; ++bx) {
                uint32_t t[16];
                decode_block(data.data() + off + (by * bw + bx) * bsize, fmt, t);
                for (int y = 0; y < 4; ++y)
                    for (int x = 0; x < 4; ++x) {
                        const size_t px = bx * 4 + x, py = by * 4 + y;
                        if (px < w && py < hgt) dst[py * w + px] = t[y * 4 + x];
                    }
            }
        off += bw * bh * bsize;
        img->levels = (int)level + 1;
    }
    *out = img;
    return VCT_OK;
}

int vct_dds_texture(const vct_image *img, vct_texture *t) {
    if (!img || !t) return VCT_ERR_INVALID;
    t->width = img->width;
    t->height = img->height;
    t->levels = img->levels;
    t->texels = img->texels.data();
    return VCT_OK;
}

void vct_dds_free(vct_image *img) { delete img; }

}  // extern "C"
