// png_loader.cpp -- PNG diffuse maps decoded to the RGBA8 levels of
// vct_set_textures, as GLHelper::createTextureFromImage uploads an image that
// is not a DDS file (src/Graphics/GLHelper.cpp:165-210):
//   * the file as stb_image returns it with its own channel count (PNG, ISO
//     15948 / W3C PNG 2nd edition): zlib-compressed scanlines, each with one
//     of the filters None, Sub, Up, Average, Paeth; Adam7 passes when
//     interlaced; gray and palette depths 1-8 (gray scaled to 8 bits by 255 /
//     (2^d - 1), palette indices expanded through PLTE), 16-bit samples to
//     their high byte; gray = 1 channel, gray + alpha = 2, RGB = 3, RGBA = 4,
//     a palette 3 (4 with a tRNS chunk); a tRNS key colour adds an alpha
//     channel (0 where a pixel equals the key, 255 elsewhere);
//   * uploaded with glTextureSubImage2D at GL's default unpack alignment of 4:
//     a row starts every round_up(width * channels, 4) bytes of that buffer
//     (bytes past its end, which the reference reads beyond its allocation,
//     are taken as 0);
//   * 1 channel is GL_R8, sampled as (r, 0, 0, 1); 3 channels GL_RGB8 (alpha
//     1); 4 channels GL_RGBA8; 2 channels receive no storage, so the texture
//     stays incomplete (no levels here);
//   * glGenerateTextureMipmap fills floor(log2(max(w, h))) + 1 levels; GL
//     leaves the filter open and recommends a box: here each texel of level
//     l + 1 is the mean of the 2 x 2 texels of level l that it covers (the
//     last row or column of an odd size has no partner and is dropped, as the
//     level's size floor(w / 2) drops it; an axis of one texel pairs it with
//     itself), rounded to the nearest byte, halves up.
#include <zlib.h>

#include <algorithm>
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

uint32_t be32(const unsigned char *p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Reverses the filters of `rows` scanlines of `rowbytes` bytes (each preceded
// by its filter byte) in place into out; bpp = bytes per complete pixel (>= 1).
bool unfilter(const unsigned char *in, int rows, size_t rowbytes, int bpp, std::vector<unsigned char> &out) {
    out.assign((size_t)rows * rowbytes, 0);
    for (int y = 0; y < rows; ++y) {
        const unsigned char *src = in + (size_t)y * (rowbytes + 1);
        const int f = src[0];
        ++src;
        unsigned char *cur = out.data() + (size_t)y * rowbytes;
        const unsigned char *prev = y > 0 ? cur - rowbytes : nullptr;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int v;
            switch (f) {
            case 0: v = src[i]; break;
            case 1: v = src[i] + a; break;
            case 2: v = src[i] + b; break;
            case 3: v = src[i] + ((a + b) >> 1); break;
            case 4: v = src[i] + paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = (unsigned char)(v & 0xff);
        }
    }
    return true;
}

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<unsigned char> plte, trns, idat;
};

int samples_of(int ctype) { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

// The 8-bit samples (the file's own channels, plus alpha from a tRNS key) of a
// reduced image of pw x ph pixels from its unfiltered rows.
void expand(const Png &p, const unsigned char *rows, size_t rowbytes, uint32_t pw, uint32_t ph, int outc,
            std::vector<unsigned char> &dst, uint32_t x0, uint32_t dx, uint32_t y0, uint32_t dy) {
    const int ns = samples_of(p.ctype);
    for (uint32_t y = 0; y < ph; ++y) {
        const unsigned char *r = rows + (size_t)y * rowbytes;
        for (uint32_t x = 0; x < pw; ++x) {
            uint32_t s[4] = {0, 0, 0, 0}, raw[4] = {0, 0, 0, 0};
            for (int k = 0; k < ns; ++k) {
                if (p.depth == 16) {
                    const size_t o = ((size_t)x * ns + k) * 2;
                    raw[k] = ((uint32_t)r[o] << 8) | r[o + 1];
                    s[k] = r[o];  // the high byte
                } else if (p.depth == 8) {
                    raw[k] = s[k] = r[(size_t)x * ns + k];
                } else {  // 1, 2 or 4 bits: one sample (gray or palette index)
                    const size_t bit = (size_t)x * p.depth;
                    const uint32_t v = (r[bit >> 3] >> (8 - p.depth - (bit & 7))) & ((1u << p.depth) - 1);
                    raw[k] = v;
                    s[k] = p.ctype == 3 ? v : v * (255u / ((1u << p.depth) - 1));
                }
            }
            unsigned char *o = &dst[((size_t)(y0 + y * dy) * p.w + (x0 + x * dx)) * outc];
            if (p.ctype == 3) {
                const uint32_t i = s[0];
                for (int k = 0; k < 3; ++k) o[k] = (size_t)i * 3 + k < p.plte.size() ? p.plte[i * 3 + k] : 0;
                if (outc == 4) o[3] = i < p.trns.size() ? p.trns[i] : 255;
                continue;
            }
            for (int k = 0; k < ns; ++k) o[k] = (unsigned char)s[k];
            if (outc > ns) {  // a tRNS key colour: alpha 0 where the pixel equals it
                bool key = true;
                for (int k = 0; k < ns; ++k) {
                    const uint32_t kv = p.trns.size() >= (size_t)(2 * k + 2)
                                            ? ((uint32_t)p.trns[2 * k] << 8) | p.trns[2 * k + 1]
                                            : 0xffffffffu;
                    key = key && raw[k] == kv;
                }
                o[ns] = key ? 0 : 255;
            }
        }
    }
}

}  // namespace

extern "C" {

int vct_png_load(const char *path, vct_image **out, char *err, int errlen) {
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
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (data.size() < 8 || std::memcmp(data.data(), sig, 8) != 0) {
        set_err(err, errlen, std::string("not a PNG file: ") + path);
        return VCT_ERR_INVALID;
    }
    Png p;
    bool ihdr = false, iend = false;
    for (size_t off = 8; off + 12 <= data.size() && !iend;) {
        const uint32_t len = be32(&data[off]);
        if (off + 12 + (size_t)len > data.size()) break;
        const unsigned char *type = &data[off + 4], *body = &data[off + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            p.w = be32(body);
            p.h = be32(body + 4);
            p.depth = body[8];
            p.ctype = body[9];
            p.interlace = body[12];
            ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            p.plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            p.trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            p.idat.insert(p.idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            iend = true;
        }
        off += 12 + (size_t)len;
    }
    const bool depth_ok = p.depth == 8 || p.depth == 16 ||
                          ((p.ctype == 0 || p.ctype == 3) && (p.depth == 1 || p.depth == 2 || p.depth == 4));
    const bool ctype_ok = p.ctype == 0 || p.ctype == 2 || p.ctype == 3 || p.ctype == 4 || p.ctype == 6;
    if (!ihdr || !ctype_ok || !depth_ok || (p.ctype == 3 && (p.depth == 16 || p.plte.empty())) || p.interlace > 1 ||
        p.w == 0 || p.h == 0 || p.w > 32768 || p.h > 32768 || p.idat.empty()) {
        set_err(err, errlen, std::string("unsupported or malformed PNG: ") + path);
        return VCT_ERR_INVALID;
    }
    const int ns = samples_of(p.ctype);
    int outc = p.ctype == 3 ? (p.trns.empty() ? 3 : 4) : ns;
    if (!p.trns.empty() && (p.ctype == 0 || p.ctype == 2)) outc = ns + 1;
    // the inflated scanlines: every pass's rows with their filter bytes
    const int bits_pp = ns * p.depth;
    auto rowbytes_of = [&](uint32_t w) { return ((size_t)w * bits_pp + 7) / 8; };
    static const uint32_t ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const uint32_t adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    size_t total = 0;
    const int npass = p.interlace ? 7 : 1;
    for (int k = 0; k < npass; ++k) {
        const uint32_t x0 = p.interlace ? ax0[k] : 0, dx = p.interlace ? adx[k] : 1;
        const uint32_t y0 = p.interlace ? ay0[k] : 0, dy = p.interlace ? ady[k] : 1;
        const uint32_t pw = p.w > x0 ? (p.w - x0 + dx - 1) / dx : 0, ph = p.h > y0 ? (p.h - y0 + dy - 1) / dy : 0;
        if (pw && ph) total += (size_t)ph * (rowbytes_of(pw) + 1);
    }
    // deflate expands at most 1032:1 (a 258-byte match per two-bit code): a
    // header announcing more than the data can hold is rejected before any
    // buffer of its size is allocated
    if (total / 1032 > p.idat.size() + 64) {
        set_err(err, errlen, std::string("PNG image data too short for its size: ") + path);
        return VCT_ERR_INVALID;
    }
    std::vector<unsigned char> raw(total);
    uLongf got = (uLongf)total;
    if (uncompress(raw.data(), &got, p.idat.data(), (uLong)p.idat.size()) != Z_OK || got != total) {
        set_err(err, errlen, std::string("bad PNG image data: ") + path);
        return VCT_ERR_INVALID;
    }
    std::vector<unsigned char> pix((size_t)p.w * p.h * outc, 0), rows;
    size_t at = 0;
    const int bpp = std::max(1, bits_pp / 8);
    for (int k = 0; k < npass; ++k) {
        const uint32_t x0 = p.interlace ? ax0[k] : 0, dx = p.interlace ? adx[k] : 1;
        const uint32_t y0 = p.interlace ? ay0[k] : 0, dy = p.interlace ? ady[k] : 1;
        const uint32_t pw = p.w > x0 ? (p.w - x0 + dx - 1) / dx : 0, ph = p.h > y0 ? (p.h - y0 + dy - 1) / dy : 0;
        if (!pw || !ph) continue;
        const size_t rb = rowbytes_of(pw);
        if (!unfilter(raw.data() + at, (int)ph, rb, bpp, rows)) {
            set_err(err, errlen, std::string("bad PNG filter type: ") + path);
            return VCT_ERR_INVALID;
        }
        expand(p, rows.data(), rb, pw, ph, outc, pix, x0, dx, y0, dy);
        at += (size_t)ph * (rb + 1);
    }
    vct_image *img = new vct_image();
    img->width = (int)p.w;
    img->height = (int)p.h;
    if (outc == 2) {  // no storage for two channels: an incomplete texture
        img->levels = 0;
        *out = img;
        return VCT_OK;
    }
    // level 0 as glTextureSubImage2D reads the tightly packed buffer with rows
    // every round_up(w * outc, 4) bytes
    const size_t stride = ((size_t)p.w * outc + 3) & ~(size_t)3;
    const size_t W = p.w, H = p.h;
    std::vector<uint32_t> l0(W * H);
    for (size_t y = 0; y < H; ++y)
        for (size_t x = 0; x < W; ++x) {
            uint32_t c[4] = {0, 0, 0, 255};
            for (int k = 0; k < outc; ++k) {
                const size_t o = y * stride + x * outc + k;
                c[k] = o < pix.size() ? pix[o] : 0u;
            }
            if (outc == 1) c[1] = c[2] = 0;  // GL_R8: (r, 0, 0, 1)
            if (outc < 4) c[3] = 255;
            l0[y * W + x] = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
        }
    int q = 0;
    while ((std::max(p.w, p.h) >> (q + 1)) > 0) ++q;
    img->texels = l0;
    std::vector<uint32_t> prev = l0;
    size_t pw = W, ph = H;
    for (int l = 1; l <= q; ++l) {
        const size_t nw = std::max<size_t>(1, W >> l), nh = std::max<size_t>(1, H >> l);
        std::vector<uint32_t> cur(nw * nh);
        for (size_t y = 0; y < nh; ++y)
            for (size_t x = 0; x < nw; ++x) {
                const size_t xs[2] = {std::min(2 * x, pw - 1), std::min(2 * x + 1, pw - 1)};
                const size_t ys[2] = {std::min(2 * y, ph - 1), std::min(2 * y + 1, ph - 1)};
                uint32_t v = 0;
                for (int k = 0; k < 4; ++k) {
                    uint32_t s = 0;
                    for (int j = 0; j < 2; ++j)
                        for (int i = 0; i < 2; ++i) s += (prev[ys[j] * pw + xs[i]] >> (8 * k)) & 0xffu;
                    v |= ((s + 2) / 4) << (8 * k);
                }
                cur[y * nw + x] = v;
            }
        img->texels.insert(img->texels.end(), cur.begin(), cur.end());
        prev.swap(cur);
        pw = nw;
        ph = nh;
    }
    img->levels = q + 1;
    *out = img;
    return VCT_OK;
}

}  // extern "C"
