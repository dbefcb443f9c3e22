// PNG decoding for the mask level set (host side; zlib inflates the IDAT
// stream, everything else is this file). The output is what the reference's
// reader (Src/Utils/PngProcess.cpp:4-72, libpng 1.6) hands to its level-set
// code after its transformations: 8-bit RGBA rows, top row first, with 16-bit
// samples cut to their high byte (png_set_strip_16), palette entries expanded
// to RGB, 1/2/4-bit gray scaled to 8 bits, gray replicated into R, G, B and a
// 0xFF filler where the image has no alpha (tRNS transparency becomes alpha).
// Interlaced (Adam7) images are refused.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ibtfo.h"
#include "ibtfo_host.h"

namespace ibtfo {

static unsigned be32(const unsigned char* p) {
    return (unsigned)p[0] << 24 | (unsigned)p[1] << 16 | (unsigned)p[2] << 8 | (unsigned)p[3];
}

static int paeth(int a, int b, int c) {
    int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

int png_decode(const unsigned char* buf, size_t len, std::vector<unsigned char>& rgba, int* width, int* height) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (len < 8 || memcmp(buf, sig, 8) != 0) { set_error("not a PNG file"); return IBTFO_E_IO; }
    size_t pos = 8;
    unsigned w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte, trns;
    bool end = false;
    while (!end) {
        if (pos + 12 > len) { set_error("PNG: truncated chunk"); return IBTFO_E_IO; }
        unsigned n = be32(buf + pos);
        const unsigned char* type = buf + pos + 4;
        if (n > len || pos + 12 + n > len) { set_error("PNG: truncated chunk"); return IBTFO_E_IO; }
        const unsigned char* d = buf + pos + 8;
        unsigned crc = (unsigned)crc32(0L, type, 4 + n);
        if (crc != be32(d + n)) {
            char name[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < 4; q++) name[q] = (type[q] >= 32 && type[q] < 127) ? (char)type[q] : '?';
            set_error("PNG: CRC mismatch in chunk %s at byte %zu", name, pos);
            return IBTFO_E_IO;
        }
        if (!memcmp(type, "IHDR", 4)) {
            if (n != 13) { set_error("PNG: bad IHDR"); return IBTFO_E_IO; }
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) { set_error("PNG: unknown compression or filter method"); return IBTFO_E_IO; }
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + n);
        } else if (!memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + n);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + n);
        } else if (!memcmp(type, "IEND", 4)) {
            end = true;
        }
        pos += 12 + n;
    }
    if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) { set_error("PNG: bad dimensions %ux%u", w, h); return IBTFO_E_IO; }
    if (interlace != 0) { set_error("PNG: interlaced images are not supported"); return IBTFO_E_IO; }
    int chans;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default: set_error("PNG: bad colour type %d", ctype); return IBTFO_E_IO;
    }
    bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                    (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                    ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) { set_error("PNG: bit depth %d invalid for colour type %d", depth, ctype); return IBTFO_E_IO; }
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) { set_error("PNG: palette image without PLTE"); return IBTFO_E_IO; }
    const size_t bits = (size_t)chans * depth;
    const size_t stride = (w * bits + 7) / 8;  // bytes per row without the filter byte
    const size_t bpp = (bits + 7) / 8;         // filter unit
    std::vector<unsigned char> raw((stride + 1) * h);
    uLongf rawlen = raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), idat.size()) != Z_OK || rawlen != raw.size()) {
        set_error("PNG: bad or short image data");
        return IBTFO_E_IO;
    }
    // undo the per-row filters in place
    std::vector<unsigned char> img(stride * h);
    for (unsigned y = 0; y < h; y++) {
        const unsigned char* in = &raw[y * (stride + 1)] + 1;
        unsigned char* out = &img[y * stride];
        const unsigned char* up = y ? &img[(y - 1) * stride] : nullptr;
        const int f = in[-1];
        // a = the byte one pixel to the left, b = the byte above, c = above-left
        // (0 outside the image); all sums modulo 256
        switch (f) {
            case 0:  // None
                memcpy(out, in, stride);
                break;
            case 1:  // Sub
                for (size_t i = 0; i < stride; i++) out[i] = (unsigned char)(in[i] + (i >= bpp ? out[i - bpp] : 0));
                break;
            case 2:  // Up
                if (up)
                    for (size_t i = 0; i < stride; i++) out[i] = (unsigned char)(in[i] + up[i]);
                else
                    memcpy(out, in, stride);
                break;
            case 3:  // Average
                for (size_t i = 0; i < stride; i++) {
                    const int a = i >= bpp ? out[i - bpp] : 0, b = up ? up[i] : 0;
                    out[i] = (unsigned char)(in[i] + ((a + b) >> 1));
                }
                break;
            case 4:  // Paeth
                for (size_t i = 0; i < stride; i++) {
                    const int a = i >= bpp ? out[i - bpp] : 0, b = up ? up[i] : 0;
                    const int c = (up && i >= bpp) ? up[i - bpp] : 0;
                    out[i] = (unsigned char)(in[i] + paeth(a, b, c));
                }
                break;
            default: set_error("PNG: bad filter type %d in row %u", f, y); return IBTFO_E_IO;
        }
    }
    // to 8-bit RGBA
    if (ctype == 6 && depth == 8) {  // already 8-bit RGBA
        rgba.assign(img.begin(), img.end());
        *width = (int)w;
        *height = (int)h;
        return IBTFO_OK;
    }
    rgba.assign((size_t)w * h * 4, 0);
    auto sample = [&](const unsigned char* row, size_t idx) -> unsigned {  // idx-th sample of a row
        if (depth == 8) return row[idx];
        if (depth == 16) return row[2 * idx];  // high byte
        size_t bit = idx * depth;
        unsigned v = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
        return v;
    };
    auto scale_gray = [&](unsigned v) -> unsigned { return depth < 8 ? v * 255u / ((1u << depth) - 1) : v; };
    for (unsigned y = 0; y < h; y++) {
        const unsigned char* row = &img[y * stride];
        for (unsigned x = 0; x < w; x++) {
            unsigned char* o = &rgba[((size_t)y * w + x) * 4];
            unsigned r, g, b, a = 255;
            if (ctype == 0) {
                unsigned v = sample(row, x);
                if (trns.size() >= 2) {
                    unsigned t = ((unsigned)trns[0] << 8 | trns[1]);
                    unsigned full = depth == 16 ? ((unsigned)row[2 * x] << 8 | row[2 * x + 1]) : v;
                    if (full == t) a = 0;
                }
                r = g = b = scale_gray(v);
            } else if (ctype == 3) {
                unsigned k = sample(row, x);
                if (3 * k + 2 >= plte.size()) { set_error("PNG: palette index out of range"); return IBTFO_E_IO; }
                r = plte[3 * k]; g = plte[3 * k + 1]; b = plte[3 * k + 2];
                if (k < trns.size()) a = trns[k];
            } else if (ctype == 2) {
                r = sample(row, 3 * x); g = sample(row, 3 * x + 1); b = sample(row, 3 * x + 2);
                if (trns.size() >= 6) {
                    // the transparent colour is matched on the full samples, before
                    // 16-bit samples are cut to 8 (as libpng expands tRNS first)
                    bool eq = true;
                    for (int q = 0; q < 3; q++) {
                        unsigned t = (unsigned)trns[2 * q] << 8 | trns[2 * q + 1];
                        unsigned full = depth == 16 ? ((unsigned)row[6 * x + 2 * q] << 8 | row[6 * x + 2 * q + 1])
                                                    : row[3 * x + q];
                        eq = eq && full == t;
                    }
                    if (eq) a = 0;
                }
            } else if (ctype == 4) {
                r = g = b = sample(row, 2 * x);
                a = sample(row, 2 * x + 1);
            } else {
                r = sample(row, 4 * x); g = sample(row, 4 * x + 1); b = sample(row, 4 * x + 2);
                a = sample(row, 4 * x + 3);
            }
            o[0] = (unsigned char)r; o[1] = (unsigned char)g; o[2] = (unsigned char)b; o[3] = (unsigned char)a;
        }
    }
    *width = (int)w;
    *height = (int)h;
    return IBTFO_OK;
}

int png_read_file(const char* path, std::vector<unsigned char>& rgba, int* width, int* height) {
    FILE* f = fopen(path, "rb");
    if (!f) { set_error("cannot open %s", path); return IBTFO_E_IO; }
    std::vector<unsigned char> buf;
    unsigned char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    return png_decode(buf.data(), buf.size(), rgba, width, height);
}

}  // namespace ibtfo

extern "C" int ibtfo_png_decode(const char* path, unsigned char* rgba, long long cap, int* width, int* height) {
    if (!path || !width || !height) { ibtfo::set_error("null argument"); return IBTFO_E_ARG; }
    std::vector<unsigned char> img;
    int rc = ibtfo::png_read_file(path, img, width, height);
    if (rc) return rc;
    if (rgba) {
        if ((long long)img.size() > cap) { ibtfo::set_error("ibtfo_png_decode: buffer of %lld bytes too small", cap); return IBTFO_E_ARG; }
        memcpy(rgba, img.data(), img.size());
    }
    return IBTFO_OK;
}
