// File formats and preprocessing of the reference's demo pipeline for C++ consumers,
// without OpenCV (header-only; PNG decoding needs zlib: link with -lz).
//
// The reference's test program (test_vfs/main.cpp:103-231) and DLL (dllhook/vfs.cpp)
// read their inputs with cv::imread(IMREAD_GRAYSCALE) (8-bit masks and images),
// cv::readOpticalFlow (Middlebury .flo vector fields), cv::equalizeHist and cv::resize.
// These functions restate those steps for the files the reference ships: grayscale,
// non-interlaced PNGs of 8 or 16 bits, binary PGM, and .flo.  vfs_amd/imageio.py is the
// same in Python.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <cmath>
#include <string>
#include <vector>

namespace vfs_io {

// Middlebury .flo: float tag 202021.25 ('PIEH'), int32 width, int32 height, then
// interleaved float32 (u, v) per pixel, row-major.
inline bool read_flo(const std::string &path, int &w, int &h, std::vector<float> &data)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    float tag = 0;
    int32_t ww = 0, hh = 0;
    bool ok = fread(&tag, 4, 1, f) == 1 && tag == 202021.25f && fread(&ww, 4, 1, f) == 1 &&
              fread(&hh, 4, 1, f) == 1 && ww > 0 && hh > 0;
    if (ok) {
        w = ww;
        h = hh;
        data.resize((size_t)w * h * 2);
        ok = fread(data.data(), sizeof(float), data.size(), f) == data.size();
    }
    fclose(f);
    return ok;
}

inline bool write_flo(const std::string &path, int w, int h, const std::vector<float> &data)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const float tag = 202021.25f;
    const int32_t ww = w, hh = h;
    bool ok = fwrite(&tag, 4, 1, f) == 1 && fwrite(&ww, 4, 1, f) == 1 && fwrite(&hh, 4, 1, f) == 1 &&
              fwrite(data.data(), sizeof(float), data.size(), f) == data.size();
    fclose(f);
    return ok;
}

// Binary PGM (P5, maxval 255).
inline bool read_pgm(const std::string &path, int &w, int &h, std::vector<uint8_t> &data)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[3] = {0};
    int maxv = 0;
    if (fscanf(f, "%2s %d %d %d", magic, &w, &h, &maxv) != 4 || strcmp(magic, "P5") || maxv != 255) {
        fclose(f);
        return false;
    }
    fgetc(f);  // the single whitespace after the header
    data.resize((size_t)w * h);
    const bool ok = fread(data.data(), 1, data.size(), f) == data.size();
    fclose(f);
    return ok;
}

// Grayscale PNG (colour type 0, bit depth 8 or 16, not interlaced): samples as uint16
// (8-bit images keep 0..255).  Returns false for any other PNG.
inline bool read_png_gray(const std::string &path, int &w, int &h, int &bit_depth, std::vector<uint16_t> &out)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<unsigned char> file;
    unsigned char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) file.insert(file.end(), buf, buf + n);
    fclose(f);
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (file.size() < 8 || memcmp(file.data(), sig, 8)) return false;
    auto be32 = [&](size_t p) {
        return (uint32_t)file[p] << 24 | (uint32_t)file[p + 1] << 16 | (uint32_t)file[p + 2] << 8 | file[p + 3];
    };
    std::vector<unsigned char> idat;
    int ctype = -1, interlace = -1;
    w = h = bit_depth = 0;
    for (size_t p = 8; p + 8 <= file.size();) {
        const uint32_t len = be32(p);
        if (p + 12 + len > file.size()) return false;
        const unsigned char *type = &file[p + 4], *body = &file[p + 8];
        if (!memcmp(type, "IHDR", 4) && len >= 13) {
            w = (int)be32(p + 8);
            h = (int)be32(p + 12);
            bit_depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + len;
    }
    if (ctype != 0 || interlace != 0 || (bit_depth != 8 && bit_depth != 16) || w <= 0 || h <= 0) return false;
    const int bpp = bit_depth / 8;
    const size_t stride = (size_t)w * bpp;
    std::vector<unsigned char> raw(h * (stride + 1));
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size()) return false;
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    out.resize((size_t)w * h);
    for (int y = 0; y < h; y++) {
        const unsigned char ft = raw[y * (stride + 1)];
        const unsigned char *line = &raw[y * (stride + 1) + 1];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int pr;
            switch (ft) {
            case 0: pr = 0; break;
            case 1: pr = a; break;
            case 2: pr = b; break;
            case 3: pr = (a + b) >> 1; break;
            case 4: {
                const int pp = a + b - c, pa = abs(pp - a), pb = abs(pp - b), pc = abs(pp - c);
                pr = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: return false;
            }
            cur[i] = (unsigned char)((line[i] + pr) & 255);
        }
        for (int x = 0; x < w; x++)
            out[(size_t)y * w + x] = bpp == 1 ? cur[x] : (uint16_t)(cur[2 * x] << 8 | cur[2 * x + 1]);
        prev.swap(cur);
    }
    return true;
}

// 8-bit grayscale PNG into bytes (cv::imread(path, IMREAD_GRAYSCALE) for such files).
inline bool read_png_u8(const std::string &path, int &w, int &h, std::vector<uint8_t> &out)
{
    std::vector<uint16_t> s;
    int depth = 0;
    if (!read_png_gray(path, w, h, depth, s) || depth != 8) return false;
    out.assign(s.begin(), s.end());
    return true;
}

// Histogram equalisation of an 8-bit image as cv::equalizeHist computes it: with i the
// lowest occupied level, lut[i] = 0 and lut[v] = round((cdf(v) - hist[i]) * 255 /
// (N - hist[i])) above it (float product, rounding half to even); an image of one level
// is left unchanged.
inline void equalize_hist(uint8_t *img, size_t n)
{
    size_t hist[256] = {0};
    for (size_t k = 0; k < n; k++) hist[img[k]]++;
    int i = 0;
    while (i < 256 && !hist[i]) i++;
    if (i == 256 || hist[i] == n) return;
    const float scale = 255.0f / (float)(n - hist[i]);
    uint8_t lut[256] = {0};
    size_t sum = 0;
    for (int v = i + 1; v < 256; v++) {
        sum += hist[v];
        const float r = std::nearbyint((float)sum * scale);
        lut[v] = (uint8_t)(r < 0.f ? 0.f : (r > 255.f ? 255.f : r));
    }
    for (size_t k = 0; k < n; k++) img[k] = lut[img[k]];
}

// Exact 2x downscale of an 8-bit image by bilinear interpolation at the coarse pixel
// centres (cv::resize, INTER_LINEAR): each output pixel is the rounded mean of its 2x2
// block, (a + b + c + d + 2) >> 2.
inline std::vector<uint8_t> resize_half(const uint8_t *img, int w, int h)
{
    const int ow = w / 2, oh = h / 2;
    std::vector<uint8_t> out((size_t)ow * oh);
    for (int y = 0; y < oh; y++)
        for (int x = 0; x < ow; x++) {
            const uint8_t *r0 = img + (size_t)(2 * y) * w + 2 * x, *r1 = r0 + w;
            out[(size_t)y * ow + x] = (uint8_t)((r0[0] + r0[1] + r1[0] + r1[1] + 2) >> 2);
        }
    return out;
}

}  // namespace vfs_io
