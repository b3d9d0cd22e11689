// Drop-in consumer of the StereoTgv class surface: replays the call sequence of the
// reference demo test_vfs/main.cpp:104-231 (test_t265_data) through
// include/stereotgv/stereotgv.h, without OpenCV.  Files replace the OpenCV I/O:
//   mask / images: binary PGM (P5, 8-bit); vector fields: Middlebury .flo
//   ("PIEH", int32 width, height, interleaved float32 x, y), as cv::readOpticalFlow.
// Outputs written to <outdir>: disparity.flo (copyDisparityToHost), u.raw
// (copy1DDisparityToHost), depth.raw and X.raw (copyStereoToHost), float32.
//
// usage: test_vfs_dropin <dir> <outdir> [nLevel fScale nWarpIters nSolverIters limitRange
//                                        equalize(0|1)]
//   <dir> holds i0.pgm i1.pgm mask.pgm tv.flo cv.flo
#include <stereotgv/stereotgv.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

bool read_pgm(const std::string &path, int &w, int &h, std::vector<uint8_t> &data)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[3] = {0};
    int maxv = 0;
    if (fscanf(f, "%2s %d %d %d", magic, &w, &h, &maxv) != 4 || strcmp(magic, "P5") || maxv != 255) {
        fclose(f);
        return false;
    }
    fgetc(f);  // single whitespace after the header
    data.resize((size_t)w * h);
    const bool ok = fread(data.data(), 1, data.size(), f) == data.size();
    fclose(f);
    return ok;
}

bool read_flo(const std::string &path, int &w, int &h, std::vector<float> &data)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    float tag = 0;
    int32_t ww = 0, hh = 0;
    bool ok = fread(&tag, 4, 1, f) == 1 && tag == 202021.25f && fread(&ww, 4, 1, f) == 1 && fread(&hh, 4, 1, f) == 1;
    if (ok) {
        w = ww;
        h = hh;
        data.resize((size_t)w * h * 2);
        ok = fread(data.data(), sizeof(float), data.size(), f) == data.size();
    }
    fclose(f);
    return ok;
}

bool write_flo(const std::string &path, int w, int h, const std::vector<float> &data)
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

bool write_raw(const std::string &path, const std::vector<float> &data)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = fwrite(data.data(), sizeof(float), data.size(), f) == data.size();
    fclose(f);
    return ok;
}

// Histogram equalisation of an 8-bit image (the preprocessing of main.cpp:196-197):
// lut[v] = round(255 * (cdf(v) - cdf_min) / (N - cdf_min)).
void equalize(std::vector<uint8_t> &img)
{
    size_t hist[256] = {0};
    for (uint8_t v : img) hist[v]++;
    size_t cdf = 0, cdf_min = 0;
    for (int v = 0; v < 256; v++)
        if (hist[v]) {
            cdf_min = hist[v];
            break;
        }
    const size_t n = img.size();
    uint8_t lut[256];
    for (int v = 0; v < 256; v++) {
        cdf += hist[v];
        const double num = (double)(cdf > cdf_min ? cdf - cdf_min : 0);
        const double den = (double)(n > cdf_min ? n - cdf_min : 1);
        int o = (int)(255.0 * num / den + 0.5);
        lut[v] = (uint8_t)(o > 255 ? 255 : o);
    }
    for (uint8_t &v : img) v = lut[v];
}

int fail(const char *what, StereoTgv *s)
{
    fprintf(stderr, "test_vfs_dropin: %s failed: %s\n", what, s ? s->lastError() : "");
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <dir> <outdir> [nLevel fScale nWarpIters nSolverIters limitRange equalize]\n",
                argv[0]);
        return 2;
    }
    const std::string dir = argv[1], out = argv[2];
    int nLevel = argc > 3 ? atoi(argv[3]) : 5;
    float fScale = argc > 4 ? (float)atof(argv[4]) : 2.0f;
    int nWarpIters = argc > 5 ? atoi(argv[5]) : 10;
    int nSolverIters = argc > 6 ? atoi(argv[6]) : 50;
    float limit = argc > 7 ? (float)atof(argv[7]) : 1.0f;
    bool equ = argc > 8 ? atoi(argv[8]) != 0 : false;

    std::vector<uint8_t> im1, im2, mask8;
    int w = 0, h = 0, w2, h2, wm, hm;
    if (!read_pgm(dir + "/i0.pgm", w, h, im1) || !read_pgm(dir + "/i1.pgm", w2, h2, im2) || w2 != w || h2 != h) {
        fprintf(stderr, "cannot read i0.pgm / i1.pgm\n");
        return 2;
    }
    // Camera parameters (main.cpp:116-128)
    float focalx = 285.722f, focaly = 286.759f, cx = 420.135f, cy = 403.394f;
    float d0 = -0.00659769f, d1 = 0.0473251f, d2 = -0.0458264f, d3 = 0.00897725f;
    float tx = -0.0641854f, ty = -0.000218299f, tz = 0.000111253f;
    float stereoScaling = 848.0f / (float)w;

    // VFS parameters (main.cpp:130-146)
    StereoTgv *stereotgv = new StereoTgv();
    float lambda = 5.0f, beta = 9.0f, gamma = 0.85f, alpha0 = 17.0f, alpha1 = 1.2f, timeStepLambda = 1.0f;
    stereotgv->limitRange = limit;
    if (stereotgv->initialize(w, h, beta, gamma, alpha0, alpha1, timeStepLambda, lambda, nLevel, fScale, nWarpIters,
                              nSolverIters))
        return fail("initialize", stereotgv);
    stereotgv->visualizeResults = true;

    // Load fisheye mask and vector fields (main.cpp:155-192)
    std::vector<float> tv, cv;
    int wt, ht, wc, hc;
    if (!read_pgm(dir + "/mask.pgm", wm, hm, mask8) || !read_flo(dir + "/tv.flo", wt, ht, tv) ||
        !read_flo(dir + "/cv.flo", wc, hc, cv) || wm != w || hm != h || wt != w || ht != h || wc != w || hc != h) {
        fprintf(stderr, "cannot read mask.pgm / tv.flo / cv.flo at %dx%d\n", w, h);
        return 2;
    }
    std::vector<float> mask(mask8.size());
    for (size_t i = 0; i < mask8.size(); i++) mask[i] = mask8[i] * (float)(1.0 / 255.0);  // convertTo(CV_32F, 1/255)
    if (stereotgv->copyMaskToDevice(mask.data())) return fail("copyMaskToDevice", stereotgv);
    if (stereotgv->loadVectorFields(tv.data(), cv.data())) return fail("loadVectorFields", stereotgv);

    // Solve stereo depth (main.cpp:194-211)
    if (equ) {
        equalize(im1);
        equalize(im2);
    }
    if (stereotgv->copyImagesToDevice(im1.data(), im2.data())) return fail("copyImagesToDevice", stereotgv);
    if (stereotgv->solveStereoForwardMasked()) return fail("solveStereoForwardMasked", stereotgv);
    std::vector<float> depth((size_t)w * h), X((size_t)w * h * 3);
    if (stereotgv->copyStereoToHost(depth.data(), X.data(), focalx / stereoScaling, focaly / stereoScaling,
                                    cx / stereoScaling, cy / stereoScaling, d0, d1, d2, d3, tx, ty, tz))
        return fail("copyStereoToHost", stereotgv);

    std::vector<float> disparity((size_t)w * h * 2), u((size_t)w * h);
    if (stereotgv->copyDisparityToHost(disparity.data())) return fail("copyDisparityToHost", stereotgv);
    if (stereotgv->copy1DDisparityToHost(u.data())) return fail("copy1DDisparityToHost", stereotgv);
    if (!write_flo(out + "/disparity.flo", w, h, disparity) || !write_raw(out + "/u.raw", u) ||
        !write_raw(out + "/depth.raw", depth) || !write_raw(out + "/X.raw", X)) {
        fprintf(stderr, "cannot write outputs to %s\n", out.c_str());
        return 2;
    }
    printf("test_vfs_dropin: solved %dx%d L%d W%d I%d, outputs in %s\n", w, h, nLevel, nWarpIters, nSolverIters,
           out.c_str());
    delete stereotgv;
    return 0;
}
