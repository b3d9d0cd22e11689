// Drop-in consumer of the StereoTgv class surface: replays the call sequence of the
// reference demo test_vfs/main.cpp:104-231 (test_t265_data) through
// include/stereotgv/stereotgv.h, without OpenCV; include/stereotgv/vfs_io.h replaces the
// OpenCV I/O: mask and images as 8-bit grayscale PNG (cv::imread IMREAD_GRAYSCALE) or
// binary PGM; vector fields as Middlebury .flo (cv::readOpticalFlow); equalizeHist and
// the 2x resize.
// Outputs written to <outdir>: disparity.flo (copyDisparityToHost), u.raw
// (copy1DDisparityToHost), depth.raw and X.raw (copyStereoToHost), vis.raw
// (copyDisparityVisToHost, flow scale 40 as main.cpp:221), float32.
//
// usage: test_vfs_dropin <dir> <outdir> [nLevel fScale nWarpIters nSolverIters limitRange
//                                        equalize(0|1) resize(0|1) camera(0|1)]
//   <dir> holds i0, i1 and mask as .png or .pgm, tv.flo and cv.flo.  resize = 1 halves
//   the (equalized) frames as main.cpp:196-199; camera 0 = the T265 intrinsics of
//   main.cpp:116-128, 1 = those of the wang run (main.cpp:235-247).
#include <stereotgv/stereotgv.h>
#include <stereotgv/vfs_io.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

bool exists(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (f) fclose(f);
    return f != nullptr;
}

// <name>.png (8-bit grayscale, cv::imread(IMREAD_GRAYSCALE)) or <name>.pgm
bool read_gray(const std::string &stem, int &w, int &h, std::vector<uint8_t> &data)
{
    if (exists(stem + ".png")) return vfs_io::read_png_u8(stem + ".png", w, h, data);
    return vfs_io::read_pgm(stem + ".pgm", w, h, data);
}

bool write_raw(const std::string &path, const std::vector<float> &data)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = fwrite(data.data(), sizeof(float), data.size(), f) == data.size();
    fclose(f);
    return ok;
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
        fprintf(stderr,
                "usage: %s <dir> <outdir> [nLevel fScale nWarpIters nSolverIters limitRange equalize resize camera]\n",
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
    bool half = argc > 9 ? atoi(argv[9]) != 0 : false;
    int camera = argc > 10 ? atoi(argv[10]) : 0;

    std::vector<uint8_t> im1, im2, mask8;
    int w = 0, h = 0, w2, h2, wm, hm;
    if (!read_gray(dir + "/i0", w, h, im1) || !read_gray(dir + "/i1", w2, h2, im2) || w2 != w || h2 != h) {
        fprintf(stderr, "cannot read i0 / i1\n");
        return 2;
    }
    // Camera parameters (main.cpp:116-128; the wang run: main.cpp:235-247)
    float focalx = 285.722f, focaly = 286.759f, cx = 420.135f, cy = 403.394f;
    float d0 = -0.00659769f, d1 = 0.0473251f, d2 = -0.0458264f, d3 = 0.00897725f;
    float tx = -0.0641854f, ty = -0.000218299f, tz = 0.000111253f;
    if (camera == 1) {
        focalx = 286.518f, focaly = 286.602f, cx = 423.142f, cy = 393.321f;
        d0 = -0.00727452f, d1 = 0.05200624f, d2 = -0.05105740f, d3 = 0.01108238f;
        tx = -0.0639505f, ty = 0.000184912f, tz = -0.0000766807f;
    }
    // Solve size: the frames, halved when resize = 1 (main.cpp:148-150)
    if (half) {
        w /= 2;
        h /= 2;
    }
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
    if (!read_gray(dir + "/mask", wm, hm, mask8) || !vfs_io::read_flo(dir + "/tv.flo", wt, ht, tv) ||
        !vfs_io::read_flo(dir + "/cv.flo", wc, hc, cv) || wm != w || hm != h || wt != w || ht != h || wc != w ||
        hc != h) {
        fprintf(stderr, "cannot read mask / tv.flo / cv.flo at %dx%d\n", w, h);
        return 2;
    }
    std::vector<float> mask(mask8.size());
    for (size_t i = 0; i < mask8.size(); i++) mask[i] = mask8[i] * (float)(1.0 / 255.0);  // convertTo(CV_32F, 1/255)
    if (stereotgv->copyMaskToDevice(mask.data())) return fail("copyMaskToDevice", stereotgv);
    if (stereotgv->loadVectorFields(tv.data(), cv.data())) return fail("loadVectorFields", stereotgv);

    // Solve stereo depth (main.cpp:194-211)
    if (equ) {
        vfs_io::equalize_hist(im1.data(), im1.size());
        vfs_io::equalize_hist(im2.data(), im2.size());
    }
    if (half) {
        im1 = vfs_io::resize_half(im1.data(), 2 * w, 2 * h);
        im2 = vfs_io::resize_half(im2.data(), 2 * w, 2 * h);
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
    // main.cpp:220-221
    std::vector<float> vis((size_t)w * h * 3);
    if (stereotgv->copyDisparityVisToHost(vis.data(), 40.0f)) return fail("copyDisparityVisToHost", stereotgv);
    if (!vfs_io::write_flo(out + "/disparity.flo", w, h, disparity) || !write_raw(out + "/u.raw", u) ||
        !write_raw(out + "/depth.raw", depth) || !write_raw(out + "/X.raw", X) || !write_raw(out + "/vis.raw", vis)) {
        fprintf(stderr, "cannot write outputs to %s\n", out.c_str());
        return 2;
    }
    printf("test_vfs_dropin: solved %dx%d L%d W%d I%d, outputs in %s\n", w, h, nLevel, nWarpIters, nSolverIters,
           out.c_str());
    delete stereotgv;
    return 0;
}
