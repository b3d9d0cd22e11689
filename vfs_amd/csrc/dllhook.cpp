// libvfs_dllhook: Create / Init / SolveStereo of the reference's DLL
// (dllhook/vfs.cpp:3-57) mapped onto the libvfs C ABI.  Host code only.
#include "../../include/vfs_dllhook.h"

#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/stereotgv/vfs_io.h"
#include "../../include/vfs.h"

struct vfs_dll_state {
    vfs_solver *s = nullptr;
    int width = 0, height = 0;
    float focal = 0, baseline = 0;
    bool ready = false;
    std::string err;
    std::vector<float> u;
};

namespace {

bool ends_with(const std::string &s, const char *suffix)
{
    const size_t n = strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

bool fail(vfs_dll_state *d, const std::string &what)
{
    d->err = what;
    if (d->s && vfs_last_error(d->s)[0]) d->err += std::string(": ") + vfs_last_error(d->s);
    d->ready = false;
    return false;
}

}  // namespace

extern "C" {

StereoTgvHandle *Create(void)
{
    vfs_dll_state *d = new vfs_dll_state;
    if (vfs_create(&d->s) != VFS_OK) {
        delete d;
        return nullptr;
    }
    return d;
}

void Init(StereoTgvHandle *d, int stereoWidth, int stereoHeight, float beta, float gamma, float alpha0, float alpha1,
          float timeStepLambda, float lambda, int nLevel, float fScale, int nWarpIters, int nSolverIters, float focal,
          float baseline, const char *fisheyeMaskFilename, const char *translationVectorFilename,
          const char *calibrationVectorFilename)
{
    if (!d) return;
    d->ready = false;
    d->err.clear();
    if (vfs_initialize(d->s, stereoWidth, stereoHeight, beta, gamma, alpha0, alpha1, timeStepLambda, lambda, nLevel,
                       fScale, nWarpIters, nSolverIters) != VFS_OK) {
        fail(d, "initialize");
        return;
    }
    d->width = stereoWidth;
    d->height = stereoHeight;
    d->baseline = baseline;
    d->focal = focal;
    // fisheye mask: imread(IMREAD_GRAYSCALE).convertTo(CV_32F, 1 / 255) (vfs.cpp:18-21)
    int w = 0, h = 0;
    std::vector<uint8_t> m8;
    const std::string mpath = fisheyeMaskFilename ? fisheyeMaskFilename : "";
    const bool mok = ends_with(mpath, ".pgm") ? vfs_io::read_pgm(mpath, w, h, m8) : vfs_io::read_png_u8(mpath, w, h, m8);
    if (!mok || w != stereoWidth || h != stereoHeight) {
        fail(d, "cannot read the fisheye mask " + mpath + " at the stereo size");
        return;
    }
    std::vector<float> mask(m8.size());
    for (size_t i = 0; i < m8.size(); i++) mask[i] = (float)m8[i] * (float)(1.0 / 255.0);
    if (vfs_copy_mask(d->s, mask.data(), w) != VFS_OK) {
        fail(d, "copyMaskToDevice");
        return;
    }
    // vector fields: readOpticalFlow (vfs.cpp:24-27)
    std::vector<float> tv, cv;
    int wt = 0, ht = 0, wc = 0, hc = 0;
    if (!translationVectorFilename || !calibrationVectorFilename ||
        !vfs_io::read_flo(translationVectorFilename, wt, ht, tv) ||
        !vfs_io::read_flo(calibrationVectorFilename, wc, hc, cv) || wt != w || ht != h || wc != w || hc != h) {
        fail(d, "cannot read the .flo vector fields at the stereo size");
        return;
    }
    if (vfs_load_vector_fields(d->s, tv.data(), 2 * w, cv.data(), 2 * w) != VFS_OK) {
        fail(d, "loadVectorFields");
        return;
    }
    d->u.resize((size_t)w * h);
    d->ready = true;
}

void SolveStereo(StereoTgvHandle *d, unsigned char *fs1, unsigned char *fs2, unsigned char *output, int stereoWidth,
                 int stereoHeight, float maxDepth)
{
    if (!d) return;
    if (!d->ready) {
        if (d->err.empty()) d->err = "SolveStereo before a successful Init";
        return;
    }
    if (!fs1 || !fs2 || stereoWidth != d->width || stereoHeight != d->height) {
        fail(d, "SolveStereo: bad frames or size");
        return;
    }
    const size_t n = (size_t)d->width * d->height;
    std::vector<uint8_t> e1(fs1, fs1 + n), e2(fs2, fs2 + n);
    vfs_io::equalize_hist(e1.data(), n);  // vfs.cpp:36-37
    vfs_io::equalize_hist(e2.data(), n);
    if (vfs_copy_images_u8(d->s, e1.data(), e2.data(), d->width) != VFS_OK || vfs_solve_masked(d->s) != VFS_OK ||
        vfs_copy_disparity_1d(d->s, d->u.data(), d->width) != VFS_OK) {
        fail(d, "solve");
        return;
    }
    if (output) {
        const float fb = d->focal * d->baseline;
        for (size_t i = 0; i < n; i++) {
            const float u = d->u[i];
            float v = 0.f;
            if (u > 0.f && maxDepth > 0.f) {
                float z = fb / u;
                z = z < maxDepth ? z : maxDepth;
                v = nearbyintf(255.f * z / maxDepth);
            }
            output[i] = (unsigned char)(v < 0.f ? 0.f : (v > 255.f ? 255.f : v));
        }
    }
    d->err.clear();
}

const char *VfsDllLastError(const StereoTgvHandle *d) { return d ? d->err.c_str() : "null handle"; }

void VfsDllDestroy(StereoTgvHandle *d)
{
    if (!d) return;
    vfs_destroy(d->s);
    delete d;
}

}  // extern "C"
