// Header-only C++ class `StereoTgv` over libvfs's C ABI (include/vfs.h).
//
// Drop-in for the reference class surface (menandro/vfs stereotgv/stereotgv.h:14-252)
// as its consumers use it (test_vfs/main.cpp:104-231, dllhook/vfs.cpp:3-48): same
// method names, argument order and meaning, and the public fields limitRange,
// visualizeResults, baseline and focal.  Differences, all at the boundary:
//   - raw-pointer overloads take caller-owned, tightly packed host arrays
//     (width x height [x channels]); cv::Mat overloads are compiled when OpenCV's
//     headers are present;
//   - methods return 0 on success and a negative VFS_ERR_* status on failure
//     (the reference returns 0 and exits the process on CUDA errors); lastError()
//     describes the failure;
//   - no HIP/CUDA type is exposed, so the reference's public device pointers are
//     replaced by readPlane().
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../vfs.h"

#if defined(__has_include)
#if __has_include(<opencv2/core.hpp>)
#include <opencv2/core.hpp>
#define STEREOTGV_HAVE_OPENCV 1
#endif
#endif

class StereoTgv {
public:
    // stereotgv.cu:32-36
    StereoTgv() { open(32, 1, 32); }
    // stereotgv.cu:38-42
    StereoTgv(int blockWidth, int blockHeight, int strideAlignment) { open(blockWidth, blockHeight, strideAlignment); }
    ~StereoTgv()
    {
        if (h_) vfs_destroy(h_);
    }
    StereoTgv(const StereoTgv &) = delete;
    StereoTgv &operator=(const StereoTgv &) = delete;

    // public fields (stereotgv.h:20-45)
    int BlockWidth = 32, BlockHeight = 1, StrideAlignment = 32;
    int width = 0, height = 0, stride = 0;
    float baseline = 0.f, focal = 0.f;
    bool visualizeResults = false;
    float beta = 0.f, gamma = 0.f, alpha0 = 0.f, alpha1 = 0.f, timestep_lambda = 0.f, lambda = 0.f, fScale = 0.f;
    int nLevels = 0, nSolverIters = 0, nWarpIters = 0;
    float limitRange = 1.0f;

    // stereotgv.h:107-109
    int initialize(int width_, int height_, float beta_, float gamma_, float alpha0_, float alpha1_,
                   float timestep_lambda_, float lambda_, int nLevels_, float fScale_, int nWarpIters_,
                   int nSolverIters_)
    {
        int rc = vfs_initialize(h_, width_, height_, beta_, gamma_, alpha0_, alpha1_, timestep_lambda_, lambda_,
                                nLevels_, fScale_, nWarpIters_, nSolverIters_);
        if (rc) return rc;
        width = width_;
        height = height_;
        int w, h;
        vfs_get_level_dims(h_, 0, &w, &h, &stride);
        beta = beta_;
        gamma = gamma_;
        alpha0 = alpha0_;
        alpha1 = alpha1_;
        timestep_lambda = timestep_lambda_;
        lambda = lambda_;
        nLevels = nLevels_;
        fScale = fScale_;
        nWarpIters = nWarpIters_;
        nSolverIters = nSolverIters_;
        return 0;
    }

    // stereotgv.h:110 -- tv, cv: width*height*2 floats (interleaved x, y)
    int loadVectorFields(const float *translationVector, const float *calibrationVector)
    {
        return vfs_load_vector_fields(h_, translationVector, 2 * width, calibrationVector, 2 * width);
    }
    // stereotgv.h:112
    int copyImagesToDevice(unsigned char *i0, unsigned char *i1) { return vfs_copy_images_u8(h_, i0, i1, width); }
    // CV_32F branch of stereotgv.h:113
    int copyImagesToDevice(const float *i0, const float *i1) { return vfs_copy_images_f32(h_, i0, i1, width); }
    // CV_8UC3 branch of stereotgv.h:113 (width*height*3 bytes)
    int copyImagesToDeviceRgb(const unsigned char *i0, const unsigned char *i1)
    {
        return vfs_copy_images_u8c3(h_, i0, i1, 3 * width);
    }
    // stereotgv.h:115 -- width*height floats, nonzero = inside the fisheye circle
    int copyMaskToDevice(const float *mask) { return vfs_copy_mask(h_, mask, width); }
    // stereotgv.h:116-117
    int solveStereoForward()
    {
        vfs_set_limit_range(h_, limitRange);
        return vfs_solve(h_);
    }
    int solveStereoForwardMasked()
    {
        vfs_set_limit_range(h_, limitRange);
        return vfs_solve_masked(h_);
    }
    // stereotgv.h:118-120 -- depth: width*height floats, X: width*height*3 floats
    int copyStereoToHost(float *depth, float *X, float focalx, float focaly, float cx, float cy, float d1, float d2,
                         float d3, float d4, float t1, float t2, float t3)
    {
        return vfs_copy_stereo(h_, depth, width, X, 3 * width, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3);
    }
    // stereotgv.h:121
    int copy1DDisparityToHost(float *u) { return vfs_copy_disparity_1d(h_, u, width); }
    // stereotgv.h:122 -- width*height*2 floats
    int copyDisparityToHost(float *flow) { return vfs_copy_flow(h_, flow, 2 * width); }
    // stereotgv.h:123 -- width*height*3 floats, (B, G, R) in [0, 1]
    int copyDisparityVisToHost(float *bgr, float flowScale)
    {
        return vfs_copy_disparity_vis(h_, bgr, 3 * width, flowScale);
    }
    // stereotgv.h:124
    int copyWarpedImageToHost(float *out) { return vfs_copy_warped_image(h_, out, width); }

    // Inspection of internal planes (the reference's public device pointers).
    int readPlane(const char *name, int level, std::vector<float> &out)
    {
        int w, h, s;
        int rc = vfs_get_level_dims(h_, level, &w, &h, &s);
        if (rc) return rc;
        out.resize((size_t)w * h * 4);
        rc = vfs_read_plane(h_, name, level, out.data());
        if (rc > 0) out.resize((size_t)w * h * rc);
        return rc;
    }
    const char *lastError() const { return vfs_last_error(h_); }
    vfs_solver *handle() const { return h_; }

#ifdef STEREOTGV_HAVE_OPENCV
    int loadVectorFields(cv::Mat translationVector, cv::Mat calibrationVector)
    {
        cv::Mat t = translationVector.isContinuous() ? translationVector : translationVector.clone();
        cv::Mat c = calibrationVector.isContinuous() ? calibrationVector : calibrationVector.clone();
        return loadVectorFields(t.ptr<float>(), c.ptr<float>());
    }
    int copyImagesToDevice(cv::Mat i0, cv::Mat i1)
    {
        cv::Mat a = i0.isContinuous() ? i0 : i0.clone();
        cv::Mat b = i1.isContinuous() ? i1 : i1.clone();
        if (a.type() == CV_8U) return copyImagesToDevice(a.ptr<unsigned char>(), b.ptr<unsigned char>());
        if (a.type() == CV_32F) return copyImagesToDevice(a.ptr<float>(), b.ptr<float>());
        if (a.type() == CV_8UC3) return copyImagesToDeviceRgb(a.ptr<unsigned char>(), b.ptr<unsigned char>());
        return VFS_ERR_ARG;
    }
    int copyMaskToDevice(cv::Mat mask)
    {
        cv::Mat m = mask.isContinuous() ? mask : mask.clone();
        return copyMaskToDevice(m.ptr<float>());
    }
    int copyStereoToHost(cv::Mat &croppedDepth, cv::Mat &croppedX, float focalx, float focaly, float cx, float cy,
                         float d1, float d2, float d3, float d4, float t1, float t2, float t3)
    {
        depth_.create(height, width, CV_32F);
        X_.create(height, width, CV_32FC3);
        int rc = copyStereoToHost(depth_.ptr<float>(), X_.ptr<float>(), focalx, focaly, cx, cy, d1, d2, d3, d4, t1,
                                  t2, t3);
        croppedDepth = depth_;
        croppedX = X_;
        return rc;
    }
    int copy1DDisparityToHost(cv::Mat &wCropped)
    {
        disp_.create(height, width, CV_32F);
        int rc = copy1DDisparityToHost(disp_.ptr<float>());
        wCropped = disp_;
        return rc;
    }
    int copyDisparityToHost(cv::Mat &wCropped)
    {
        flow_.create(height, width, CV_32FC2);
        int rc = copyDisparityToHost(flow_.ptr<float>());
        wCropped = flow_;
        return rc;
    }
    int copyDisparityVisToHost(cv::Mat &wCropped, float flowScale)
    {
        vis_.create(height, width, CV_32FC3);
        int rc = copyDisparityVisToHost(vis_.ptr<float>(), flowScale);
        wCropped = vis_;
        return rc;
    }
    int copyWarpedImageToHost(cv::Mat &wCropped)
    {
        cv::Mat out(height, width, CV_32F);
        int rc = copyWarpedImageToHost(out.ptr<float>());
        wCropped = out;
        return rc;
    }

private:
    cv::Mat depth_, X_, disp_, flow_, vis_;
#endif

private:
    vfs_solver *h_ = nullptr;

    void open(int bw, int bh, int align)
    {
        BlockWidth = bw;
        BlockHeight = bh;
        StrideAlignment = align;
        if (vfs_create_ex(&h_, bw, bh, align) != VFS_OK)
            throw std::runtime_error(std::string("StereoTgv: ") + vfs_last_error(nullptr));
    }
};
