// Launchers for the gfx950 StereoTgv kernels (vfs_kernels.hip).
// Every plane is a pitched row-major float (or float2 / float4) array; `w, h, s`
// are the level's width, height and stride (elements).  All launchers are
// asynchronous on `st`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vfs {

struct Dims {
    int w, h, s;
};

// Scalar parameters of one inner iteration (host-computed schedule, stereotgv.cu:449-477).
struct IterParams {
    float k1;      // alpha1 * sigma / eta_p
    float k0;      // alpha0 * sigma / eta_q
    float tau;
    float mu;
    float lambda;
    float alpha0;
    float alpha1;
    int first;     // 1: p and q were just zeroed (stereotgv.cu:431-432), read them as 0
    int write_tp;  // 1: also store Tp (read later only where a neighbour is outside the mask)
};

void launch_fill_zero(void *ptr, size_t bytes, hipStream_t st);
void launch_cv8u(const uint8_t *in, int in_pitch, float *out, Dims d, hipStream_t st);
void launch_cv8uc3(const uint8_t *in, int in_pitch, float *out, Dims d, hipStream_t st);
void launch_warp(const float *src, const float2 *uv, float *out, Dims d, hipStream_t st);
void launch_downscale2(const float *a, const float *b, Dims fine, float *oa, float *ob, Dims coarse,
                       hipStream_t st);
void launch_downscale_f2(const float2 *src, Dims fine, float2 *out, Dims coarse, hipStream_t st);
void launch_downscale_nn(const float *src, Dims fine, float *out, Dims coarse, hipStream_t st);
void launch_flowvec(const float *u, const float2 *tv, float2 *uv, Dims d, hipStream_t st);
void launch_tensor(const float *I0, const float *mask, float beta, float gamma, float alpha0,
                   float alpha1, int xlimit, float *a, float *b, float *c, Dims d, hipStream_t st);
void launch_warp_prep(const float2 *uv, const float2 *tv, const float *I1, const float *u,
                      float2 *tv2, float *i1w, float *u_last, float2 *v_, Dims d, hipStream_t st);
void launch_deriv(const float *I0, const float *I1w, const float2 *tv, float *Iu, float *Iz, Dims d,
                  hipStream_t st);
void launch_dual_tp(int masked, const float *mask, const float *u_, const float2 *v_, const float *a,
                    const float *b, const float *c, const IterParams &ip, float2 *p, float4 *q,
                    float2 *Tp, Dims d, hipStream_t st);
void launch_thresh_primal(int masked, const float *mask, const float2 *Tp, const float *u_,
                          const float2 *v_, const float *Iu, const float *Iz, const float *a,
                          const float *b, const float *c, const float2 *p, const float4 *q,
                          const float *u, const IterParams &ip, float *us, float *u_s, float2 *v_s,
                          Dims d, hipStream_t st);
void launch_median_update(const float *u, const float *u_last, const float2 *tv2, float limit,
                          float *out_u, float *u_, float2 *warpUV, Dims d, hipStream_t st);
void launch_update(float *u, const float *u_last, const float2 *tv2, float limit, float *u_,
                   float2 *warpUV, Dims d, hipStream_t st);
void launch_upscale3(const float *u, const float *u_, const float2 *uv, Dims coarse, float scale,
                     float *ou, float *ou_, float2 *ouv, Dims fine, hipStream_t st);
void launch_tgv_iter(int masked, const float *mask, const float *u_, const float2 *v_, const float *a,
                     const float *b, const float *c, const float *Iu, const float *Iz, const float *u,
                     const float2 *p_in, const float4 *q_in, float2 *p_out, float4 *q_out, float2 *Tp, float *us,
                     float *u_s, float2 *v_s, const IterParams &ip, Dims d, hipStream_t st);
void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
void launch_convert_kb(const float2 *uv, double focalx, double focaly, double cx, double cy, double d1,
                       double d2, double d3, double d4, double t1, double t2, double t3, float *X,
                       float *depth, Dims d, hipStream_t st);

}  // namespace vfs
