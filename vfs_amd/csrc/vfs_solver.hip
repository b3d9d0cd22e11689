// Host side of libvfs: the StereoTgv solver object behind the C ABI of include/vfs.h.
//
// Buffer model.  The reference allocates every work plane at level-0 size and reuses it
// at each pyramid level with that level's stride, swapping pointers between ping-pong
// partners on the host (stereotgv.cu:44-168, 389-515).  The same model is kept here so
// that values left outside the mask by earlier iterations and coarser levels are read
// exactly where the reference reads them.  Every work plane lives in one device arena
// that is zeroed (by a fill kernel) at the start of every solve: each solve equals the
// reference's first solve after initialize() on zeroed memory.
//
// Execution.  The solve schedule is open-loop (fixed iteration counts, host-computed
// step sizes, stereotgv.cu:449-477), so the whole solve is issued once under HIP stream
// capture and replayed as one graph launch per frame.  The pointer swaps are resolved
// while capturing; the roles at the end of the capture are recorded for the copy-outs.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/vfs.h"
#include "vfs_kernels.h"

using vfs::Dims;

namespace {

constexpr int kMaxLevels = 16;

}  // namespace

struct vfs_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    int blockWidth = 32, blockHeight = 1, strideAlign = 32;
    bool initialized = false;
    bool has_mask = false, has_fields = false, has_images = false;
    std::string err;

    // parameters (stereotgv.h:35-45)
    int width = 0, height = 0, stride = 0, nLevels = 0, nWarpIters = 0, nSolverIters = 0;
    float beta = 0, gamma = 0, alpha0 = 0, alpha1 = 0, timestep_lambda = 0, lambda = 0, fScale = 0;
    float limitRange = 1.0f;
    Dims lv[kMaxLevels] = {};

    // persistent device planes (not zeroed per solve)
    float *pI0[kMaxLevels] = {};   // I0 pyramid (level 0 = input)
    float *pI1[kMaxLevels] = {};   // I1 pyramid; level 0 = calibrated I1
    float *in_i1 = nullptr;        // raw I1 input slot (warped by cv into pI1[0])
    float *pMask[kMaxLevels] = {};
    float2 *pTv[kMaxLevels] = {};
    float2 *cv = nullptr;
    uint8_t *staging = nullptr;    // u8 / u8c3 input upload
    size_t staging_bytes = 0;
    float *outX = nullptr, *outDepth = nullptr;
    bool raw_i1_fresh = false;     // in_i1 holds a frame not yet calibrated

    // work arena (zeroed at every solve)
    void *arena = nullptr;
    size_t arena_bytes = 0;
    float *bufU[2] = {}, *bufU_[2] = {}, *u_last = nullptr, *i1warp = nullptr, *Iu = nullptr, *Iz = nullptr;
    float *a = nullptr, *b = nullptr, *c = nullptr;
    float2 *bufV_[2] = {}, *bufP[2] = {}, *Tp = nullptr, *tv2 = nullptr, *bufUV[2] = {};
    float4 *bufQ[2] = {};

    // roles after the last issued solve
    float *out_u = nullptr, *out_u_ = nullptr;
    float2 *out_uv = nullptr, *out_v_ = nullptr, *out_p = nullptr;
    float4 *out_q = nullptr;
    bool fused = true;  // one launch per inner iteration (else dual+Tp, threshold+primal)

    // graph replay
    bool use_graph = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    float graph_limit[2] = {0, 0};
    struct Roles {
        float *u, *u_;
        float2 *uv, *v_, *p;
        float4 *q;
    };
    Roles graph_roles[2] = {};
    bool graph_fused[2] = {true, true};

    std::vector<void *> allocs;
};

namespace {

thread_local std::string g_create_error;

int fail(vfs_solver *s, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(vfs_solver *s, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    else g_create_error = buf;
    return code;
}

#define HIPCHK(s, expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail((s), VFS_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int align_up(int n, int m)
{
    int r = n % m;
    return r ? n + m - r : n;
}

int dev_alloc(vfs_solver *s, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(s, VFS_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    s->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes, s->stream);
    if (e != hipSuccess) return fail(s, VFS_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    return VFS_OK;
}

void drop_graphs(vfs_solver *s)
{
    for (int m = 0; m < 2; m++) {
        if (s->graph[m]) (void)hipGraphExecDestroy(s->graph[m]);
        s->graph[m] = nullptr;
    }
}

void release(vfs_solver *s)
{
    drop_graphs(s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    for (int l = 0; l < kMaxLevels; l++) {
        s->pI0[l] = s->pI1[l] = s->pMask[l] = nullptr;
        s->pTv[l] = nullptr;
    }
    s->in_i1 = nullptr;
    s->cv = nullptr;
    s->staging = nullptr;
    s->staging_bytes = 0;
    s->arena = nullptr;
    s->initialized = s->has_mask = s->has_fields = s->has_images = false;
}

int set_device(vfs_solver *s)
{
    HIPCHK(s, hipSetDevice(s->device));
    return VFS_OK;
}

int check_launch(vfs_solver *s)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(s, VFS_ERR_DEVICE, "kernel launch: %s", hipGetErrorString(e));
    return VFS_OK;
}

// Issues the whole solve (stereotgv.cu:389-515 when masked, :256-387 otherwise) on the
// solver's stream, starting from the canonical buffer roles.  Records the final roles.
int issue_solve(vfs_solver *s, int masked, vfs_solver::Roles *fin)
{
    hipStream_t st = s->stream;
    const Dims d0 = s->lv[0];
    // fresh state: zero the work arena
    vfs::launch_fill_zero(s->arena, s->arena_bytes, st);
    // calibration warp into the level-0 I1 slot (stereotgv.cu:392-393)
    vfs::launch_warp(s->in_i1, s->cv, s->pI1[0], d0, st);
    // pyramids (stereotgv.cu:400-405)
    for (int l = 1; l < s->nLevels; l++)
        vfs::launch_downscale2(s->pI0[l - 1], s->pI1[l - 1], s->lv[l - 1], s->pI0[l], s->pI1[l], s->lv[l], st);

    float *u = s->bufU[0], *us = s->bufU[1];
    float *u_ = s->bufU_[0], *u_s = s->bufU_[1];
    float2 *v_ = s->bufV_[0], *v_s = s->bufV_[1];
    float2 *uv = s->bufUV[0], *uvs = s->bufUV[1];
    int pq = 0;  // which p / q buffer holds the current duals

    for (int l = s->nLevels - 1; l >= 0; l--) {
        const Dims d = s->lv[l];
        const float *mask = masked ? s->pMask[l] : nullptr;
        if (l == s->nLevels - 1) vfs::launch_flowvec(u, s->pTv[l], uv, d, st);
        const int xlimit = s->blockWidth * ((d.h + s->blockWidth - 1) / s->blockWidth);
        vfs::launch_tensor(s->pI0[l], mask, s->beta, s->gamma, s->alpha0, s->alpha1, xlimit, s->a, s->b, s->c, d,
                           st);
        for (int wi = 0; wi < s->nWarpIters; wi++) {
            vfs::launch_warp_prep(uv, s->pTv[l], s->pI1[l], u, s->tv2, s->i1warp, s->u_last, v_, d, st);
            vfs::launch_deriv(s->pI0[l], s->i1warp, s->pTv[l], s->Iu, s->Iz, d, st);
            float tau = 1.0f;
            float sigma = 1.0f / tau;
            for (int it = 0; it < s->nSolverIters; it++) {
                float mu;
                if (sigma < 1000.0f) mu = 1.0f / sqrtf(1 + 0.7f * tau * s->timestep_lambda);
                else mu = 1;
                vfs::IterParams ip;
                ip.k1 = s->alpha1 * sigma / 3.0f;
                ip.k0 = s->alpha0 * sigma / 2.0f;
                ip.tau = tau;
                ip.mu = mu;
                ip.lambda = s->lambda;
                ip.alpha0 = s->alpha0;
                ip.alpha1 = s->alpha1;
                ip.first = (it == 0);
                ip.write_tp = (it == s->nSolverIters - 1);
                if (s->fused) {
                    vfs::launch_tgv_iter(masked, mask, u_, v_, s->a, s->b, s->c, s->Iu, s->Iz, u, s->bufP[pq],
                                         s->bufQ[pq], s->bufP[pq ^ 1], s->bufQ[pq ^ 1], s->Tp, us, u_s, v_s, ip, d,
                                         st);
                    pq ^= 1;
                } else {
                    vfs::launch_dual_tp(masked, mask, u_, v_, s->a, s->b, s->c, ip, s->bufP[pq], s->bufQ[pq], s->Tp,
                                        d, st);
                    vfs::launch_thresh_primal(masked, mask, s->Tp, u_, v_, s->Iu, s->Iz, s->a, s->b, s->c,
                                              s->bufP[pq], s->bufQ[pq], u, ip, us, u_s, v_s, d, st);
                }
                std::swap(u, us);
                std::swap(u_, u_s);
                std::swap(v_, v_s);
                sigma = sigma / mu;
                tau = tau * mu;
            }
            if (masked) {
                vfs::launch_median_update(u, s->u_last, s->tv2, s->limitRange, us, u_, uv, d, st);
                std::swap(u, us);
            } else {
                vfs::launch_update(u, s->u_last, s->tv2, s->limitRange, u_, uv, d, st);
            }
        }
        if (l > 0) {
            vfs::launch_upscale3(u, u_, uv, d, s->fScale, us, u_s, uvs, s->lv[l - 1], st);
            std::swap(u, us);
            std::swap(u_, u_s);
            std::swap(uv, uvs);
        }
    }
    fin->u = u;
    fin->u_ = u_;
    fin->uv = uv;
    fin->v_ = v_;
    fin->p = s->bufP[pq];
    fin->q = s->bufQ[pq];
    return check_launch(s);
}

int solve_common(vfs_solver *s, int masked)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "solve before initialize");
    if (!s->has_fields) return fail(s, VFS_ERR_STATE, "solve before loadVectorFields");
    if (masked && !s->has_mask) return fail(s, VFS_ERR_STATE, "masked solve before copyMaskToDevice");
    if (!s->has_images) return fail(s, VFS_ERR_STATE, "solve before copyImagesToDevice");
    int rc = set_device(s);
    if (rc) return rc;
    const int m = masked ? 1 : 0;
    // A second solve without new images re-calibrates the already calibrated I1,
    // as the reference's swap of pI1[0] with d_i1calibrated does (stereotgv.cu:392-393).
    if (!s->raw_i1_fresh) {
        const Dims d0 = s->lv[0];
        HIPCHK(s, hipMemcpyAsync(s->in_i1, s->pI1[0], sizeof(float) * (size_t)d0.s * d0.h,
                                 hipMemcpyDeviceToDevice, s->stream));
    }
    vfs_solver::Roles roles{};
    if (!s->use_graph) {
        rc = issue_solve(s, masked, &roles);
        if (rc) return rc;
    } else {
        if (s->graph[m] && (s->graph_limit[m] != s->limitRange || s->graph_fused[m] != s->fused)) {
            (void)hipGraphExecDestroy(s->graph[m]);
            s->graph[m] = nullptr;
        }
        if (!s->graph[m]) {
            HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            rc = issue_solve(s, masked, &roles);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(s, VFS_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&s->graph[m], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->graph[m] = nullptr;
                return fail(s, VFS_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
            }
            s->graph_limit[m] = s->limitRange;
            s->graph_fused[m] = s->fused;
            s->graph_roles[m] = roles;
        }
        HIPCHK(s, hipGraphLaunch(s->graph[m], s->stream));
        roles = s->graph_roles[m];
    }
    s->out_u = roles.u;
    s->out_u_ = roles.u_;
    s->out_uv = roles.uv;
    s->out_v_ = roles.v_;
    s->out_p = roles.p;
    s->out_q = roles.q;
    s->raw_i1_fresh = false;
    return VFS_OK;
}

int copy_out_plane(vfs_solver *s, const void *dev, int nch, int level, float *out, int pitch_elems)
{
    const Dims d = s->lv[level];
    HIPCHK(s, hipMemcpy2DAsync(out, sizeof(float) * (size_t)pitch_elems, dev, sizeof(float) * nch * (size_t)d.s,
                               sizeof(float) * nch * (size_t)d.w, d.h, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return VFS_OK;
}

int upload_u8(vfs_solver *s, const uint8_t *i0, const uint8_t *i1, int pitch, int nch)
{
    const size_t row = (size_t)s->width * nch;
    const size_t need = 2 * row * s->height;
    if (s->staging_bytes < need) {
        int rc = dev_alloc(s, (void **)&s->staging, need);
        if (rc) return rc;
        s->staging_bytes = need;
    }
    HIPCHK(s, hipMemcpy2DAsync(s->staging, row, i0, pitch, row, s->height, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpy2DAsync(s->staging + row * s->height, row, i1, pitch, row, s->height, hipMemcpyHostToDevice,
                               s->stream));
    const Dims d0 = s->lv[0];
    if (nch == 1) {
        vfs::launch_cv8u(s->staging, (int)row, s->pI0[0], d0, s->stream);
        vfs::launch_cv8u(s->staging + row * s->height, (int)row, s->in_i1, d0, s->stream);
    } else {
        vfs::launch_cv8uc3(s->staging, (int)row, s->pI0[0], d0, s->stream);
        vfs::launch_cv8uc3(s->staging + row * s->height, (int)row, s->in_i1, d0, s->stream);
    }
    int rc = check_launch(s);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_images = true;
    s->raw_i1_fresh = true;
    return VFS_OK;
}

}  // namespace

extern "C" {

int vfs_create_ex(vfs_solver **out, int blockWidth, int blockHeight, int strideAlignment)
{
    if (!out) return fail(nullptr, VFS_ERR_ARG, "null handle pointer");
    *out = nullptr;
    if (blockWidth < 1 || blockHeight < 1 || strideAlignment < 1)
        return fail(nullptr, VFS_ERR_ARG, "block/alignment must be positive");
    vfs_solver *s = new (std::nothrow) vfs_solver();
    if (!s) return fail(nullptr, VFS_ERR_NOMEM, "host allocation failed");
    s->blockWidth = blockWidth;
    s->blockHeight = blockHeight;
    s->strideAlign = strideAlignment;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fail(nullptr, VFS_ERR_DEVICE, "no usable HIP device: %s", hipGetErrorString(e));
        delete s;
        return VFS_ERR_DEVICE;
    }
    *out = s;
    return VFS_OK;
}

int vfs_create(vfs_solver **out) { return vfs_create_ex(out, 32, 1, 32); }

int vfs_destroy(vfs_solver *s)
{
    if (!s) return VFS_ERR_ARG;
    (void)hipSetDevice(s->device);
    release(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return VFS_OK;
}

const char *vfs_last_error(const vfs_solver *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int vfs_initialize(vfs_solver *s, int width, int height, float beta, float gamma, float alpha0, float alpha1,
                   float timestep_lambda, float lambda, int nLevels, float fScale, int nWarpIters, int nSolverIters)
{
    if (!s) return VFS_ERR_ARG;
    s->err.clear();
    if (width < 1 || height < 1) return fail(s, VFS_ERR_ARG, "bad size %dx%d", width, height);
    if (nLevels < 1 || nLevels > kMaxLevels) return fail(s, VFS_ERR_ARG, "nLevels %d outside [1,%d]", nLevels, kMaxLevels);
    if (nWarpIters < 0 || nSolverIters < 0) return fail(s, VFS_ERR_ARG, "negative iteration count");
    if (!(fScale > 0.f)) return fail(s, VFS_ERR_ARG, "fScale must be positive");
    int rc = set_device(s);
    if (rc) return rc;
    release(s);
    s->width = width;
    s->height = height;
    s->stride = align_up(width, s->strideAlign);
    s->beta = beta;
    s->gamma = gamma;
    s->alpha0 = alpha0;
    s->alpha1 = alpha1;
    s->timestep_lambda = timestep_lambda;
    s->lambda = lambda;
    s->fScale = fScale;
    s->nLevels = nLevels;
    s->nWarpIters = nWarpIters;
    s->nSolverIters = nSolverIters;
    // pyramid dims (stereotgv.cu:74-95)
    int nh = height, nw = width;
    for (int l = 0; l < nLevels; l++) {
        if (nw < 1 || nh < 1) return fail(s, VFS_ERR_ARG, "level %d collapses to %dx%d", l, nw, nh);
        s->lv[l] = Dims{nw, nh, align_up(nw, s->strideAlign)};
        nh = (int)((float)nh / fScale);
        nw = (int)((float)nw / fScale);
    }
    const size_t N = (size_t)s->stride * height;
    for (int l = 0; l < nLevels; l++) {
        const size_t n = (size_t)s->lv[l].s * s->lv[l].h;
        if ((rc = dev_alloc(s, (void **)&s->pI0[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, (void **)&s->pI1[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, (void **)&s->pMask[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, (void **)&s->pTv[l], n * sizeof(float2)))) return rc;
    }
    if ((rc = dev_alloc(s, (void **)&s->in_i1, N * sizeof(float)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->cv, N * sizeof(float2)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->outX, 3 * N * sizeof(float)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->outDepth, N * sizeof(float)))) return rc;
    // work arena: 37 floats per level-0 pixel, each plane 256 B aligned
    const size_t plane = (N * sizeof(float) + 255) & ~(size_t)255;
    const int nplanes = 5 /*u,us,u_,u_s,u_last*/ + 4 /*i1warp,Iu,Iz,.*/ + 3 /*a,b,c*/ + 4 /*v_,v_s*/ + 4 /*p x2*/ +
                        2 /*Tp*/ + 2 /*tv2*/ + 4 /*uv,uvs*/ + 8 /*q x2*/;
    s->arena_bytes = plane * nplanes;
    if ((rc = dev_alloc(s, &s->arena, s->arena_bytes))) return rc;
    char *ar = (char *)s->arena;
    auto take = [&](int nfl) {
        char *p = ar;
        ar += plane * nfl;
        return (void *)p;
    };
    s->bufU[0] = (float *)take(1);
    s->bufU[1] = (float *)take(1);
    s->bufU_[0] = (float *)take(1);
    s->bufU_[1] = (float *)take(1);
    s->u_last = (float *)take(1);
    s->i1warp = (float *)take(1);
    s->Iu = (float *)take(1);
    s->Iz = (float *)take(1);
    take(1);  // spare (keeps the 2- and 4-channel planes 512 B aligned)
    s->a = (float *)take(1);
    s->b = (float *)take(1);
    s->c = (float *)take(1);
    s->bufV_[0] = (float2 *)take(2);
    s->bufV_[1] = (float2 *)take(2);
    s->bufP[0] = (float2 *)take(2);
    s->bufP[1] = (float2 *)take(2);
    s->Tp = (float2 *)take(2);
    s->tv2 = (float2 *)take(2);
    s->bufUV[0] = (float2 *)take(2);
    s->bufUV[1] = (float2 *)take(2);
    s->bufQ[0] = (float4 *)take(4);
    s->bufQ[1] = (float4 *)take(4);
    s->out_u = s->bufU[0];
    s->out_u_ = s->bufU_[0];
    s->out_uv = s->bufUV[0];
    s->out_v_ = s->bufV_[0];
    s->out_p = s->bufP[0];
    s->out_q = s->bufQ[0];
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->initialized = true;
    return VFS_OK;
}

int vfs_set_limit_range(vfs_solver *s, float limit)
{
    if (!s) return VFS_ERR_ARG;
    s->limitRange = limit;
    return VFS_OK;
}

int vfs_get_limit_range(const vfs_solver *s, float *limit)
{
    if (!s || !limit) return VFS_ERR_ARG;
    *limit = s->limitRange;
    return VFS_OK;
}

int vfs_copy_mask(vfs_solver *s, const float *mask, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copyMaskToDevice before initialize");
    if (!mask || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad mask pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->pMask[0], sizeof(float) * (size_t)d0.s, mask, sizeof(float) * (size_t)pitch,
                               sizeof(float) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    for (int l = 1; l < s->nLevels; l++)
        vfs::launch_downscale_nn(s->pMask[l - 1], s->lv[l - 1], s->pMask[l], s->lv[l], s->stream);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_mask = true;
    return VFS_OK;
}

int vfs_load_vector_fields(vfs_solver *s, const float *tv, int tv_pitch, const float *cvh, int cv_pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "loadVectorFields before initialize");
    if (!tv || !cvh || tv_pitch < 2 * s->width || cv_pitch < 2 * s->width)
        return fail(s, VFS_ERR_ARG, "bad vector field pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->pTv[0], sizeof(float2) * (size_t)d0.s, tv, sizeof(float) * (size_t)tv_pitch,
                               sizeof(float2) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpy2DAsync(s->cv, sizeof(float2) * (size_t)d0.s, cvh, sizeof(float) * (size_t)cv_pitch,
                               sizeof(float2) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    for (int l = 1; l < s->nLevels; l++)
        vfs::launch_downscale_f2(s->pTv[l - 1], s->lv[l - 1], s->pTv[l], s->lv[l], s->stream);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_fields = true;
    return VFS_OK;
}

int vfs_copy_images_f32(vfs_solver *s, const float *i0, const float *i1, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copyImagesToDevice before initialize");
    if (!i0 || !i1 || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad image pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->pI0[0], sizeof(float) * (size_t)d0.s, i0, sizeof(float) * (size_t)pitch,
                               sizeof(float) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpy2DAsync(s->in_i1, sizeof(float) * (size_t)d0.s, i1, sizeof(float) * (size_t)pitch,
                               sizeof(float) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_images = true;
    s->raw_i1_fresh = true;
    return VFS_OK;
}

int vfs_copy_images_u8(vfs_solver *s, const uint8_t *i0, const uint8_t *i1, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copyImagesToDevice before initialize");
    if (!i0 || !i1 || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad image pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return upload_u8(s, i0, i1, pitch, 1);
}

int vfs_copy_images_u8c3(vfs_solver *s, const uint8_t *i0, const uint8_t *i1, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copyImagesToDevice before initialize");
    if (!i0 || !i1 || pitch < 3 * s->width) return fail(s, VFS_ERR_ARG, "bad image pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return upload_u8(s, i0, i1, pitch, 3);
}

int vfs_solve_masked(vfs_solver *s) { return solve_common(s, 1); }
int vfs_solve(vfs_solver *s) { return solve_common(s, 0); }

int vfs_copy_disparity_1d(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out_plane(s, s->out_u, 1, 0, out, pitch);
}

int vfs_copy_flow(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < 2 * s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out_plane(s, s->out_uv, 2, 0, out, pitch);
}

int vfs_copy_warped_image(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out_plane(s, s->i1warp, 1, 0, out, pitch);
}

int vfs_copy_stereo(vfs_solver *s, float *depth, int depth_pitch, float *X, int X_pitch, float focalx, float focaly,
                    float cx, float cy, float d1, float d2, float d3, float d4, float t1, float t2, float t3)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if ((depth && depth_pitch < s->width) || (X && X_pitch < 3 * s->width))
        return fail(s, VFS_ERR_ARG, "bad output pitch");
    int rc = set_device(s);
    if (rc) return rc;
    const Dims d0 = s->lv[0];
    vfs::launch_convert_kb(s->out_uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3, s->outX, s->outDepth, d0,
                           s->stream);
    if ((rc = check_launch(s))) return rc;
    if (depth && (rc = copy_out_plane(s, s->outDepth, 1, 0, depth, depth_pitch))) return rc;
    if (X && (rc = copy_out_plane(s, s->outX, 3, 0, X, X_pitch))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return VFS_OK;
}

int vfs_get_level_dims(const vfs_solver *s, int level, int *w, int *h, int *stride)
{
    if (!s || !w || !h || !stride) return VFS_ERR_ARG;
    if (!s->initialized || level < 0 || level >= s->nLevels) return VFS_ERR_ARG;
    *w = s->lv[level].w;
    *h = s->lv[level].h;
    *stride = s->lv[level].s;
    return VFS_OK;
}

int vfs_read_plane(vfs_solver *s, const char *name, int level, float *out)
{
    if (!s || !name || !out) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "read before initialize");
    if (level < 0 || level >= s->nLevels) return fail(s, VFS_ERR_ARG, "level %d out of range", level);
    const void *src = nullptr;
    int nch = 1;
    const std::string n(name);
    if (n == "u") src = s->out_u;
    else if (n == "u_") src = s->out_u_;
    else if (n == "u_last") src = s->u_last;
    else if (n == "a") src = s->a;
    else if (n == "b") src = s->b;
    else if (n == "c") src = s->c;
    else if (n == "Iu") src = s->Iu;
    else if (n == "Iz") src = s->Iz;
    else if (n == "i1warp") src = s->i1warp;
    else if (n == "I0") src = s->pI0[level];
    else if (n == "I1") src = s->pI1[level];
    else if (n == "mask") src = s->pMask[level];
    else if (n == "warpUV") { src = s->out_uv; nch = 2; }
    else if (n == "tv2") { src = s->tv2; nch = 2; }
    else if (n == "tv") { src = s->pTv[level]; nch = 2; }
    else if (n == "p") { src = s->out_p; nch = 2; }
    else if (n == "v_") { src = s->out_v_; nch = 2; }
    else if (n == "Tp") { src = s->Tp; nch = 2; }
    else if (n == "q") { src = s->out_q; nch = 4; }
    if (!src) return fail(s, VFS_ERR_ARG, "unknown plane '%s'", name);
    int rc = set_device(s);
    if (rc) return rc;
    rc = copy_out_plane(s, src, nch, level, out, s->lv[level].w * nch);
    return rc ? rc : nch;
}

int vfs_synchronize(vfs_solver *s)
{
    if (!s) return VFS_ERR_ARG;
    int rc = set_device(s);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return VFS_OK;
}

int vfs_set_option(vfs_solver *s, const char *key, int value)
{
    if (!s || !key) return VFS_ERR_ARG;
    if (!strcmp(key, "graph")) {
        s->use_graph = value != 0;
        return VFS_OK;
    }
    if (!strcmp(key, "fused")) {
        s->fused = value != 0;
        return VFS_OK;
    }
    return fail(s, VFS_ERR_ARG, "unknown option '%s'", key);
}

int vfs_time_solves(vfs_solver *s, int masked, int n, float *ms_total)
{
    if (!s || !ms_total || n < 1) return VFS_ERR_ARG;
    int rc = set_device(s);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(s, hipEventCreate(&e0));
    HIPCHK(s, hipEventCreate(&e1));
    HIPCHK(s, hipEventRecord(e0, s->stream));
    for (int i = 0; i < n && !rc; i++) rc = solve_common(s, masked);
    HIPCHK(s, hipEventRecord(e1, s->stream));
    HIPCHK(s, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(s, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms_total = ms;
    return VFS_OK;
}

int vfs_time_inner_iterations(vfs_solver *s, int level, int reps, float *ms_total, int *launches_per_iter)
{
    if (!s || !ms_total || reps < 1) return VFS_ERR_ARG;
    if (!s->initialized || level < 0 || level >= s->nLevels) return fail(s, VFS_ERR_ARG, "bad level");
    int rc = set_device(s);
    if (rc) return rc;
    const Dims d = s->lv[level];
    const float *mask = s->pMask[level];
    hipEvent_t e0, e1;
    HIPCHK(s, hipEventCreate(&e0));
    HIPCHK(s, hipEventCreate(&e1));
    float *u = s->bufU[0], *us = s->bufU[1], *u_ = s->bufU_[0], *u_s = s->bufU_[1];
    float2 *v_ = s->bufV_[0], *v_s = s->bufV_[1];
    HIPCHK(s, hipEventRecord(e0, s->stream));
    float tau = 1.0f, sigma = 1.0f;
    for (int it = 0; it < reps; it++) {
        const float mu = 1.0f / sqrtf(1 + 0.7f * tau * s->timestep_lambda);
        vfs::IterParams ip{s->alpha1 * sigma / 3.0f, s->alpha0 * sigma / 2.0f, tau, mu, s->lambda, s->alpha0,
                           s->alpha1, 0, 0};
        if (s->fused) {
            vfs::launch_tgv_iter(1, mask, u_, v_, s->a, s->b, s->c, s->Iu, s->Iz, u, s->bufP[it & 1], s->bufQ[it & 1],
                                 s->bufP[(it & 1) ^ 1], s->bufQ[(it & 1) ^ 1], s->Tp, us, u_s, v_s, ip, d, s->stream);
        } else {
            vfs::launch_dual_tp(1, mask, u_, v_, s->a, s->b, s->c, ip, s->bufP[0], s->bufQ[0], s->Tp, d, s->stream);
            vfs::launch_thresh_primal(1, mask, s->Tp, u_, v_, s->Iu, s->Iz, s->a, s->b, s->c, s->bufP[0], s->bufQ[0],
                                      u, ip, us, u_s, v_s, d, s->stream);
        }
        std::swap(u, us);
        std::swap(u_, u_s);
        std::swap(v_, v_s);
        if ((it + 1) % 50 == 0) {
            tau = 1.0f;
            sigma = 1.0f;
        } else {
            sigma /= mu;
            tau *= mu;
        }
    }
    HIPCHK(s, hipEventRecord(e1, s->stream));
    HIPCHK(s, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(s, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if ((rc = check_launch(s))) return rc;
    *ms_total = ms;
    if (launches_per_iter) *launches_per_iter = s->fused ? 1 : 2;
    return VFS_OK;
}

int vfs_time_copy(vfs_solver *s, size_t bytes, int reps, float *ms_total)
{
    if (!s || !ms_total || reps < 1 || bytes < 16) return VFS_ERR_ARG;
    int rc = set_device(s);
    if (rc) return rc;
    void *src = nullptr, *dst = nullptr;
    HIPCHK(s, hipMalloc(&src, bytes));
    hipError_t e = hipMalloc(&dst, bytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return fail(s, VFS_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    vfs::launch_fill_zero(src, bytes, s->stream);
    hipEvent_t e0, e1;
    HIPCHK(s, hipEventCreate(&e0));
    HIPCHK(s, hipEventCreate(&e1));
    HIPCHK(s, hipEventRecord(e0, s->stream));
    for (int i = 0; i < reps; i++) vfs::launch_copy(src, dst, bytes, s->stream);
    HIPCHK(s, hipEventRecord(e1, s->stream));
    HIPCHK(s, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(s, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if ((rc = check_launch(s))) return rc;
    *ms_total = ms;
    return VFS_OK;
}

}  // extern "C"
