// Host side of libvfs: the StereoTgv solver object behind the C ABI of include/vfs.h.
//
// Buffer model.  The reference allocates every work plane at level-0 size and reuses it
// at each pyramid level with that level's stride, swapping pointers between ping-pong
// partners on the host (stereotgv.cu:44-168, 389-515).  The same model is kept here so
// that values left outside the mask by earlier iterations and coarser levels are read
// exactly where the reference reads them.  Each solver's work planes live in one arena
// (slot layout in vfs_engine.h) that is zeroed by a fill kernel at the start of every
// solve: each solve equals the reference's first solve after initialize() on zeroed
// memory.
//
// Execution.  The solve schedule is open-loop (fixed iteration counts, host-computed
// step sizes, stereotgv.cu:449-477), so a solve is issued once under HIP stream capture
// and replayed as one graph launch.  The ping-pong roles are bits passed to the kernels;
// they evolve identically for every pair, so a batch of solvers (vfs_batch) runs the
// same schedule with every launch covering all of its pairs.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/vfs.h"
#include "vfs_engine.h"

using vfs::Dims;
using vfs::kMaxLevels;

namespace {

thread_local std::string g_create_error;

// Device-side launch set: descriptors + tile tables + captured graphs.
struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    vfs::PairDesc *d_desc = nullptr;
    int *d_tiles = nullptr;
    int npairs = 0;
    int tiles_total[kMaxLevels] = {};
    int max_plane = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    float graph_limit[2] = {0, 0};
    int graph_roles[2] = {0, 0};
};

struct Params {
    int nLevels, nWarpIters, nSolverIters, blockWidth, storage;
    float beta, gamma, alpha0, alpha1, timestep_lambda, lambda, fScale, limitRange;
};

}  // namespace

struct vfs_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // recorded after every solve of this pair (own or batched)
    int blockWidth = 32, blockHeight = 1, strideAlign = 32;
    bool initialized = false, has_mask = false, has_fields = false, has_images = false;
    std::string err;

    int width = 0, height = 0, stride = 0, nLevels = 0, nWarpIters = 0, nSolverIters = 0;
    float beta = 0, gamma = 0, alpha0 = 0, alpha1 = 0, timestep_lambda = 0, lambda = 0, fScale = 0;
    float limitRange = 1.0f;
    int storage = vfs::kF32;
    bool use_graph = true;
    Dims lv[kMaxLevels] = {};

    // persistent planes
    float *in_i0 = nullptr, *in_i1 = nullptr, *i0p0 = nullptr;
    void *I0p[kMaxLevels] = {}, *I1p[kMaxLevels] = {};
    float *mask[kMaxLevels] = {};
    float2 *tv[kMaxLevels] = {};
    float2 *cv = nullptr;
    uint8_t *staging = nullptr;
    size_t staging_bytes = 0;
    float *outX = nullptr, *outDepth = nullptr;
    bool raw_i1_fresh = false;  // in_i1 holds a frame not yet calibrated

    // work arena
    float *arena = nullptr;
    long long plane = 0;
    int roles = 0;  // ping-pong roles after the last solve

    Engine eng;  // single-pair launch set
    bool desc_dirty = true;
    std::vector<void *> allocs;

    Params params() const
    {
        return Params{nLevels, nWarpIters, nSolverIters, blockWidth, storage, beta, gamma, alpha0, alpha1,
                      timestep_lambda, lambda, fScale, limitRange};
    }
    vfs::PairDesc desc() const
    {
        vfs::PairDesc d{};
        d.arena = arena;
        d.plane = plane;
        d.in_i0 = in_i0;
        d.in_i1 = in_i1;
        d.cv = cv;
        for (int l = 0; l < nLevels; l++) {
            d.w[l] = lv[l].w;
            d.h[l] = lv[l].h;
            d.s[l] = lv[l].s;
            d.I0[l] = I0p[l];
            d.I1[l] = I1p[l];
            d.mask[l] = mask[l];
            d.tv[l] = tv[l];
        }
        d.I0[0] = storage == vfs::kF32 ? (void *)in_i0 : (void *)i0p0;
        return d;
    }
};

struct vfs_batch {
    Engine eng;
    std::vector<vfs_solver *> members;
    std::string err;
};

namespace {

int fail_str(std::string *err, int code, const char *fmt, va_list ap)
{
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, ap);
    if (err) *err = buf;
    else g_create_error = buf;
    return code;
}

int fail(vfs_solver *s, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(vfs_solver *s, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = fail_str(s ? &s->err : nullptr, code, fmt, ap);
    va_end(ap);
    return r;
}

int failb(vfs_batch *b, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int failb(vfs_batch *b, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = fail_str(b ? &b->err : nullptr, code, fmt, ap);
    va_end(ap);
    return r;
}

#define HIPCHK(s, expr)                                                                                      \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return fail((s), VFS_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

#define HIPCHKB(b, expr)                                                                                     \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return failb((b), VFS_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));     \
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

void engine_drop_graphs(Engine &e)
{
    for (int m = 0; m < 2; m++) {
        if (e.graph[m]) (void)hipGraphExecDestroy(e.graph[m]);
        e.graph[m] = nullptr;
    }
}

void engine_release(Engine &e)
{
    engine_drop_graphs(e);
    if (e.d_desc) (void)hipFree(e.d_desc);
    if (e.d_tiles) (void)hipFree(e.d_tiles);
    e.d_desc = nullptr;
    e.d_tiles = nullptr;
    e.npairs = 0;
}

// Uploads descriptors and tile prefix tables for `descs` (all with nLevels levels).
hipError_t engine_setup(Engine &e, const std::vector<vfs::PairDesc> &descs, int nLevels)
{
    engine_release(e);
    const int n = (int)descs.size();
    std::vector<int> tiles((size_t)kMaxLevels * (n + 1), 0);
    e.max_plane = 0;
    for (int l = 0; l < kMaxLevels; l++) e.tiles_total[l] = 0;
    for (int l = 0; l < nLevels; l++) {
        int acc = 0;
        for (int i = 0; i < n; i++) {
            tiles[(size_t)l * (n + 1) + i] = acc;
            const int tx = (descs[i].w[l] + vfs::kTileW - 1) / vfs::kTileW;
            const int ty = (descs[i].h[l] + vfs::kTileH - 1) / vfs::kTileH;
            acc += tx * ty;
        }
        tiles[(size_t)l * (n + 1) + n] = acc;
        e.tiles_total[l] = acc;
    }
    for (const auto &d : descs) e.max_plane = std::max(e.max_plane, (int)d.plane);
    hipError_t r = hipMalloc(&e.d_desc, sizeof(vfs::PairDesc) * n);
    if (r != hipSuccess) return r;
    r = hipMalloc(&e.d_tiles, sizeof(int) * tiles.size());
    if (r != hipSuccess) return r;
    r = hipMemcpyAsync(e.d_desc, descs.data(), sizeof(vfs::PairDesc) * n, hipMemcpyHostToDevice, e.stream);
    if (r != hipSuccess) return r;
    r = hipMemcpyAsync(e.d_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, e.stream);
    if (r != hipSuccess) return r;
    e.npairs = n;
    return hipStreamSynchronize(e.stream);
}

vfs::Launch engine_launch(const Engine &e)
{
    vfs::Launch L;
    L.desc = e.d_desc;
    L.tiles = e.d_tiles;
    L.npairs = e.npairs;
    for (int l = 0; l < kMaxLevels; l++) L.tiles_total[l] = e.tiles_total[l];
    L.max_plane_floats = e.max_plane;
    L.stream = e.stream;
    return L;
}

// Issues one solve of every pair of the launch set (stereotgv.cu:389-515 when masked,
// :256-387 otherwise), starting from zeroed arenas and role bits 0.  Returns the roles
// at the end.
int issue_solve(const Engine &e, const Params &P, int masked)
{
    const vfs::Launch L = engine_launch(e);
    const int st = P.storage;
    vfs::launch_zero_arenas(L, e.stream);
    vfs::launch_calibrate(L, st);
    for (int l = 1; l < P.nLevels; l++) vfs::launch_downscale_images(L, l, st);
    int roles = 0;
    for (int l = P.nLevels - 1; l >= 0; l--) {
        if (l == P.nLevels - 1) vfs::launch_coarsest_init(L, l, roles);
        vfs::launch_tensor(L, l, masked, P.blockWidth, P.beta, P.gamma, st);
        for (int wi = 0; wi < P.nWarpIters; wi++) {
            vfs::launch_warp_prep(L, l, roles, st);
            vfs::launch_deriv(L, l, st);
            float tau = 1.0f;
            float sigma = 1.0f / tau;
            for (int it = 0; it < P.nSolverIters; it++) {
                float mu;
                if (sigma < 1000.0f) mu = 1.0f / sqrtf(1 + 0.7f * tau * P.timestep_lambda);
                else mu = 1;
                vfs::IterParams ip;
                ip.k1 = P.alpha1 * sigma / 3.0f;  // eta_p = 3 (stereotgv.cu:416)
                ip.k0 = P.alpha0 * sigma / 2.0f;  // eta_q = 2 (stereotgv.cu:417)
                ip.tau = tau;
                ip.mu = mu;
                ip.lambda = P.lambda;
                ip.alpha0 = P.alpha0;
                ip.alpha1 = P.alpha1;
                ip.first = (it == 0);
                ip.write_tp = (it == P.nSolverIters - 1);
                vfs::launch_iter(L, l, masked, roles, ip, st);
                roles ^= vfs::RB_U | vfs::RB_UB | vfs::RB_V | vfs::RB_PQ;
                sigma = sigma / mu;
                tau = tau * mu;
            }
            vfs::launch_warp_end(L, l, masked, roles, P.limitRange);
            if (masked) roles ^= vfs::RB_U;
        }
        if (l > 0) {
            vfs::launch_upscale(L, l, roles, P.fScale);
            roles ^= vfs::RB_U | vfs::RB_UB | vfs::RB_UV;
        }
    }
    return roles;
}

// Runs one solve on the engine's stream: replays (capturing first if needed) the graph
// for this (masked, limit) or launches directly.  Returns the final roles through *roles.
int engine_solve(Engine &e, const Params &P, int masked, bool use_graph, std::string *err, int *roles)
{
    const int m = masked ? 1 : 0;
    auto ferr = [&](const char *what, hipError_t r) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(r));
        *err = buf;
        return VFS_ERR_DEVICE;
    };
    if (!use_graph) {
        *roles = issue_solve(e, P, masked);
        hipError_t r = hipGetLastError();
        return r == hipSuccess ? VFS_OK : ferr("kernel launch", r);
    }
    if (e.graph[m] && e.graph_limit[m] != P.limitRange) {
        (void)hipGraphExecDestroy(e.graph[m]);
        e.graph[m] = nullptr;
    }
    if (!e.graph[m]) {
        hipError_t r = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal);
        if (r != hipSuccess) return ferr("hipStreamBeginCapture", r);
        const int fr = issue_solve(e, P, masked);
        hipGraph_t g = nullptr;
        r = hipStreamEndCapture(e.stream, &g);
        if (r != hipSuccess) return ferr("hipStreamEndCapture", r);
        r = hipGraphInstantiate(&e.graph[m], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (r != hipSuccess) {
            e.graph[m] = nullptr;
            return ferr("hipGraphInstantiate", r);
        }
        e.graph_limit[m] = P.limitRange;
        e.graph_roles[m] = fr;
    }
    hipError_t r = hipGraphLaunch(e.graph[m], e.stream);
    if (r != hipSuccess) return ferr("hipGraphLaunch", r);
    *roles = e.graph_roles[m];
    return VFS_OK;
}

void release(vfs_solver *s)
{
    engine_release(s->eng);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    for (int l = 0; l < kMaxLevels; l++) {
        s->I0p[l] = s->I1p[l] = nullptr;
        s->mask[l] = nullptr;
        s->tv[l] = nullptr;
    }
    s->in_i0 = s->in_i1 = s->i0p0 = nullptr;
    s->cv = nullptr;
    s->staging = nullptr;
    s->staging_bytes = 0;
    s->arena = nullptr;
    s->initialized = s->has_mask = s->has_fields = s->has_images = false;
    s->desc_dirty = true;
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

int ready_to_solve(vfs_solver *s, int masked, std::string *err)
{
    const char *why = nullptr;
    if (!s->initialized) why = "solve before initialize";
    else if (!s->has_fields) why = "solve before loadVectorFields";
    else if (masked && !s->has_mask) why = "masked solve before copyMaskToDevice";
    else if (!s->has_images) why = "solve before copyImagesToDevice";
    if (why) {
        *err = why;
        return VFS_ERR_STATE;
    }
    return VFS_OK;
}

// A second solve without new images re-calibrates the already calibrated I1, as the
// reference's swap of pI1[0] with d_i1calibrated does (stereotgv.cu:392-393): the
// calibrated level-0 image becomes the raw input again.
hipError_t refresh_raw_i1(vfs_solver *s, hipStream_t st)
{
    if (s->raw_i1_fresh) return hipSuccess;
    const Dims d0 = s->lv[0];
    const size_t n = (size_t)d0.s * d0.h;
    if (s->storage == vfs::kF32)
        return hipMemcpyAsync(s->in_i1, s->I1p[0], sizeof(float) * n, hipMemcpyDeviceToDevice, st);
    vfs::launch_widen(s->I1p[0], s->storage, 1, s->in_i1, d0, st);
    return hipGetLastError();
}

int ensure_engine(vfs_solver *s)
{
    if (!s->desc_dirty) return VFS_OK;
    std::vector<vfs::PairDesc> d{s->desc()};
    hipError_t r = engine_setup(s->eng, d, s->nLevels);
    if (r != hipSuccess) return fail(s, VFS_ERR_DEVICE, "engine setup: %s", hipGetErrorString(r));
    s->desc_dirty = false;
    return VFS_OK;
}

int solve_common(vfs_solver *s, int masked)
{
    if (!s) return VFS_ERR_ARG;
    int rc = ready_to_solve(s, masked, &s->err);
    if (rc) return rc;
    if ((rc = set_device(s))) return rc;
    if ((rc = ensure_engine(s))) return rc;
    HIPCHK(s, refresh_raw_i1(s, s->stream));
    int roles = 0;
    rc = engine_solve(s->eng, s->params(), masked, s->use_graph, &s->err, &roles);
    if (rc) return rc;
    s->roles = roles;
    s->raw_i1_fresh = false;
    HIPCHK(s, hipEventRecord(s->done, s->stream));
    return VFS_OK;
}

// Waits (on the solver's stream) for the last solve of this pair, wherever it ran.
int wait_done(vfs_solver *s)
{
    HIPCHK(s, hipStreamWaitEvent(s->stream, s->done, 0));
    return VFS_OK;
}

int copy_out(vfs_solver *s, const void *dev, int nch, int level, float *out, int pitch_elems)
{
    const Dims d = s->lv[level];
    int rc = wait_done(s);
    if (rc) return rc;
    HIPCHK(s, hipMemcpy2DAsync(out, sizeof(float) * (size_t)pitch_elems, dev, sizeof(float) * nch * (size_t)d.s,
                               sizeof(float) * nch * (size_t)d.w, d.h, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return VFS_OK;
}

// Storage-typed plane -> float host array (through a temporary float plane).
int copy_out_st(vfs_solver *s, const void *dev, int nch, int level, float *out, int pitch_elems)
{
    if (s->storage == vfs::kF32) return copy_out(s, dev, nch, level, out, pitch_elems);
    const Dims d = s->lv[level];
    float *tmp = nullptr;
    HIPCHK(s, hipMalloc(&tmp, sizeof(float) * nch * (size_t)d.s * d.h));
    int rc = wait_done(s);
    if (!rc) {
        vfs::launch_widen(dev, s->storage, nch, tmp, d, s->stream);
        rc = check_launch(s);
    }
    if (!rc) rc = copy_out(s, tmp, nch, level, out, pitch_elems);
    (void)hipFree(tmp);
    return rc;
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
    int rc = wait_done(s);
    if (rc) return rc;
    HIPCHK(s, hipMemcpy2DAsync(s->staging, row, i0, pitch, row, s->height, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpy2DAsync(s->staging + row * s->height, row, i1, pitch, row, s->height, hipMemcpyHostToDevice,
                               s->stream));
    const Dims d0 = s->lv[0];
    if (nch == 1) {
        vfs::launch_cv8u(s->staging, (int)row, s->in_i0, d0, s->stream);
        vfs::launch_cv8u(s->staging + row * s->height, (int)row, s->in_i1, d0, s->stream);
    } else {
        vfs::launch_cv8uc3(s->staging, (int)row, s->in_i0, d0, s->stream);
        vfs::launch_cv8uc3(s->staging + row * s->height, (int)row, s->in_i1, d0, s->stream);
    }
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_images = true;
    s->raw_i1_fresh = true;
    return VFS_OK;
}

int time_events(vfs_solver *s, hipStream_t st, hipEvent_t *e0, hipEvent_t *e1)
{
    HIPCHK(s, hipEventCreate(e0));
    HIPCHK(s, hipEventCreate(e1));
    HIPCHK(s, hipEventRecord(*e0, st));
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
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(s->done, s->stream);
    if (e != hipSuccess) {
        fail(nullptr, VFS_ERR_DEVICE, "no usable HIP device: %s", hipGetErrorString(e));
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return VFS_ERR_DEVICE;
    }
    s->eng.stream = s->stream;
    s->eng.device = s->device;
    *out = s;
    return VFS_OK;
}

int vfs_create(vfs_solver **out) { return vfs_create_ex(out, 32, 1, 32); }

int vfs_destroy(vfs_solver *s)
{
    if (!s) return VFS_ERR_ARG;
    (void)hipSetDevice(s->device);
    release(s);
    if (s->done) (void)hipEventDestroy(s->done);
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
    if (nLevels < 1 || nLevels > kMaxLevels)
        return fail(s, VFS_ERR_ARG, "nLevels %d outside [1,%d]", nLevels, kMaxLevels);
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
        if ((rc = dev_alloc(s, &s->I0p[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, &s->I1p[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, (void **)&s->mask[l], n * sizeof(float)))) return rc;
        if ((rc = dev_alloc(s, (void **)&s->tv[l], n * sizeof(float2)))) return rc;
    }
    s->i0p0 = (float *)s->I0p[0];
    if ((rc = dev_alloc(s, (void **)&s->in_i0, N * sizeof(float)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->in_i1, N * sizeof(float)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->cv, N * sizeof(float2)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->outX, 3 * N * sizeof(float)))) return rc;
    if ((rc = dev_alloc(s, (void **)&s->outDepth, N * sizeof(float)))) return rc;
    // work arena: PL_COUNT slots of >= N floats, each 256 B aligned
    s->plane = (long long)((N + 63) & ~(size_t)63);
    if ((rc = dev_alloc(s, (void **)&s->arena, sizeof(float) * (size_t)vfs::PL_COUNT * s->plane))) return rc;
    s->roles = 0;
    s->desc_dirty = true;
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

int vfs_copy_mask(vfs_solver *s, const float *m, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copyMaskToDevice before initialize");
    if (!m || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad mask pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = wait_done(s))) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->mask[0], sizeof(float) * (size_t)d0.s, m, sizeof(float) * (size_t)pitch,
                               sizeof(float) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    for (int l = 1; l < s->nLevels; l++)
        vfs::launch_downscale_nn(s->mask[l - 1], s->lv[l - 1], s->mask[l], s->lv[l], s->stream);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_mask = true;
    return VFS_OK;
}

int vfs_load_vector_fields(vfs_solver *s, const float *tvh, int tv_pitch, const float *cvh, int cv_pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "loadVectorFields before initialize");
    if (!tvh || !cvh || tv_pitch < 2 * s->width || cv_pitch < 2 * s->width)
        return fail(s, VFS_ERR_ARG, "bad vector field pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = wait_done(s))) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->tv[0], sizeof(float2) * (size_t)d0.s, tvh, sizeof(float) * (size_t)tv_pitch,
                               sizeof(float2) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpy2DAsync(s->cv, sizeof(float2) * (size_t)d0.s, cvh, sizeof(float) * (size_t)cv_pitch,
                               sizeof(float2) * (size_t)d0.w, d0.h, hipMemcpyHostToDevice, s->stream));
    for (int l = 1; l < s->nLevels; l++)
        vfs::launch_downscale_f2(s->tv[l - 1], s->lv[l - 1], s->tv[l], s->lv[l], s->stream);
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
    if ((rc = wait_done(s))) return rc;
    const Dims d0 = s->lv[0];
    HIPCHK(s, hipMemcpy2DAsync(s->in_i0, sizeof(float) * (size_t)d0.s, i0, sizeof(float) * (size_t)pitch,
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

static float *slot_ptr(vfs_solver *s, int k) { return s->arena + (long long)k * s->plane; }

int vfs_copy_disparity_1d(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out(s, slot_ptr(s, (s->roles & vfs::RB_U) ? vfs::PL_U1 : vfs::PL_U0), 1, 0, out, pitch);
}

int vfs_copy_flow(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < 2 * s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out(s, slot_ptr(s, (s->roles & vfs::RB_UV) ? vfs::PL_UV1 : vfs::PL_UV0), 2, 0, out, pitch);
}

int vfs_copy_warped_image(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    return copy_out_st(s, slot_ptr(s, vfs::PL_I1W), 1, 0, out, pitch);
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
    if ((rc = wait_done(s))) return rc;
    const Dims d0 = s->lv[0];
    const float2 *uv = (const float2 *)slot_ptr(s, (s->roles & vfs::RB_UV) ? vfs::PL_UV1 : vfs::PL_UV0);
    vfs::launch_convert_kb(uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3, s->outX, s->outDepth, d0,
                           s->stream);
    if ((rc = check_launch(s))) return rc;
    if (depth && (rc = copy_out(s, s->outDepth, 1, 0, depth, depth_pitch))) return rc;
    if (X && (rc = copy_out(s, s->outX, 3, 0, X, X_pitch))) return rc;
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
    const int r = s->roles;
    const std::string n(name);
    const void *src = nullptr;
    int nch = 1;
    bool typed = false;  // stored in the storage type
    using namespace vfs;
    if (n == "u") src = slot_ptr(s, (r & RB_U) ? PL_U1 : PL_U0);
    else if (n == "u_") src = slot_ptr(s, (r & RB_UB) ? PL_UB1 : PL_UB0);
    else if (n == "u_last") src = slot_ptr(s, PL_ULAST);
    else if (n == "a") { src = slot_ptr(s, PL_A); typed = true; }
    else if (n == "b") { src = slot_ptr(s, PL_B); typed = true; }
    else if (n == "c") { src = slot_ptr(s, PL_C); typed = true; }
    else if (n == "Iu") { src = slot_ptr(s, PL_IU); typed = true; }
    else if (n == "Iz") { src = slot_ptr(s, PL_IZ); typed = true; }
    else if (n == "i1warp") { src = slot_ptr(s, PL_I1W); typed = true; }
    else if (n == "I0") { src = s->desc().I0[level]; typed = true; }
    else if (n == "I1") { src = s->I1p[level]; typed = true; }
    else if (n == "mask") src = s->mask[level];
    else if (n == "warpUV") { src = slot_ptr(s, (r & RB_UV) ? PL_UV1 : PL_UV0); nch = 2; }
    else if (n == "tv2") { src = slot_ptr(s, PL_TV2); nch = 2; }
    else if (n == "tv") { src = s->tv[level]; nch = 2; }
    else if (n == "p") { src = slot_ptr(s, (r & RB_PQ) ? PL_P1 : PL_P0); nch = 2; typed = true; }
    else if (n == "v_") { src = slot_ptr(s, (r & RB_V) ? PL_V1 : PL_V0); nch = 2; }
    else if (n == "Tp") { src = slot_ptr(s, PL_TP); nch = 2; }
    else if (n == "q") { src = slot_ptr(s, (r & RB_PQ) ? PL_Q1 : PL_Q0); nch = 4; typed = true; }
    if (!src) return fail(s, VFS_ERR_ARG, "unknown plane '%s'", name);
    int rc = set_device(s);
    if (rc) return rc;
    rc = typed ? copy_out_st(s, src, nch, level, out, s->lv[level].w * nch)
               : copy_out(s, src, nch, level, out, s->lv[level].w * nch);
    return rc ? rc : nch;
}

int vfs_synchronize(vfs_solver *s)
{
    if (!s) return VFS_ERR_ARG;
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = wait_done(s))) return rc;
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
    if (!strcmp(key, "storage")) {
        if (value < vfs::kF32 || value > vfs::kBF16) return fail(s, VFS_ERR_ARG, "storage must be 0, 1 or 2");
        if (value != s->storage) {
            s->storage = value;
            s->desc_dirty = true;
        }
        return VFS_OK;
    }
    return fail(s, VFS_ERR_ARG, "unknown option '%s'", key);
}

int vfs_get_option(const vfs_solver *s, const char *key, int *value)
{
    if (!s || !key || !value) return VFS_ERR_ARG;
    if (!strcmp(key, "graph")) *value = s->use_graph ? 1 : 0;
    else if (!strcmp(key, "storage")) *value = s->storage;
    else return VFS_ERR_ARG;
    return VFS_OK;
}

int vfs_time_solves(vfs_solver *s, int masked, int n, float *ms_total)
{
    if (!s || !ms_total || n < 1) return VFS_ERR_ARG;
    int rc = set_device(s);
    if (rc) return rc;
    hipEvent_t e0, e1;
    if ((rc = time_events(s, s->stream, &e0, &e1))) return rc;
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
    if ((rc = ensure_engine(s))) return rc;
    if ((rc = wait_done(s))) return rc;
    const vfs::Launch L = engine_launch(s->eng);
    hipEvent_t e0, e1;
    if ((rc = time_events(s, s->stream, &e0, &e1))) return rc;
    int roles = 0;
    float tau = 1.0f, sigma = 1.0f;
    for (int it = 0; it < reps; it++) {
        const float mu = 1.0f / sqrtf(1 + 0.7f * tau * s->timestep_lambda);
        vfs::IterParams ip{s->alpha1 * sigma / 3.0f, s->alpha0 * sigma / 2.0f, tau, mu, s->lambda, s->alpha0,
                           s->alpha1, 0, 0};
        vfs::launch_iter(L, level, 1, roles, ip, s->storage);
        roles ^= vfs::RB_U | vfs::RB_UB | vfs::RB_V | vfs::RB_PQ;
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
    HIPCHK(s, hipEventRecord(s->done, s->stream));
    *ms_total = ms;
    if (launches_per_iter) *launches_per_iter = 1;
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
    if ((rc = time_events(s, s->stream, &e0, &e1))) return rc;
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

// ------------------------------------------------------------------ batches

int vfs_batch_create(vfs_batch **out, vfs_solver *const *solvers, int n)
{
    if (!out) return failb(nullptr, VFS_ERR_ARG, "null handle pointer");
    *out = nullptr;
    if (!solvers || n < 1) return failb(nullptr, VFS_ERR_ARG, "empty batch");
    const vfs_solver *f = solvers[0];
    if (!f || !f->initialized) return failb(nullptr, VFS_ERR_STATE, "batch member 0 not initialized");
    for (int i = 0; i < n; i++) {
        const vfs_solver *s = solvers[i];
        if (!s || !s->initialized) return failb(nullptr, VFS_ERR_STATE, "batch member %d not initialized", i);
        if (s->device != f->device) return failb(nullptr, VFS_ERR_ARG, "batch members on different devices");
        if (s->nLevels != f->nLevels || s->nWarpIters != f->nWarpIters || s->nSolverIters != f->nSolverIters ||
            s->beta != f->beta || s->gamma != f->gamma || s->alpha0 != f->alpha0 || s->alpha1 != f->alpha1 ||
            s->timestep_lambda != f->timestep_lambda || s->lambda != f->lambda || s->fScale != f->fScale ||
            s->blockWidth != f->blockWidth)
            return failb(nullptr, VFS_ERR_ARG, "batch member %d has a different schedule or parameters", i);
        for (int j = 0; j < i; j++)
            if (solvers[j] == s) return failb(nullptr, VFS_ERR_ARG, "solver listed twice in a batch");
    }
    vfs_batch *b = new (std::nothrow) vfs_batch();
    if (!b) return failb(nullptr, VFS_ERR_NOMEM, "host allocation failed");
    b->members.assign(solvers, solvers + n);
    b->eng.device = f->device;
    hipError_t e = hipSetDevice(f->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->eng.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        failb(nullptr, VFS_ERR_DEVICE, "stream: %s", hipGetErrorString(e));
        delete b;
        return VFS_ERR_DEVICE;
    }
    *out = b;
    return VFS_OK;
}

int vfs_batch_destroy(vfs_batch *b)
{
    if (!b) return VFS_ERR_ARG;
    (void)hipSetDevice(b->eng.device);
    if (b->eng.stream) (void)hipStreamSynchronize(b->eng.stream);
    engine_release(b->eng);
    if (b->eng.stream) (void)hipStreamDestroy(b->eng.stream);
    delete b;
    return VFS_OK;
}

const char *vfs_batch_last_error(const vfs_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int vfs_batch_solve(vfs_batch *b, int masked)
{
    if (!b) return VFS_ERR_ARG;
    HIPCHKB(b, hipSetDevice(b->eng.device));
    const vfs_solver *f = b->members[0];
    Params P = f->params();
    bool rebuild = b->eng.npairs != (int)b->members.size();
    for (size_t i = 0; i < b->members.size(); i++) {
        vfs_solver *s = b->members[i];
        std::string why;
        if (ready_to_solve(s, masked, &why)) return failb(b, VFS_ERR_STATE, "member %zu: %s", i, why.c_str());
        if (s->storage != f->storage || s->limitRange != f->limitRange)
            return failb(b, VFS_ERR_ARG, "member %zu differs in storage or limitRange", i);
        if (s->desc_dirty) rebuild = true;
    }
    if (rebuild) {
        std::vector<vfs::PairDesc> descs;
        for (vfs_solver *s : b->members) {
            descs.push_back(s->desc());
            // a member's own single-pair launch set is rebuilt on its next own solve
        }
        hipError_t r = engine_setup(b->eng, descs, P.nLevels);
        if (r != hipSuccess) return failb(b, VFS_ERR_DEVICE, "engine setup: %s", hipGetErrorString(r));
    }
    // order after each member's pending work, then refresh re-used raw inputs
    for (vfs_solver *s : b->members) {
        HIPCHKB(b, hipStreamWaitEvent(b->eng.stream, s->done, 0));
        HIPCHKB(b, refresh_raw_i1(s, b->eng.stream));
    }
    int roles = 0;
    int rc = engine_solve(b->eng, P, masked, f->use_graph, &b->err, &roles);
    if (rc) return rc;
    for (vfs_solver *s : b->members) {
        s->roles = roles;
        s->raw_i1_fresh = false;
        HIPCHKB(b, hipEventRecord(s->done, b->eng.stream));
    }
    return VFS_OK;
}

int vfs_batch_synchronize(vfs_batch *b)
{
    if (!b) return VFS_ERR_ARG;
    HIPCHKB(b, hipSetDevice(b->eng.device));
    HIPCHKB(b, hipStreamSynchronize(b->eng.stream));
    return VFS_OK;
}

int vfs_batch_time_solves(vfs_batch *b, int masked, int n, float *ms_total)
{
    if (!b || !ms_total || n < 1) return VFS_ERR_ARG;
    HIPCHKB(b, hipSetDevice(b->eng.device));
    hipEvent_t e0, e1;
    HIPCHKB(b, hipEventCreate(&e0));
    HIPCHKB(b, hipEventCreate(&e1));
    HIPCHKB(b, hipEventRecord(e0, b->eng.stream));
    int rc = 0;
    for (int i = 0; i < n && !rc; i++) rc = vfs_batch_solve(b, masked);
    HIPCHKB(b, hipEventRecord(e1, b->eng.stream));
    HIPCHKB(b, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHKB(b, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms_total = ms;
    return VFS_OK;
}

}  // extern "C"
