// gfx950 device code of the StereoTgv TGV-L1 fisheye-stereo solver (included by the
// instantiation units tgv_inst.hip and by kernels.hip).
//
// Every per-pixel op of the solve runs inside one of two kernels over a *set of frame
// pairs* (see vfs_engine.h): the tick kernel k_tick (one op per launch, workgroups of 256
// work-items own tiles of the launch's shape, 64x8 by default) and k_iterk (two or four
// inner iterations per launch over a staged, recomputed halo).  A prefix table maps the
// flat workgroup index to (pair, tile), so pairs of different resolutions share a launch.
//
// Texture fetches of the reference (linear filter, mirror addressing, normalized
// coordinates) are evaluated in registers: exact bilinear weights at pixel-index
// position p = x_norm * dim - 0.5 over the symmetric extension of the image.
//
// Storage type T (float, __half, __hip_bfloat16) applies to the image pyramid, the
// warped image, the derivatives, the tensor and the duals p, q; all arithmetic is FP32.
#pragma once
#include "vfs_engine.h"

#include <math.h>

#include <type_traits>

// Every product and sum is rounded as written (no multiply-add fusion chosen by the
// compiler): a pixel's value is computed by the same rounding sequence in every tile
// shape and in every inlined copy of an op (tile pixel or recomputed halo), so results
// do not depend on the launch shape.
#pragma clang fp contract(off)

namespace vfs {

namespace {

constexpr int NT = 256;  // work-items per workgroup

// Work-item layout of tile shape SH: BX x BY work-items, RPT rows each.
template <int SH>
struct Sh {
    static constexpr int TW = kShape[SH].tw, TH = kShape[SH].th;
    static constexpr int BX = TW, BY = NT / TW, RPT = TH / BY;
    static_assert(BX * BY == NT && (RPT == 1 || RPT == 2) && BY * RPT == TH, "tile shape");
};

// ------------------------------------------------------------------ storage

template <typename T>
struct Cv;
template <>
struct Cv<float> {
    static __device__ __forceinline__ float f(float v) { return v; }
    static __device__ __forceinline__ float t(float v) { return v; }
};
template <>
struct Cv<__half> {
    static __device__ __forceinline__ float f(__half v) { return __half2float(v); }
    static __device__ __forceinline__ __half t(float v) { return __float2half(v); }
};
template <>
struct Cv<__hip_bfloat16> {
    static __device__ __forceinline__ float f(__hip_bfloat16 v) { return __bfloat162float(v); }
    static __device__ __forceinline__ __hip_bfloat16 t(float v) { return __float2bfloat16(v); }
};

template <typename T>
struct alignas(2 * sizeof(T)) V2 {
    T x, y;
};
template <typename T>
struct alignas(4 * sizeof(T)) V4 {
    T x, y, z, w;
};

template <typename T>
__device__ __forceinline__ float ldT(const T *__restrict__ p, long i)
{
    return Cv<T>::f(p[i]);
}
template <typename T>
__device__ __forceinline__ void stT(T *__restrict__ p, long i, float v)
{
    p[i] = Cv<T>::t(v);
}
template <typename T>
__device__ __forceinline__ float2 ld2T(const V2<T> *__restrict__ p, long i)
{
    const V2<T> v = p[i];
    return make_float2(Cv<T>::f(v.x), Cv<T>::f(v.y));
}
template <typename T>
__device__ __forceinline__ void st2T(V2<T> *__restrict__ p, long i, float2 v)
{
    V2<T> o;
    o.x = Cv<T>::t(v.x);
    o.y = Cv<T>::t(v.y);
    p[i] = o;
}
template <typename T>
__device__ __forceinline__ float4 ld4T(const V4<T> *__restrict__ p, long i)
{
    const V4<T> v = p[i];
    return make_float4(Cv<T>::f(v.x), Cv<T>::f(v.y), Cv<T>::f(v.z), Cv<T>::f(v.w));
}
template <typename T>
__device__ __forceinline__ void st4T(V4<T> *__restrict__ p, long i, float4 v)
{
    V4<T> o;
    o.x = Cv<T>::t(v.x);
    o.y = Cv<T>::t(v.y);
    o.z = Cv<T>::t(v.z);
    o.w = Cv<T>::t(v.w);
    p[i] = o;
}

// ------------------------------------------------------------------ sampling

__device__ __forceinline__ int mir(int k, int n)
{
    if ((unsigned)k < (unsigned)n) return k;
    if (k < 0) k = -1 - k;
    if (k >= n) {
        const int P = 2 * n;
        k %= P;
        if (k >= n) k = P - 1 - k;
    }
    return k;
}

__device__ __forceinline__ float clampc(float v) { return fminf(fmaxf(v, -1e8f), 1e8f); }

// Bilinear fetch with mirror addressing at pixel-index position (px, py).
template <typename T>
__device__ __forceinline__ float bil(const T *__restrict__ img, int w, int h, int s, float px, float py)
{
    if (__builtin_isnan(px) || __builtin_isnan(py)) return __builtin_nanf("");
    px = clampc(px);
    py = clampc(py);
    const float fx = floorf(px), fy = floorf(py);
    const float ax = px - fx, ay = py - fy;
    const int i = (int)fx, j = (int)fy;
    const int i0 = mir(i, w), i1 = mir(i + 1, w);
    const int j0 = mir(j, h), j1 = mir(j + 1, h);
    const T *r0 = img + (size_t)j0 * s;
    const T *r1 = img + (size_t)j1 * s;
    const float top = (1.f - ax) * Cv<T>::f(r0[i0]) + ax * Cv<T>::f(r0[i1]);
    const float bot = (1.f - ax) * Cv<T>::f(r1[i0]) + ax * Cv<T>::f(r1[i1]);
    return (1.f - ay) * top + ay * bot;
}

__device__ __forceinline__ float2 bil2(const float2 *__restrict__ img, int w, int h, int s, float px, float py)
{
    if (__builtin_isnan(px) || __builtin_isnan(py)) {
        const float n = __builtin_nanf("");
        return make_float2(n, n);
    }
    px = clampc(px);
    py = clampc(py);
    const float fx = floorf(px), fy = floorf(py);
    const float ax = px - fx, ay = py - fy;
    const int i = (int)fx, j = (int)fy;
    const int i0 = mir(i, w), i1 = mir(i + 1, w);
    const int j0 = mir(j, h), j1 = mir(j + 1, h);
    const float2 *r0 = img + (size_t)j0 * s;
    const float2 *r1 = img + (size_t)j1 * s;
    const float2 a = r0[i0], b = r0[i1], c = r1[i0], d = r1[i1];
    float2 o;
    o.x = (1.f - ay) * ((1.f - ax) * a.x + ax * b.x) + ay * ((1.f - ax) * c.x + ax * d.x);
    o.y = (1.f - ay) * ((1.f - ax) * a.y + ax * b.y) + ay * ((1.f - ax) * c.y + ax * d.y);
    return o;
}

template <typename T>
__device__ __forceinline__ float texel(const T *__restrict__ img, int w, int h, int s, int x, int y)
{
    return Cv<T>::f(img[(size_t)mir(y, h) * s + mir(x, w)]);
}

// Correctly rounded single-precision square root (the reference's sqrtf, IEEE-rounded):
// the hardware square root is within one ulp of it; the rounding boundaries on either
// side of that result (midpoints to the neighbouring floats, 25 significant bits) are
// squared exactly in double and compared with x.
__device__ __forceinline__ float sqrt_rn(float x)
{
    const float r = __builtin_sqrtf(x);
    if (!(x > 0.f) || __builtin_isinf(x)) return r;  // 0, negative, infinite, NaN
    const double dx = (double)x;
    const float up = __int_as_float(__float_as_int(r) + 1);
    const double mu = 0.5 * ((double)r + (double)up);
    if (mu * mu < dx) return up;
    const float dn = __int_as_float(__float_as_int(r) - 1);
    const double md = 0.5 * ((double)dn + (double)r);
    if (md * md > dx) return dn;
    return r;
}

// Circle gate of the unmasked solve (updateDualVariables.cu:13-22, thresholdingL1.cu:10-19,
// updatePrimalVariables.cu:14-23): true when sqrtf((y - h/2)^2 + (x - w/2)^2) >= w / 2.2,
// evaluated in float with the correctly rounded square root, as the reference does.
__device__ __forceinline__ bool outside_circle(int x, int y, int w, int h)
{
    const float desired = (float)w / 2.20f;
    const float hw = (float)w / 2.0f, hh = (float)h / 2.0f;
    const float dy = (float)y - hh, dx = (float)x - hw;
    return sqrt_rn(dy * dy + dx * dx) >= desired;
}

__device__ __forceinline__ float2 ld0(const float2 *__restrict__ b, long i)
{
    return i < 0 ? make_float2(0.f, 0.f) : b[i];
}

// ------------------------------------------------------------------ tiling

struct Tile {
    const PairDesc *d;
    float *arena;        // the pair's work arena
    long long plane;     // floats per plane slot
    const float *mask;   // mask plane of the tile level
    int x0, y0, w, h, s;
};

__device__ __forceinline__ float *slot(const Tile &t, int k) { return t.arena + (long long)k * t.plane; }

template <typename P>
__device__ __forceinline__ P *slotp(const Tile &t, int k)
{
    return reinterpret_cast<P *>(t.arena + (long long)k * t.plane);
}

// Iterates the calling work-item over its pixels of the tile (S::RPT rows).
#define FOR_TILE_PIXELS(S, t, x, y)                                                   \
    for (int y = (t).y0 + (int)threadIdx.y, x = (t).x0 + (int)threadIdx.x;            \
         y < (t).y0 + S::TH && y < (t).h && x < (t).w; y += S::BY)

// ------------------------------------------------------------------ solve ops
//
// Each op works on one tile (shape S) of one pair; k_tick (below) dispatches them.

// Fresh state: zero every work plane of the pair at the tile's level-0 positions
// (the tile grid spans the stride, so row padding is cleared too).
template <class S>
__device__ __forceinline__ void op_zero(const Tile &t)
{
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
#pragma unroll 5
        for (int k = 0; k < PL_COUNT; k++) slot(t, k)[pos] = 0.f;
    }
}

// Before the multi-iteration launches of a warp at a coarse level of a warm solve: the
// spare buffers of u, u_ and v_ take the current buffers' values beyond the level's flat
// extent (stride x height), where level-0 data of the previous solve lies and every
// buffer differs.  A multi-iteration launch makes the spare buffers current there
// without writing them; with this copy the current buffers keep their values, as the
// reference's buffers do beyond the level's extent.  The tile covers the level-0 planes.
template <class S>
__device__ __forceinline__ void op_sync(const Tile &t, int level, int roles)
{
    const long ext = (long)t.d->s[level] * t.d->h[level];
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *us = slot(t, slot_u(rspare(roles, G_U)));
    const float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float *u_s = slot(t, slot_ub(rspare(roles, G_UB)));
    const float2 *v = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    float2 *vs = slotp<float2>(t, slot_v(rspare(roles, G_V)));
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        if (pos >= ext) {
            us[pos] = u[pos];
            u_s[pos] = u_[pos];
            vs[pos] = v[pos];
        }
    }
}

// OP_SYNC's copy for one chunk of kSyncChunk flat positions beyond the level's extent
// (the extra workgroups of a derivatives op with EF_SYNC): chunk k covers positions
// ext + k kSyncChunk .. of the level-0 planes.
template <class S>
__device__ __forceinline__ void op_sync_chunk(const Tile &t, int level, int roles, int chunk)
{
    const long ext = (long)t.d->s[level] * t.d->h[level];
    const long end = (long)t.d->s[0] * t.d->h[0];
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *us = slot(t, slot_u(rspare(roles, G_U)));
    const float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float *u_s = slot(t, slot_ub(rspare(roles, G_UB)));
    const float2 *v = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    float2 *vs = slotp<float2>(t, slot_v(rspare(roles, G_V)));
    const int tid = threadIdx.y * S::BX + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSyncChunk / NT; k++) {
        const long pos = ext + (long)chunk * kSyncChunk + k * NT + tid;
        if (pos < end) {
            us[pos] = u[pos];
            u_s[pos] = u_[pos];
            vs[pos] = v[pos];
        }
    }
}

// Calibration warp of I1 (stereotgv.cu:392-393, warpImage.cu:9-23) into the level-0
// pyramid slot; with 16-bit storage also narrows I0 into its pyramid slot.
template <class S, typename T>
__device__ __forceinline__ void op_calib(const Tile &t)
{
    const float *i1 = t.d->in_i1;
    T *o1 = (T *)t.d->I1[0];
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        const float2 c = t.d->cv[pos];
        stT(o1, pos, bil(i1, t.w, t.h, t.s, (float)x + c.x, (float)y + c.y));
        if (sizeof(T) != sizeof(float)) stT((T *)t.d->I0[0], pos, t.d->in_i0[pos]);
    }
}

// downscale.cu:10-30 for I0 and I1 (stereotgv.cu:400-405): mean of 4 bilinear taps at
// +-1/4 coarse pixel around the coarse pixel centre, fetched from the finer level.
template <class S, typename T>
__device__ __forceinline__ void op_down(const Tile &t, int level)
{
    const int W = t.d->w[level - 1], H = t.d->h[level - 1], SF = t.d->s[level - 1];
    const T *a = (const T *)t.d->I0[level - 1];
    const T *b = (const T *)t.d->I1[level - 1];
    T *oa = (T *)t.d->I0[level];
    T *ob = (T *)t.d->I1[level];
    const float rx = (float)W / (float)t.w, ry = (float)H / (float)t.h;
    const float ox = 0.25f * rx, oy = 0.25f * ry;
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const float cx = ((float)x + 0.5f) * rx - 0.5f, cy = ((float)y + 0.5f) * ry - 0.5f;
        const long pos = (long)y * t.s + x;
        stT(oa, pos, 0.25f * (bil(a, W, H, SF, cx - ox, cy) + bil(a, W, H, SF, cx + ox, cy) +
                              bil(a, W, H, SF, cx, cy - oy) + bil(a, W, H, SF, cx, cy + oy)));
        stT(ob, pos, 0.25f * (bil(b, W, H, SF, cx - ox, cy) + bil(b, W, H, SF, cx + ox, cy) +
                              bil(b, W, H, SF, cx, cy - oy) + bil(b, W, H, SF, cx, cy + oy)));
    }
}

// Coarsest-level init warpUV = u * tv (stereotgv.cu:419-421, warpImage.cu:136-150).
template <class S>
__device__ __forceinline__ void op_cinit(const Tile &t, int level, int roles)
{
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    const float2 *tv = t.d->tv[level];
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        const float us = u[pos];
        const float2 v = tv[pos];
        uv[pos] = make_float2(us * v.x, us * v.y);
    }
}

// Warm start: u, u_ and warpUV of the finer level downscaled to `level` and multiplied
// by `scale` (= 1/fScale) into the other slot of each pair: the 4-tap Downscale with a
// scale factor of downscale.cu:148-237.
template <class S>
__device__ __forceinline__ void op_winit(const Tile &t, int level, int roles, float scale)
{
    const int W = t.d->w[level - 1], H = t.d->h[level - 1], SF = t.d->s[level - 1];
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *ou = slot(t, slot_u(roth(roles, G_U)));
    const float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float *ou_ = slot(t, slot_ub(roth(roles, G_UB)));
    const float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    float2 *ouv = slotp<float2>(t, slot_uv(roles, false));
    const float rx = (float)W / (float)t.w, ry = (float)H / (float)t.h;
    const float ox = 0.25f * rx, oy = 0.25f * ry;
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const float cx = ((float)x + 0.5f) * rx - 0.5f, cy = ((float)y + 0.5f) * ry - 0.5f;
        const long pos = (long)y * t.s + x;
        ou[pos] = scale * 0.25f * (bil(u, W, H, SF, cx - ox, cy) + bil(u, W, H, SF, cx + ox, cy) +
                                   bil(u, W, H, SF, cx, cy - oy) + bil(u, W, H, SF, cx, cy + oy));
        ou_[pos] = scale * 0.25f * (bil(u_, W, H, SF, cx - ox, cy) + bil(u_, W, H, SF, cx + ox, cy) +
                                    bil(u_, W, H, SF, cx, cy - oy) + bil(u_, W, H, SF, cx, cy + oy));
        const float2 v0 = bil2(uv, W, H, SF, cx - ox, cy), v1 = bil2(uv, W, H, SF, cx + ox, cy);
        const float2 v2 = bil2(uv, W, H, SF, cx, cy - oy), v3 = bil2(uv, W, H, SF, cx, cy + oy);
        ouv[pos] = make_float2(scale * 0.25f * (v0.x + v1.x + v2.x + v3.x), scale * 0.25f * (v0.y + v1.y + v2.y + v3.y));
    }
}

// 3x3 binomial of the level image at (x, y), mirror addressing (calcTensor.cu:245-273).
template <typename T>
__device__ __forceinline__ float gauss3(const T *__restrict__ I, int w, int h, int s, int x, int y)
{
    float r = 0.25f * texel(I, w, h, s, x, y);
    r += 0.0625f * texel(I, w, h, s, x - 1, y - 1);
    r += 0.0625f * texel(I, w, h, s, x - 1, y + 1);
    r += 0.0625f * texel(I, w, h, s, x + 1, y - 1);
    r += 0.0625f * texel(I, w, h, s, x + 1, y + 1);
    r += 0.125f * texel(I, w, h, s, x - 1, y);
    r += 0.125f * texel(I, w, h, s, x + 1, y);
    r += 0.125f * texel(I, w, h, s, x, y - 1);
    r += 0.125f * texel(I, w, h, s, x, y + 1);
    return r;
}

// computeDerivatives.cu:88-129: the 5-point derivative along t = tv/|tv| of image `I` at
// pixel (x, y) (one of the two halves Iu averages).
template <typename T>
__device__ __forceinline__ float deriv_along(const T *__restrict__ I, const float2 *__restrict__ tv, int w, int h,
                                             int s, int x, int y)
{
    const float2 v = tv[(long)y * s + x];
    const float r = sqrt_rn(v.x * v.x + v.y * v.y);
    const float dx = v.x / r, dy = v.y / r;
    const float X = (float)x, Y = (float)y;
    float t = bil(I, w, h, s, X - 2.f * dx, Y - 2.f * dy);
    t -= bil(I, w, h, s, X - dx, Y - dy) * 8.f;
    t += bil(I, w, h, s, X + dx, Y + dy) * 8.f;
    t -= bil(I, w, h, s, X + 2.f * dx, Y + 2.f * dy);
    return t / 12.f;
}

// The smoothed image as the tensor kernel samples it: the masked blur writes only
// mask != 0 pixels of a zeroed plane, and the texture fetch mirrors at the level edge.
template <typename T, bool MASKED>
__device__ __forceinline__ float smoothed(const T *__restrict__ I, const float *__restrict__ mask, int w, int h,
                                          int s, int x, int y)
{
    const int mx = mir(x, w), my = mir(y, h);
    if (MASKED && mask[(size_t)my * s + mx] == 0.f) return 0.f;
    return gauss3(I, w, h, s, mx, my);
}

// Fused GaussianMasked + CalcTensorMasked (calcTensor.cu:154-273) or the unmasked pair
// (calcTensor.cu:7-128).  Writes a, b, c for every in-image pixel: the tensor where the
// reference computes it, 0 where the reference leaves its per-level memset (mask == 0,
// or columns x >= xlimit: calcTensor.cu:81,227 size the launch by the HEIGHT).  The
// preconditioners eta (preconditioning.cu) are recomputed from (a, b, c) where needed.
template <class S, typename T, bool MASKED>
__device__ __forceinline__ void op_tensor(const Tile &t, int level, int xblock, float beta, float gamma)
{
    const T *I0 = (const T *)t.d->I0[level];
    const float *mask = t.mask;
    T *at = (T *)slot(t, PL_A);
    T *bt = (T *)slot(t, PL_B);
    T *ct = (T *)slot(t, PL_C);
    float *i0d = slot(t, PL_I0D);
    const float2 *tv = t.d->tv[level];
    const int xlimit = xblock * ((t.h + xblock - 1) / xblock);
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        // the I0 half of every warp's derivative at this level (op_deriv)
        i0d[pos] = deriv_along(I0, tv, t.w, t.h, t.s, x, y);
        float a = 0.f, b = 0.f, c = 0.f;
        const bool active = MASKED ? (mask[pos] != 0.f) : true;
        if (active && x < xlimit) {
            const float s11 = smoothed<T, MASKED>(I0, mask, t.w, t.h, t.s, x + 1, y + 1);
            const float s01 = smoothed<T, MASKED>(I0, mask, t.w, t.h, t.s, x, y + 1);
            const float s10 = smoothed<T, MASKED>(I0, mask, t.w, t.h, t.s, x + 1, y);
            const float gx = s11 - s01;
            const float gy = s11 - s10;
            const float absg = sqrt_rn(gx * gx + gy * gy);
            float nx = gx / absg, ny = gy / absg;
            if (absg < 1e-8f) {
                nx = 1.f;
                ny = 0.f;
            }
            const float tx = ny, ty = -nx;
            float wt = expf(-beta * powf(absg, gamma));
            if (!(wt > 1e-8f)) wt = 1e-8f;
            a = wt * nx * nx + tx * tx;
            c = wt * nx * ny + tx * ty;
            b = wt * ny * ny + ty * ty;
        }
        stT(at, pos, a);
        stT(bt, pos, b);
        stT(ct, pos, c);
    }
}

// Per-warp preparation (stereotgv.cu:431-447): tv2 = bilinear(tv, x + warpUV)
// (warpImage.cu:96-111), I1w = bilinear(I1, x + warpUV) (warpImage.cu:9-23),
// u_last = u (copyClone.cu:3-12), v_ = 0 (Clone of the zeroed v).
template <class S, typename T>
__device__ __forceinline__ void op_prep(const Tile &t, int level, int roles)
{
    const float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float2 *v_ = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    const float2 *tv = t.d->tv[level];
    const T *I1 = (const T *)t.d->I1[level];
    float2 *tv2 = slotp<float2>(t, PL_TV2);
    T *i1w = (T *)slot(t, PL_I1W);
    float *ulast = slot(t, PL_ULAST);
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        const float2 d = uv[pos];
        const float px = (float)x + d.x, py = (float)y + d.y;
        tv2[pos] = bil2(tv, t.w, t.h, t.s, px, py);
        stT(i1w, pos, bil(I1, t.w, t.h, t.s, px, py));
        ulast[pos] = u[pos];
        v_[pos] = make_float2(0.f, 0.f);
    }
}

// computeDerivatives.cu:88-129 at pixel (x, y): Iu = (I0's derivative along the trajectory,
// computed once per level by the tensor op into PL_I0D, + I1w's) / 2 and Iz = I1w - I0 at
// the texel centre.
template <typename T>
__device__ __forceinline__ float2 deriv_px(const T *__restrict__ I0, const T *__restrict__ I1w,
                                           const float *__restrict__ i0d, const float2 *__restrict__ tv, int w, int h,
                                           int s, int x, int y)
{
    const long pos = (long)y * s + x;
    const float t1 = deriv_along(I1w, tv, w, h, s, x, y);
    return make_float2((i0d[pos] + t1) * 0.5f, ldT(I1w, pos) - ldT(I0, pos));
}

template <class S, typename T>
__device__ __forceinline__ void op_deriv(const Tile &t, int level)
{
    const T *I0 = (const T *)t.d->I0[level];
    const T *I1w = (const T *)slot(t, PL_I1W);
    const float2 *tv = t.d->tv[level];
    T *Iu = (T *)slot(t, PL_IU);
    T *Iz = (T *)slot(t, PL_IZ);
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        const float2 d = deriv_px(I0, I1w, slot(t, PL_I0D), tv, t.w, t.h, t.s, x, y);
        stT(Iu, pos, d.x);
        stT(Iz, pos, d.y);
    }
}

// The warp's derivatives at staged position (x, y) (clamped into the level) for an
// inner-iteration op that carries EF_DERIV: computed as op_deriv does, rounded to the
// storage type as the planes hold them, and stored when the position is in the tile.
template <typename T>
__device__ __forceinline__ float2 staged_deriv(const Tile &t, int level, int x, int y, bool in_tile)
{
    const float2 d = deriv_px((const T *)t.d->I0[level], (const T *)slot(t, PL_I1W), slot(t, PL_I0D),
                              t.d->tv[level], t.w, t.h, t.s, x, y);
    const float iu = Cv<T>::f(Cv<T>::t(d.x)), iz = Cv<T>::f(Cv<T>::t(d.y));
    if (in_tile) {
        const long pos = (long)y * t.s + x;
        stT((T *)slot(t, PL_IU), pos, iu);
        stT((T *)slot(t, PL_IZ), pos, iz);
    }
    return make_float2(iu, iz);
}

// The inner iteration divides and normalises several times per pixel; it multiplies by
// the hardware reciprocal and reciprocal square root (about 1 ulp) instead of the
// correctly rounded expansions, which are several times longer.  NaN and infinity
// propagate as in IEEE division.

// The inner iteration (dual_step, thresh_core, primal_step) writes every product that is
// summed as an explicit fused multiply-add (fmaf: one rounding), as the reference's
// compiler fuses them; explicit, so that every copy of the code rounds alike.

// Per-pixel constants of the primal step (fixed within a warp: they depend on the
// tensor only): the reciprocals of eta_u (the threshold's step scale) and of the two
// eta_v, each 1 where the reference divides by zero only in its zero branch.  eta_v2 == 0
// (alpha0 == 0 and a = c = 0), the reference's other zero branch, is r2 = -0 (the
// reciprocal itself is never -0) with r1 = 1: kept in the float, so that no per-pixel
// flag has to live in scalar registers across the iterations.
struct PrimalCoef {
    float rt, r1, r2;
    float ri;  // 1 / Iu (0 for Iu = 0) of the threshold
};
__device__ __forceinline__ bool eta_v2_zero(const PrimalCoef &k) { return __float_as_uint(k.r2) == 0x80000000u; }

__device__ __forceinline__ PrimalCoef primal_coef(const IterParams &ip, float a, float b, float c, float iu)
{
    const float aa = ip.alpha1 * ip.alpha1;
    const float etau = fmaf(b + c, b + c, fmaf(a + c, a + c, fmaf(2 * c, c, fmaf(a, a, b * b)))) * aa;
    const float a04 = 4 * ip.alpha0 * ip.alpha0;
    const float e1 = fmaf(aa, fmaf(b, b, c * c), a04);
    const float e2 = fmaf(aa, fmaf(a, a, c * c), a04);
    PrimalCoef k;
    k.rt = etau == 0.f ? 1.0f : __builtin_amdgcn_rcpf(etau);
    const bool e2z = e2 == 0.f;
    k.r1 = (e1 == 0.f || e2z) ? 1.0f : __builtin_amdgcn_rcpf(e1);
    k.r2 = e2z ? -0.0f : __builtin_amdgcn_rcpf(e2);
    k.ri = iu == 0.f ? 0.f : __builtin_amdgcn_rcpf(iu);
    return k;
}

// L1 proximal step (thresholdingL1.cu:180-215), linearised at the current u; k.rt is
// 1 / eta_u (1 for eta_u = 0, where the reference takes tau itself).
__device__ __forceinline__ float thresh_core(float divp, float tau, const PrimalCoef &k, float u_pos, float upos,
                                             float Iu, float Iz, float lambda)
{
    const float te = tau * k.rt;
    const float uhat = fmaf(te, divp, u_pos);
    const float dun = uhat - upos;
    const float rho = fmaf(Iu, dun, Iz);
    const float lt = lambda * te;
    const float upper = lt * (Iu * Iu);
    const float lower = -upper;
    // The three cases (none for NaN) as selects between candidates that are all computed:
    // the same values as branches, without the divergent control flow per pixel.  For
    // Iu = 0 (k.ri = 0) the inside case is dun - rho * 0, dun itself up to the sign of a
    // zero.
    const bool inside = (rho <= upper) && (rho >= lower);
    const float d_in = fmaf(-rho, k.ri, dun);
    const float d_lo = fmaf(lt, Iu, dun);
    const float d_hi = fmaf(-lt, Iu, dun);
    const float d_out = (rho < lower) ? d_lo : ((rho > upper) ? d_hi : 0.f);
    return upos + (inside ? d_in : d_out);
}


// Operands of the dual update at one pixel, gathered with in-buffer (clamped) addresses
// and no data-dependent branch, so that all loads of a work-item issue back to back.
struct DualIn {
    long pos;
    bool ok;  // inside the level image
    bool act; // processed: in the mask (masked form) / in the image (unmasked form)
    bool hasR, hasD;
    float a, b, c, u0, uR, uD, uL, uU;
    float2 v0, vR, vD, vL, vU, po;
    float4 qo;
};

template <bool MASKED, typename T>
__device__ __forceinline__ void dual_load(DualIn &r, int x, int y, int w, int h, int s, const float *__restrict__ mask,
                                          const float *__restrict__ u_, const float2 *__restrict__ v_,
                                          const T *__restrict__ at, const T *__restrict__ bt, const T *__restrict__ ct,
                                          const V2<T> *__restrict__ p_in, const V4<T> *__restrict__ q_in, bool first)
{
    r.ok = x >= 0 && y >= 0 && x < w && y < h;
    const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
    const long pos = (long)yc * s + xc;
    r.pos = pos;
    const long pR = (xc + 1 < w) ? pos + 1 : pos;
    const long pD = (yc + 1 < h) ? pos + s : pos;
    const long pL = pos >= 1 ? pos - 1 : pos;
    const long pU = pos >= s ? pos - s : pos;
    if (MASKED) {
        const float m0 = mask[pos], mR = mask[pR], mD = mask[pD];
        r.act = r.ok && m0 != 0.f;
        r.hasR = (xc + 1 < w) && mR != 0.f;
        r.hasD = (yc + 1 < h) && mD != 0.f;
    } else {
        r.act = r.ok;
        r.hasR = xc + 1 < w;
        r.hasD = yc + 1 < h;
    }
    r.a = ldT(at, pos);
    r.b = ldT(bt, pos);
    r.c = ldT(ct, pos);
    r.u0 = u_[pos];
    r.uR = u_[pR];
    r.uD = u_[pD];
    r.uL = pos >= 1 ? u_[pL] : 0.f;   // flat index -1 reads 0
    r.uU = pos >= s ? u_[pU] : 0.f;
    r.v0 = v_[pos];
    r.vR = v_[pR];
    r.vD = v_[pD];
    r.vL = pos >= 1 ? v_[pL] : make_float2(0.f, 0.f);
    r.vU = pos >= s ? v_[pU] : make_float2(0.f, 0.f);
    if (first) {
        r.po = make_float2(0.f, 0.f);
        r.qo = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        r.po = ld2T(p_in, pos);
        r.qo = ld4T(q_in, pos);
    }
}

// T p with T = [[a, c], [c, b]] (thresholdingL1.cu:242-243): Tp of the dual step, and the v
// step's tensor term (updatePrimalVariables.cu:216-226), the same value.
__device__ __forceinline__ float2 tensor_times(float a, float b, float c, float2 p)
{
    return make_float2(fmaf(a, p.x, c * p.y), fmaf(c, p.x, b * p.y));
}

// updateDualVariables.cu:152-236 (masked) / :28-98 (unmasked): ascent on p and q with
// unit-ball projection, then Tp = T p (thresholdingL1.cu:242-243).  `gate_out`: unmasked
// form, the pixel lies outside the circle gate (p = q = 0 there).
template <bool MASKED>
__device__ __forceinline__ void dual_step(const DualIn &r, bool gate_out, const IterParams &ip, float2 &pn, float4 &qn,
                                          float2 &tpn)
{
    if (!MASKED && gate_out) {
        pn = make_float2(0.f, 0.f);
        qn = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        const float ux = r.hasR ? (r.uR - r.u0 - r.v0.x) : (r.u0 - r.uL - r.v0.x);
        const float uy = r.hasD ? (r.uD - r.u0 - r.v0.y) : (r.u0 - r.uU - r.v0.y);
        const float dtx = fmaf(r.a, ux, r.c * uy);
        const float dty = fmaf(r.c, ux, r.b * uy);
        const float px = fmaf(ip.k1, dtx, r.po.x);
        const float py = fmaf(ip.k1, dty, r.po.y);
        // projection onto the unit ball: p / max(|p|, 1), with the hardware reciprocal
        // square root
        const float sp = fmaf(px, px, py * py);
        const float irp = sp > 1.0f ? __builtin_amdgcn_rsqf(sp) : 1.0f;
        pn = make_float2(px * irp, py * irp);
        float g0, g1, g2, g3;
        if (r.hasR) {
            g0 = r.vR.x - r.v0.x;
            g3 = r.vR.y - r.v0.y;
        } else {
            g0 = r.v0.x - r.vL.x;
            g3 = r.v0.y - r.vL.y;
        }
        if (r.hasD) {
            g1 = r.vD.y - r.v0.y;
            g2 = r.vD.x - r.v0.x;
        } else {
            g1 = r.v0.y - r.vU.y;
            g2 = r.v0.x - r.vU.x;
        }
        const float q0 = fmaf(ip.k0, g0, r.qo.x), q1 = fmaf(ip.k0, g1, r.qo.y), q2 = fmaf(ip.k0, g2, r.qo.z),
                    q3 = fmaf(ip.k0, g3, r.qo.w);
        const float sq = fmaf(q0, q0, fmaf(q1, q1, fmaf(q2, q2, q3 * q3)));
        const float irq = sq > 1.0f ? __builtin_amdgcn_rsqf(sq) : 1.0f;
        qn = make_float4(q0 * irq, q1 * irq, q2 * irq, q3 * irq);
    }
    tpn = tensor_times(r.a, r.b, r.c, pn);
}

// Thresholding and primal step at one pixel (thresholdingL1.cu:176-216 / :60-86,
// updatePrimalVariables.cu:165-245 / :30-113) from the new duals: divp is the divergence
// of Tp, qc4 / tp the pixel's new q / Tp (T p of the new p: the v step's tensor term,
// computed once in the dual step), ql / qu the new q of its left / upper neighbours (read
// only where they are inside the mask).  Returns u, u_ and v_ of the next iteration.
__device__ __forceinline__ void primal_core(const IterParams &ip, const PrimalCoef &k, float u_pos, float uu,
                                            float iu, float iz, float2 vo, bool hR, bool hD, bool hL, bool hU,
                                            float divp, float4 qc4, float4 ql, float4 qu, float2 tp, float &un,
                                            float &u_s, float2 &v_s)
{
    un = thresh_core(divp, ip.tau, k, u_pos, uu, iu, iz, ip.lambda);
    float qcx = 0.f, qcw = 0.f, qcy = 0.f, qcz = 0.f;
    if (hR) {
        qcx = qc4.x;
        qcw = qc4.w;
    }
    if (hD) {
        qcy = qc4.y;
        qcz = qc4.z;
    }
    float qwx = 0.f, qww = 0.f, qny = 0.f, qnz = 0.f;
    if (hL) {
        qwx = ql.x;
        qww = ql.w;
    }
    if (hU) {
        qny = qu.y;
        qnz = qu.z;
    }
    const float dqx = (qcx - qwx) + (qcz - qnz);
    const float dqy = (qcw - qww) + (qcy - qny);
    const float gx = fmaf(ip.alpha1, tp.x, ip.alpha0 * dqx);
    const float gy = fmaf(ip.alpha1, tp.y, ip.alpha0 * dqy);
    // eta_v1, eta_v2 >= 4 alpha0^2 > 0 for alpha0 != 0; the reference's zero branches
    // (updatePrimalVariables.cu:228-239, the second writes .x: there vx takes tau itself,
    // k.r1 = 1) are kept for alpha0 == 0.
    const float vx = fmaf(ip.tau * k.r1, gx, vo.x);
    const float vy = eta_v2_zero(k) ? 0.f : fmaf(ip.tau * k.r2, gy, vo.y);
    u_s = fmaf(ip.mu, un - u_pos, un);
    v_s = make_float2(fmaf(ip.mu, vx - vo.x, vx), fmaf(ip.mu, vy - vo.y, vy));
}

__device__ __forceinline__ void primal_step(const IterParams &ip, float a, float b, float c, float u_pos, float uu,
                                            float iu, float iz, float2 vo, bool hR, bool hD, bool hL, bool hU,
                                            float divp, float4 qc4, float4 ql, float4 qu, float2 tp, float &un,
                                            float &u_s, float2 &v_s)
{
    primal_core(ip, primal_coef(ip, a, b, c, iu), u_pos, uu, iu, iz, vo, hR, hD, hL, hU, divp, qc4, ql, qu, tp, un,
                u_s, v_s);
}

// Divergence of Tp at a pixel (thresholdingL1.cu:157-175 masked / :29-45 unmasked): tp0
// the pixel's new Tp, tpLx / tpUy the left / upper neighbour's.
template <bool MASKED>
__device__ __forceinline__ float div_tp(int x, int y, int w, int h, bool hR, bool hD, bool hL, bool hU, float2 tp0,
                                        float tpLx, float tpUy)
{
    float dx, dy;
    if (MASKED) {
        if (hL && hR) dx = tp0.x - tpLx;
        else if (!hR) dx = -tpLx;
        else dx = tp0.x;
        if (hU && hD) dy = tp0.y - tpUy;
        else if (!hD) dy = -tpUy;
        else dy = tp0.y;
    } else {
        if (hL) dx = tp0.x - tpLx;
        else if (x == w - 1) dx = -tpLx;
        else dx = tp0.x;
        if (hU) dy = tp0.y - tpUy;
        else if (y == h - 1) dy = -tpUy;
        else dy = tp0.y;
    }
    return dx + dy;
}

// One TGV-L1 inner iteration in one launch: dual ascent + projection
// (updateDualVariables.cu:122-237 / :4-101), Tp (thresholdingL1.cu:234-244), L1
// thresholding (thresholdingL1.cu:128-216 / :3-86) and the primal step with
// over-relaxation (updatePrimalVariables.cu:138-245 / :3-113).
// Phase A computes the new p, q and Tp of the tile and of its left column and top row
// (the threshold and primal of a pixel read Tp and q of its left and upper neighbours)
// into LDS: each work-item takes the S::RPT tile pixels it finishes in phase B, and the
// first TW + TH + 1 work-items also one halo position.  Phase B finishes thresholding and
// the primal step.  Old duals are read from one buffer and new ones written to the other,
// so every workgroup recomputes its halo from the previous iteration's values (the
// reference's Jacobi order) and the result does not depend on the tile shape.  Every
// load is issued before any data-dependent branch.
template <class S, bool MASKED, typename T>
__device__ __forceinline__ void op_iter(const Tile &t, int roles, const IterParams &ip,
                                        float2 (*sP)[S::TW + 1], float4 (*sQ)[S::TW + 1], float2 (*sT)[S::TW + 1])
{
    constexpr int EXW = S::TW + 1;
    constexpr int HALO = EXW + S::TH;  // top row (incl. corner) + left column positions
    static_assert(HALO <= NT, "halo positions per workgroup");
    constexpr int BX = S::BX, BY = S::BY, RPT = S::RPT;
    const int w = t.w, h = t.h, s = t.s, x0 = t.x0, y0 = t.y0;
    const float *__restrict__ mask = t.mask;
    const float *__restrict__ u = slot(t, slot_u(rcur(roles, G_U)));
    float *__restrict__ us = slot(t, slot_u(roth(roles, G_U)));
    const float *__restrict__ u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float *__restrict__ u_s = slot(t, slot_ub(roth(roles, G_UB)));
    const float2 *__restrict__ v_ = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    float2 *__restrict__ v_s = slotp<float2>(t, slot_v(roth(roles, G_V)));
    const V2<T> *__restrict__ p_in = slotp<V2<T>>(t, slot_p(rcur(roles, G_PQ)));
    V2<T> *__restrict__ p_out = slotp<V2<T>>(t, slot_p(roth(roles, G_PQ)));
    const V4<T> *__restrict__ q_in = slotp<V4<T>>(t, slot_q(rcur(roles, G_PQ)));
    V4<T> *__restrict__ q_out = slotp<V4<T>>(t, slot_q(roth(roles, G_PQ)));
    float2 *__restrict__ Tp = slotp<float2>(t, PL_TP);
    const T *__restrict__ at = (const T *)slot(t, PL_A);
    const T *__restrict__ bt = (const T *)slot(t, PL_B);
    const T *__restrict__ ct = (const T *)slot(t, PL_C);
    const T *__restrict__ Iu = (const T *)slot(t, PL_IU);
    const T *__restrict__ Iz = (const T *)slot(t, PL_IZ);
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int x = x0 + threadIdx.x;

    // ---- loads: the tile pixels (rows threadIdx.y + BY * r) + one halo position
    DualIn in[RPT];
    float uu[RPT], iu[RPT], iz[RPT], tpl[RPT], tpu[RPT];
    bool hasL[RPT], hasU[RPT], gout[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = y0 + threadIdx.y + BY * r;
        gout[r] = !MASKED && outside_circle(x, y, w, h);
        dual_load<MASKED, T>(in[r], x, y, w, h, s, mask, u_, v_, at, bt, ct, p_in, q_in, ip.first);
        const long pos = in[r].pos;
        uu[r] = u[pos];
        iu[r] = ldT(Iu, pos);
        iz[r] = ldT(Iz, pos);
        const int xc = min(x, w - 1), yc = min(y, h - 1);
        if (MASKED) {
            hasL[r] = xc >= 1 && mask[xc >= 1 ? pos - 1 : pos] != 0.f;
            hasU[r] = yc >= 1 && mask[yc >= 1 ? pos - s : pos] != 0.f;
        } else {
            hasL[r] = xc >= 1;
            hasU[r] = yc >= 1;
        }
        // Tp-plane values of the left / upper neighbour: read only where that neighbour
        // is outside the mask or the image; fetched here on latency-bound levels
        tpl[r] = 0.f;
        tpu[r] = 0.f;
        if (ip.preload_tp) {
            tpl[r] = ld0(Tp, pos - 1).x;
            tpu[r] = ld0(Tp, pos - s).y;
        }
    }
    DualIn hin;
    int hx = 0, hy = 0, hex = 0, hey = 0;
    const bool has_halo = tid < HALO;
    if (has_halo) {
        if (tid < EXW) {
            hex = tid;
            hey = 0;
        } else {
            hex = 0;
            hey = tid - EXW + 1;
        }
        hx = x0 - 1 + hex;
        hy = y0 - 1 + hey;
        dual_load<MASKED, T>(hin, hx, hy, w, h, s, mask, u_, v_, at, bt, ct, p_in, q_in, ip.first);
    }

    // ---- phase A: new duals + Tp into LDS (and the tile's to global memory)
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        if (!in[r].act) continue;
        float2 pn, tpn;
        float4 qn;
        dual_step<MASKED>(in[r], gout[r], ip, pn, qn, tpn);
        const int ex = threadIdx.x + 1, ey = threadIdx.y + BY * r + 1;
        sP[ey][ex] = pn;
        sQ[ey][ex] = qn;
        sT[ey][ex] = tpn;
        st2T(p_out, in[r].pos, pn);
        st4T(q_out, in[r].pos, qn);
        if (ip.write_tp) Tp[in[r].pos] = tpn;
    }
    if (has_halo && hin.act) {
        float2 pn, tpn;
        float4 qn;
        dual_step<MASKED>(hin, !MASKED && outside_circle(hx, hy, w, h), ip, pn, qn, tpn);
        sP[hey][hex] = pn;
        sQ[hey][hex] = qn;
        sT[hey][hex] = tpn;
    }
    __syncthreads();

    // ---- phase B: thresholding + primal step on the tile
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const DualIn &q = in[r];
        const int y = y0 + threadIdx.y + BY * r;
        if (!q.ok) continue;
        const long pos = q.pos;
        if (!MASKED && gout[r]) {
            us[pos] = 0.f;
            u_s[pos] = 0.f;
            v_s[pos] = make_float2(0.f, 0.f);
            continue;
        }
        if (!q.act) continue;
        const bool hR = q.hasR, hD = q.hasD, hL = hasL[r], hU = hasU[r];
        const int ex = threadIdx.x + 1, ey = threadIdx.y + BY * r + 1;
        const float a = q.a, b = q.b, c = q.c;
        const float2 tp0 = sT[ey][ex];
        // a neighbour outside the mask (or the image) is read from the Tp plane, which
        // holds what earlier levels left there, as the reference does
        const float tpLx = hL ? sT[ey][ex - 1].x : (ip.preload_tp ? tpl[r] : ld0(Tp, pos - 1).x);
        const float tpUy = hU ? sT[ey - 1][ex].y : (ip.preload_tp ? tpu[r] : ld0(Tp, pos - s).y);
        const float divp = div_tp<MASKED>(x, y, w, h, hR, hD, hL, hU, tp0, tpLx, tpUy);
        const float4 ql = hL ? sQ[ey][ex - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 qu = hU ? sQ[ey - 1][ex] : make_float4(0.f, 0.f, 0.f, 0.f);
        float un, u_s_v;
        float2 v_s_v;
        primal_step(ip, a, b, c, q.u0, uu[r], iu[r], iz[r], q.v0, hR, hD, hL, hU, divp, sQ[ey][ex], ql, qu, tp0, un,
                    u_s_v, v_s_v);
        us[pos] = un;
        u_s[pos] = u_s_v;
        v_s[pos] = v_s_v;
    }
}

// ------------------------------------------------------------------ persistent launches
//
// Primitives of the persistent ops (OP_ITERPK below, OP_ITERP further down): exchange words
// between the workgroups of one launch, and its bookkeeping.

// Bounded waits of the persistent op, in ticks of the 100 MHz real-time counter.
constexpr unsigned long long kArriveTicks = 500000ull;  // every workgroup resident: 5 ms
constexpr unsigned long long kPollTicks = 5000000ull;   // a neighbour's words: 50 ms
// Where a work-item stages one position, its neighbours publish their ring words as it
// finishes its own step, and a load issued at once mostly finds them stale and costs
// another round trip through the device-coherent cache (0.6-1 us between two workgroups,
// measured).  Such work-items first sleep this long (s_sleep units of 64 clocks, about
// 0.4 us) before their first load: measured at 848x800, 2.9-3.6 -> 2.6-2.7 us per
// iteration in 27x25 tiles, 3.3 -> 2.3 in 14x13, 2.5-2.7 -> 2.0-2.1 in 8x8 and 7x7 (one
// iteration per exchange), 1.66 -> 1.49 in 8x8 blocks of four.
constexpr int kRingDelay = 16;

// The exchange words and the bookkeeping are device memory reached through descriptor
// pointers: accessed as global (not flat) memory, so that waiting for LDS operations
// does not wait for exchange loads in flight as well.
typedef __attribute__((address_space(1))) unsigned long long g_u64;
typedef __attribute__((address_space(1))) unsigned g_u32;
__device__ __forceinline__ unsigned long long xload(const unsigned long long *p)
{
    return __hip_atomic_load((const g_u64 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore(unsigned long long *p, float v, unsigned tag)
{
    __hip_atomic_store((g_u64 *)p, (unsigned long long)__float_as_uint(v) | ((unsigned long long)tag << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned xload32(const unsigned *p)
{
    return __hip_atomic_load((const g_u32 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore32(unsigned *p, unsigned v)
{
    __hip_atomic_store((g_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned xadd32(unsigned *p, unsigned v)
{
    return __hip_atomic_fetch_add((g_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned xswap32(unsigned *p, unsigned v)
{
    return __hip_atomic_exchange((g_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long rtime() { return __builtin_amdgcn_s_memrealtime(); }

// Every workgroup of a persistent launch must be resident: it counts its arrival and waits
// (bounded) for the others; s_word[0] receives the launch counter, s_word[1] whether to
// leave at once (a launch gave up since the host last cleared the flag, or not every
// workgroup arrived in time), and s_word[2], the workgroup's give-up flag
// (persist_any), starts clear.
__device__ __forceinline__ void persist_arrive(XSync *__restrict__ xs, int ntiles, unsigned *s_word)
{
    if (threadIdx.x == 0) {
        const unsigned ep = xload32(&xs->epoch);
        unsigned ab = xload32(&xs->abort);
        xadd32(&xs->arrive, 1u);
        if (!ab) {
            const unsigned long long t0 = rtime();
            while (xload32(&xs->arrive) < (unsigned)ntiles) {
                if (rtime() - t0 > kArriveTicks) {
                    if (xswap32(&xs->abort, 1u) == 0) xstore32(&xs->why, ((unsigned)ntiles << 12) | (1u << 8));
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        s_word[0] = ep;
        s_word[1] = ab;
        s_word[2] = 0u;
    }
    __syncthreads();
}

// Whether some work-item of the workgroup gave up a wait (`bad` != 0), as a barrier: the
// work-items that did set the workgroup's flag (LDS) before it, and every work-item reads
// it after it.  The flag is never cleared: the workgroup leaves once it is set.  Every
// call is followed by another barrier before the next one (the launches' phase barriers),
// so no work-item sets the flag for a later call while another still reads it for this
// one.  One barrier, where a workgroup reduction (__syncthreads_or) takes three.
__device__ __forceinline__ bool persist_any(int bad, unsigned *s_word)
{
    if (bad) s_word[2] = 1u;
    __syncthreads();
    return s_word[2] != 0u;
}

// A work-item that gave up waiting (`bad`: 2 ring values, 3 a neighbour's progress word)
// sets the abort flag, and records why unless another workgroup did first.
__device__ __forceinline__ void persist_give_up(XSync *__restrict__ xs, int ntiles, int bad, int step)
{
    if (bad && xswap32(&xs->abort, 1u) == 0)
        xstore32(&xs->why, ((unsigned)ntiles << 12) | ((unsigned)bad << 8) | ((unsigned)step & 255u));
}

// The last workgroup to finish resets the arrivals and advances the launch counter.
__device__ __forceinline__ void persist_finish(XSync *__restrict__ xs, int ntiles, unsigned epoch)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned d = xadd32(&xs->done, 1u);
        if (d == (unsigned)ntiles - 1) {
            xstore32(&xs->arrive, 0u);
            xstore32(&xs->done, 0u);
            xstore32(&xs->epoch, epoch + 1u);
        }
    }
}

// Waits (bounded) until the word at `p` carries `tag`; false if it gave up (waited too
// long, or another workgroup gave up).
__device__ __forceinline__ bool persist_wait_tag(const unsigned long long *p, unsigned tag, XSync *__restrict__ xs)
{
    unsigned long long t0 = 0;
    for (int polls = 0; (unsigned)(xload(p) >> 32) != tag; polls++) {
        if (polls == 0) t0 = rtime();
        if ((polls & 63) == 63 && (rtime() - t0 > kPollTicks || xload32(&xs->abort) != 0)) return false;
    }
    return true;
}

// ------------------------------------------------------------------ K iterations per op
//
// Temporal blocking of the inner iteration (OP_ITER2, OP_ITER4, OP_ITER8): one workgroup advances
// its tile by K iterations.  Iteration j (r = K - 1 - j iterations after it) computes the
// duals over the tile grown by 2r + 1 pixels to the left / top and r to the right /
// bottom, and the primal step over the tile grown by 2r / r: the primal step of a pixel
// reads the new q and Tp of its left and upper neighbours, the dual step the state of all
// four neighbours.  The workgroup stages the state over the tile grown by 2K / K pixels
// (every global load issued before the first LDS store): each work-item keeps what only
// its own positions read in registers, and what neighbours read in LDS; the K iterations
// then run in place.  Every pixel goes through the same dual_step / primal_step code as
// in op_iter, and p and q are rounded to the storage type between iterations as the
// planes round them, so the result equals K single iterations bit for bit.
// Buffers (K even): iteration K - 2 stores u, u_ and v_ to the `other` buffers (as a
// single iteration's would), iteration K - 1 its whole state to the spare buffers
// together with, outside the mask, the current buffers' values (never written by an
// iteration); the host then
// rotates the roles (spare -> current).  Out of the mask, iteration j reads the current
// buffers' values for even j and the other buffers' values for odd j, as the reference's
// two ping-pong buffers give them.  Other workgroups of the launch read the current
// buffers and, outside the mask only, the other buffers, so no write of this op is read
// by it.  The spare buffers also take the current buffers' values in the row padding
// (columns w .. s - 1, written by the tiles of the last tile column), so that the new
// current buffers equal the old ones wherever no iteration writes: coarser levels'
// values there are read again at finer levels.  The host uses this op only where no
// mask pixel lies in column 0: there the reference's flat-index left neighbour (pos - 1)
// is the previous row's last pixel.
// The step sizes of iteration j come from the pair's schedule table at index it0 + j.
// Work-items: see BlkK::NT.
template <int SH, int K, bool HALF>
struct BlkK {
    static constexpr int TW = kShape[SH].tw, TH = kShape[SH].th;
    static constexpr int RO = -2 * K, RW = TW + 3 * K, RH = TH + 3 * K;  // staged region
    static constexpr int NR = RW * RH;
    // work-items (half_workgroup_ok)
    static constexpr int NT = HALF ? 512 : NR >= 1024 ? 1024 : (NR + 63) / 64 * 64;
    static_assert(!HALF || half_workgroup_ok(SH, K), "half workgroup");
    static constexpr int NK = (NR + NT - 1) / NT;  // staged positions per work-item
    static_assert(multi_iter_shape(SH), "multi-iteration tile shape");
};

// Values that neighbours read: u_ and v_ (inside the mask the iteration's state; outside
// it the value of the buffer the reference reads at that step), q, Tp (inside the mask
// the iteration's new Tp, outside it the Tp plane) and the mask.
template <int SH, int K>
struct LdsK {
    using B = BlkK<SH, K, false>;
    float ub[B::NR];
    float2 v[B::NR];
    float4 q[B::NR];
    float2 tp[B::NR];
    unsigned char mk[B::NR];
};

// A work-item's own staged position: everything only that position reads.
struct OwnK {
    float a, b, c, iu, iz;
    float u, ub;        // u and u_ (current buffers, then each iteration's)
    float2 v;           // v_
    float2 p;           // p: the last dual step's value (q: in LDS)
    float ubo;          // outside the mask: u_ and v_ of the other buffers
    float2 vo;
};

// Rounds a dual value to the storage type: the planes hold p and q in it, and every
// iteration's dual step reads them from there (its primal step uses the unrounded values).
template <typename T>
__device__ __forceinline__ float rT(float v)
{
    return Cv<T>::f(Cv<T>::t(v));
}

// Exchange words of the blocked persistent op (OP_ITERPK), one tile and parity: u, u_, v_.x,
// v_.y, p.x, p.y, q.x, q.y, q.z, q.w of every tile position (value-major, row-major within
// a value plane), then the tile's progress word (HB).
template <int SH>
struct StripK {
    static constexpr int TW = kShape[SH].tw, TH = kShape[SH].th;
    static constexpr int VP = TW * TH, HB = 10 * VP, WORDS = HB + 1;
    static_assert(WORDS == persistk_words(SH), "strip layout");
};

// The blocked persistent op's launch: `blocks` blocks of K iterations (the last one of K -
// `skip`), this workgroup's tile among the launch's `ntiles`, the pair's exchange words and
// bookkeeping, and two words of LDS for the arrival.
struct PersistK {
    int blocks = 1, skip = 0;
    int tile = 0, ntiles = 0;
    XSync *xs = nullptr;
    unsigned long long *xbuf = nullptr;
    unsigned *s_word = nullptr;
};

// The step sizes of iteration j: sched[j] = (sigma, tau, mu, alpha1 sigma / 3); `ip0` holds the
// schedule-independent fields and the flags of the op.
// PERSIST (OP_ITERPK): the workgroups of the launch are persistent, one per tile.  A
// workgroup stages its tile grown by 2K / K once and runs pk.blocks blocks of K iterations
// in place, the last one shortened to K - pk.skip iterations (an even number: the block's
// last iterations, whose regions the staged state covers).  After each block but the last it publishes the state of the bands of its tile
// that neighbouring rings cover (columns < K and >= tw - 2K, rows < K and >= th - 2K; the
// tile is at least 2K on each side, so a ring reads its 8 neighbours only) and then
// refreshes its own ring (everything staged outside the tile that lies in the image and the
// mask) from its neighbours' words; outside the mask nothing changes.  Words carry tags,
// are double-buffered by block parity and every tile waits for all its neighbours'
// progress words, as in OP_ITERP.  Only the last block stores the state (other and spare
// buffers, as the K-iteration op does), so the result equals the same iterations run one
// per launch.
template <int SH, int K, bool HALF, bool MASKED, typename T, bool PERSIST = false>
__device__ __forceinline__ void op_iterk(const Tile &t, int roles, const float4 *__restrict__ sched,
                                         const IterParams &ip0, LdsK<SH, K> &L, const PersistK &pk = PersistK())
{
    static_assert(K % 2 == 0, "K-iteration ops rotate the roles like two iterations");
    using B = BlkK<SH, K, HALF>;
    constexpr int NT = B::NT, NK = B::NK;
    static_assert(!PERSIST || (!HALF && NK <= 2 && persistk_shape(SH, K)), "persistent K-iteration shape");
    const int w = t.w, h = t.h, s = t.s, x0 = t.x0, y0 = t.y0;
    const int tid = threadIdx.x;
    unsigned epoch = 0;
    if constexpr (PERSIST) {
        // every workgroup of the launch must be resident
        persist_arrive(pk.xs, pk.ntiles, pk.s_word);
        epoch = pk.s_word[0];
        if (pk.s_word[1]) {
            persist_finish(pk.xs, pk.ntiles, epoch);
            return;
        }
    }
    const float *__restrict__ mask = t.mask;
    const float *__restrict__ u_cur = slot(t, slot_u(rcur(roles, G_U)));
    float *__restrict__ u_oth = slot(t, slot_u(roth(roles, G_U)));
    float *__restrict__ u_spr = slot(t, slot_u(rspare(roles, G_U)));
    const float *__restrict__ ub_cur = slot(t, slot_ub(rcur(roles, G_UB)));
    float *__restrict__ ub_oth = slot(t, slot_ub(roth(roles, G_UB)));
    float *__restrict__ ub_spr = slot(t, slot_ub(rspare(roles, G_UB)));
    const float2 *__restrict__ v_cur = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    float2 *__restrict__ v_oth = slotp<float2>(t, slot_v(roth(roles, G_V)));
    float2 *__restrict__ v_spr = slotp<float2>(t, slot_v(rspare(roles, G_V)));
    const V2<T> *__restrict__ p_cur = slotp<V2<T>>(t, slot_p(rcur(roles, G_PQ)));
    V2<T> *__restrict__ p_spr = slotp<V2<T>>(t, slot_p(rspare(roles, G_PQ)));
    const V4<T> *__restrict__ q_cur = slotp<V4<T>>(t, slot_q(rcur(roles, G_PQ)));
    V4<T> *__restrict__ q_spr = slotp<V4<T>>(t, slot_q(rspare(roles, G_PQ)));
    float2 *__restrict__ Tp = slotp<float2>(t, PL_TP);
    const T *__restrict__ at = (const T *)slot(t, PL_A);
    const T *__restrict__ bt = (const T *)slot(t, PL_B);
    const T *__restrict__ ct = (const T *)slot(t, PL_C);
    const T *__restrict__ Iu = (const T *)slot(t, PL_IU);
    const T *__restrict__ Iz = (const T *)slot(t, PL_IZ);
    // The duals of the `other` buffers are never read (an op reads the current ones and
    // each warp starts from zero): the op stores p and q after its last iteration only.

    // ---- the warp's derivatives (EF_DERIV) at the staged positions, first and on their
    // own (their registers are free again before the staging), into this work-item's own
    // Tp slots in LDS until the staging reads them
    if (ip0.deriv_level >= 0) {
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int i = tid + k * NT;
            if (i >= B::NR) continue;
            const int x = x0 + B::RO + i % B::RW, y = y0 + B::RO + i / B::RW;
            const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
            const bool own = xc == x && yc == y && x >= x0 && x < x0 + B::TW && y >= y0 && y < y0 + B::TH;
            L.tp[i] = staged_deriv<T>(t, ip0.deriv_level, xc, yc, own);
        }
    }

    // ---- stage: own values into registers, shared ones into LDS (clamped positions
    // outside the image); every load is issued before the first LDS store
    OwnK o[NK];
    int ox[NK], oy[NK];
    float omk[NK];
    float2 otp[NK];
    float4 oq[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT;
        ox[k] = x0 + B::RO + i % B::RW;
        oy[k] = y0 + B::RO + i / B::RW;
        if (i < B::NR) {
            const int xc = min(max(ox[k], 0), w - 1), yc = min(max(oy[k], 0), h - 1);
            const long pos = (long)yc * s + xc;
            omk[k] = MASKED ? mask[pos] : 1.f;
            o[k].ub = ub_cur[pos];
            o[k].v = v_cur[pos];
            o[k].u = u_cur[pos];
            o[k].a = ldT(at, pos);
            o[k].b = ldT(bt, pos);
            o[k].c = ldT(ct, pos);
            if (ip0.deriv_level >= 0) {
                const float2 d = L.tp[i];  // computed above
                o[k].iu = d.x;
                o[k].iz = d.y;
            } else {
                o[k].iu = ldT(Iu, pos);
                o[k].iz = ldT(Iz, pos);
            }
            if (ip0.first) {
                o[k].p = make_float2(0.f, 0.f);
                oq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                o[k].p = ld2T(p_cur, pos);
                oq[k] = ld4T(q_cur, pos);
            }
        }
    }
    // outside the mask: the other buffers' u_, v_ and the Tp plane (read by neighbours)
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT;
        o[k].ubo = 0.f;
        o[k].vo = make_float2(0.f, 0.f);
        otp[k] = make_float2(0.f, 0.f);
        if (MASKED && i < B::NR && omk[k] == 0.f) {
            const int xc = min(max(ox[k], 0), w - 1), yc = min(max(oy[k], 0), h - 1);
            const long pos = (long)yc * s + xc;
            o[k].ubo = ub_oth[pos];
            o[k].vo = v_oth[pos];
            otp[k] = Tp[pos];
        }
    }
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT;
        if (i < B::NR) {
            L.ub[i] = o[k].ub;
            L.v[i] = o[k].v;
            L.q[i] = oq[k];
            L.tp[i] = otp[k];
            L.mk[i] = omk[k] != 0.f ? 1 : 0;
        }
    }
    // row padding of the last tile column: the spare buffers take the current values
    if (x0 + B::TW >= w && s > w) {
        const int pw = s - w, rows = min(B::TH, h - y0);
        for (int i = tid; i < pw * rows; i += NT) {
            const long pos = (long)(y0 + i / pw) * s + w + i % pw;
            u_spr[pos] = u_cur[pos];
            ub_spr[pos] = ub_cur[pos];
            v_spr[pos] = v_cur[pos];
        }
    }
    // A workgroup whose staged region lies inside the image and (masked form) inside the
    // mask, or (unmasked form) inside the circle, takes every neighbour branch the same
    // way: it runs the iterations without the boundary tests (same arithmetic).  In the
    // persistent form with 8x8 tiles every workgroup takes the path with them: measured
    // at 848x800 in blocks of four, 1.40 against 1.47 us per iteration at 106x100 and 1.27
    // against 1.42 at 53x50 (in 14x13 tiles the split is the faster one).
    constexpr bool kFullPath = !(PERSIST && SH == 7);
    bool inner = true;
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT, x = ox[k], y = oy[k];
        if (i < B::NR)
            inner = inner && x >= 0 && y >= 0 && x < w && y < h && omk[k] != 0.f &&
                    (MASKED || !outside_circle(x, y, w, h));
    }
    const bool full = kFullPath ? __syncthreads_and(inner) != 0 : (__syncthreads(), false);
    // the primal step's per-pixel constants, once for the K iterations
    PrimalCoef kc[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) kc[k] = primal_coef(ip0, o[k].a, o[k].b, o[k].c, o[k].iu);

    auto act = [&](int k) { return !MASKED || L.mk[k] != 0; };
    auto in_tile = [&](int x, int y) { return x >= x0 && x < x0 + B::TW && y >= y0 && y < y0 + B::TH; };
    // Per own position, once per launch (the mask does not change): in the image, in the
    // mask, and which of the four neighbours are in the image and the mask (read by the
    // path with boundary tests only).
    enum : unsigned { F_IN = 1, F_ACT = 2, F_R = 4, F_D = 8, F_L = 16, F_U = 32, F_OUT = 64 };
    unsigned fl[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT, x = ox[k], y = oy[k];
        fl[k] = 0;
        if (!full && i < B::NR && x >= 0 && y >= 0 && x < w && y < h) {
            // neighbours outside the staged region are never asked for
            const int col = i % B::RW, row = i / B::RW;
            fl[k] = F_IN | (act(i) ? F_ACT : 0u);
            if (x + 1 < w && col + 1 < B::RW && act(i + 1)) fl[k] |= F_R;
            if (y + 1 < h && row + 1 < B::RH && act(i + B::RW)) fl[k] |= F_D;
            if (x >= 1 && col >= 1 && act(i - 1)) fl[k] |= F_L;
            if (y >= 1 && row >= 1 && act(i - B::RW)) fl[k] |= F_U;
            if (!MASKED && outside_circle(x, y, w, h)) fl[k] |= F_OUT;  // the unmasked form's gate
        }
    }
    // Per own position, once per launch, the number of iterations whose dual / primal
    // region holds it (the dual's none outside the image or the mask on the path with
    // boundary tests): iteration j, r = K - 1 - j, takes the duals of the tile grown by
    // 2r + 1 to the left / top and r to the right / bottom, and the primal steps of the
    // tile grown by 2r and r.  The offsets from (x0, y0) depend on the position index only.
    // Kernels held to 128 registers (three positions per work-item, half workgroups) test
    // the regions per phase instead: the counts' registers would spill there.
    constexpr bool kCounts = NK <= 2 && !HALF;
    int nd[NK], np[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int i = tid + k * NT;
        const int dx = i % B::RW + B::RO, dy = i / B::RW + B::RO;
        const int rdx = max(max(0, (-dx) / 2), dx - (B::TW - 1)), rdy = max(max(0, (-dy) / 2), dy - (B::TH - 1));
        const int rpx = max(max(0, (1 - dx) / 2), dx - (B::TW - 1)), rpy = max(max(0, (1 - dy) / 2), dy - (B::TH - 1));
        const bool staged = i < B::NR;
        nd[k] = staged && (full || ((fl[k] & F_ACT) && (fl[k] & F_IN))) ? K - max(rdx, rdy) : 0;
        np[k] = staged ? K - max(rpx, rpy) : 0;
    }

    // ---- the persistent form's exchange roles, once per launch: xr[k] >= 0 for a ring
    // position that neighbours update (its tile's value-0 word of parity 0), -2 for a tile
    // position in the published bands, -1 otherwise; hb: the work-items 0..7 wait for the
    // neighbours' progress words (that neighbour's word of parity 0, or -1 at the image
    // border); pf / pfh: with one position per work-item, its ring words and the progress
    // word loaded at the end of each block but the last and checked at the next block's
    // start (polled again while older).
    using X = StripK<SH>;
    int xr[NK], hb = -1;
    unsigned long long *__restrict__ mine = nullptr;
    const unsigned tag0 = epoch << 8;
    unsigned long long pf[NK == 1 ? 10 : 1], pfh = 0;
#pragma unroll
    for (int v = 0; v < (NK == 1 ? 10 : 1); v++) pf[v] = 0;
    if constexpr (PERSIST) {
        constexpr int TW = B::TW, TH = B::TH;
        const int ntx = (w + TW - 1) / TW, nty = (h + TH - 1) / TH;
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int i = tid + k * NT, x = ox[k], y = oy[k];
            xr[k] = -1;
            if (i >= B::NR || x < 0 || y < 0 || x >= w || y >= h || !act(i)) continue;
            const int dx = x - x0, dy = y - y0;
            if (in_tile(x, y)) {
                if (dx < K || dx >= TW - 2 * K || dy < K || dy >= TH - 2 * K) xr[k] = -2;
            } else {
                const int sx = x / TW, sy = y / TH;
                xr[k] = (sy * ntx + sx) * 2 * X::WORDS + (y - sy * TH) * TW + (x - sx * TW);
            }
        }
        if (tid < 8) {
            const int ddx = tid < 3 ? tid - 1 : tid < 5 ? (tid == 3 ? -1 : 1) : tid - 6;
            const int ddy = tid < 3 ? -1 : tid < 5 ? 0 : 1;
            const int nx = x0 / TW + ddx, ny = y0 / TH + ddy;
            if (nx >= 0 && ny >= 0 && nx < ntx && ny < nty) hb = (ny * ntx + nx) * 2 * X::WORDS + X::HB;
        }
        mine = pk.xbuf + (long)pk.tile * 2 * X::WORDS;
    }

    // one block of K iterations (`blk`: its index, iterations blk K ..; `last`: the op's last
    // block, which stores the state; a compile-time flag, so that no store address is kept
    // across the blocks before it).  The persistent form's blocks but the last publish the
    // bands in their last iteration's primal step, as each band position is computed.
    // `skip`: the block runs iterations skip .. K - 1 only (even: the iterations keep the
    // parity of their index in the warp, which the out-of-mask values follow)
    auto iterate = [&](auto full_c, int blk, auto last_c, int skip) {
        constexpr bool FULL = decltype(full_c)::value;
        constexpr bool last = decltype(last_c)::value;
        // unrolled: each iteration's regions and stores are known at compile time (fewer
        // registers and tests than a loop over j)
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j < skip) continue;
            IterParams ip = ip0;
            const float4 sc = sched[blk * K + j - skip];
            ip.k1 = sc.w;  // alpha1 sigma / 3, as op_iter forms it
            ip.k0 = ip0.alpha0 * sc.x / 2.0f;
            ip.tau = sc.y;
            ip.mu = sc.z;
            ip.first = j == skip && blk == 0 && ip0.first;
            ip.write_tp = j == K - 1 && last && ip0.write_tp;
            const int r = K - 1 - j;

            // ---- duals over the tile grown by 2r + 1 (left / top) and r (right / bottom)
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int i = tid + k * NT, x = ox[k], y = oy[k];
                if (kCounts) {
                    if (nd[k] <= j) continue;
                } else {
                    if (i >= B::NR) continue;
                    if (!FULL && !(fl[k] & F_ACT && fl[k] & F_IN)) continue;
                    if (x < x0 - 2 * r - 1 || y < y0 - 2 * r - 1 || x > x0 + B::TW - 1 + r || y > y0 + B::TH - 1 + r)
                        continue;
                }
                const int kR = (FULL || x + 1 < w) ? i + 1 : i, kD = (FULL || y + 1 < h) ? i + B::RW : i;
                DualIn rr;
                rr.ok = true;
                rr.act = true;
                rr.hasR = FULL || (fl[k] & F_R) != 0;
                rr.hasD = FULL || (fl[k] & F_D) != 0;
                rr.a = o[k].a;
                rr.b = o[k].b;
                rr.c = o[k].c;
                rr.u0 = o[k].ub;
                rr.uR = L.ub[kR];
                rr.uD = L.ub[kD];
                // backward differences (read where the right / lower neighbour is outside
                // the mask; column 0 inside the mask is not run here)
                rr.uL = (!FULL && x >= 1) ? L.ub[i - 1] : 0.f;
                rr.uU = (!FULL && y >= 1) ? L.ub[i - B::RW] : 0.f;
                rr.v0 = o[k].v;
                rr.vR = L.v[kR];
                rr.vD = L.v[kD];
                rr.vL = (!FULL && x >= 1) ? L.v[i - 1] : make_float2(0.f, 0.f);
                rr.vU = (!FULL && y >= 1) ? L.v[i - B::RW] : make_float2(0.f, 0.f);
                // the previous iteration's duals as the planes hold them
                rr.po = make_float2(rT<T>(o[k].p.x), rT<T>(o[k].p.y));
                const float4 qo = L.q[i];
                rr.qo = make_float4(rT<T>(qo.x), rT<T>(qo.y), rT<T>(qo.z), rT<T>(qo.w));
                float2 pn, tpn;
                float4 qn;
                dual_step<MASKED>(rr, !FULL && (fl[k] & F_OUT) != 0, ip, pn, qn, tpn);
                o[k].p = pn;  // the primal step of this iteration reads them unrounded
                L.q[i] = qn;
                L.tp[i] = tpn;
                if (r == 0 && last && in_tile(x, y)) {
                    const long pos = (long)y * s + x;
                    st2T(p_spr, pos, pn);
                    st4T(q_spr, pos, qn);
                    if (ip.write_tp) Tp[pos] = tpn;
                }
            }
            __syncthreads();

            // ---- thresholding + primal step over the tile grown by 2r (left / top) and
            // r; outside the mask, the stored u_ and v_ switch to the buffer the next
            // iteration reads there
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int i = tid + k * NT, x = ox[k], y = oy[k];
                if ((!kCounts || !FULL) && i >= B::NR) continue;
                if (!FULL && !(fl[k] & F_IN)) continue;
                const long pos = (long)y * s + x;
                if (!FULL && MASKED && !(fl[k] & F_ACT)) {
                    if (r == 0 && last) {
                        // outside the mask: the spare buffers take the current buffers' values
                        if (in_tile(x, y)) {
                            u_spr[pos] = o[k].u;
                            ub_spr[pos] = o[k].ub;
                            v_spr[pos] = o[k].v;
                        }
                    } else {
                        L.ub[i] = (j & 1) ? o[k].ub : o[k].ubo;
                        L.v[i] = (j & 1) ? o[k].v : o[k].vo;
                    }
                    continue;
                }
                if (kCounts) {
                    if (np[k] <= j) continue;
                } else if (x < x0 - 2 * r || y < y0 - 2 * r || x > x0 + B::TW - 1 + r || y > y0 + B::TH - 1 + r) {
                    continue;
                }
                float un, u_s;
                float2 vs;
                if (!FULL && !MASKED && (fl[k] & F_OUT)) {
                    un = 0.f;
                    u_s = 0.f;
                    vs = make_float2(0.f, 0.f);
                } else {
                    const bool hR = FULL || (fl[k] & F_R) != 0, hD = FULL || (fl[k] & F_D) != 0;
                    const bool hL = FULL || (fl[k] & F_L) != 0, hU = FULL || (fl[k] & F_U) != 0;
                    // the left / upper neighbour's Tp: this iteration's inside the mask,
                    // the Tp plane's outside it
                    const float tpLx = (FULL || x >= 1) ? L.tp[i - 1].x : 0.f;
                    const float tpUy = (FULL || y >= 1) ? L.tp[i - B::RW].y : 0.f;
                    const float2 tp0 = L.tp[i];
                    const float divp = div_tp<MASKED>(x, y, w, h, hR, hD, hL, hU, tp0, tpLx, tpUy);
                    const float4 ql = hL ? L.q[i - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const float4 qu = hU ? L.q[i - B::RW] : make_float4(0.f, 0.f, 0.f, 0.f);
                    // the v step's T p is the staged Tp (the same value); kernels held to 128
                    // registers recompute it from their registers instead of keeping it
                    const float2 tpv = kCounts ? tp0 : tensor_times(o[k].a, o[k].b, o[k].c, o[k].p);
                    primal_core(ip, kc[k], o[k].ub, o[k].u, o[k].iu, o[k].iz, o[k].v, hR, hD, hL, hU, divp, L.q[i], ql,
                                qu, tpv, un, u_s, vs);
                }
                o[k].u = un;
                o[k].ub = u_s;
                o[k].v = vs;
                L.ub[i] = u_s;
                L.v[i] = vs;
                if constexpr (PERSIST && !last) {
                    if (r == 0 && xr[k] == -2) {
                        // a band position: its state after the block (q: this work-item's own
                        // LDS slot, written by this iteration's dual step)
                        constexpr int TW = B::TW;
                        const unsigned tg = tag0 | (unsigned)blk;
                        unsigned long long *pp = mine + (blk & 1) * X::WORDS + (y - y0) * TW + (x - x0);
                        const float4 q = L.q[i];
                        xstore(pp, un, tg);
                        xstore(pp + X::VP, u_s, tg);
                        xstore(pp + 2 * X::VP, vs.x, tg);
                        xstore(pp + 3 * X::VP, vs.y, tg);
                        xstore(pp + 4 * X::VP, o[k].p.x, tg);
                        xstore(pp + 5 * X::VP, o[k].p.y, tg);
                        xstore(pp + 6 * X::VP, q.x, tg);
                        xstore(pp + 7 * X::VP, q.y, tg);
                        xstore(pp + 8 * X::VP, q.z, tg);
                        xstore(pp + 9 * X::VP, q.w, tg);
                    }
                }
                if (last && r == 1 && in_tile(x, y)) {
                    u_oth[pos] = un;
                    ub_oth[pos] = u_s;
                    v_oth[pos] = vs;
                } else if (last && r == 0) {
                    u_spr[pos] = un;
                    ub_spr[pos] = u_s;
                    v_spr[pos] = vs;
                }
            }
            if constexpr (PERSIST && !last && NK == 1) {
                // the words the next block's start reads
                if (r == 0) {
                    const int par = blk & 1;
                    if (xr[0] >= 0) __builtin_amdgcn_s_sleep(kRingDelay);
                    if (xr[0] >= 0) {
#pragma unroll
                        for (int v = 0; v < 10; v++) pf[v] = xload(pk.xbuf + xr[0] + par * X::WORDS + v * X::VP);
                    }
                    if (hb >= 0) pfh = xload(pk.xbuf + hb + par * X::WORDS);
                }
            }
            __syncthreads();
        }
    };
    if constexpr (!PERSIST) {
        if (full) iterate(std::true_type{}, 0, std::true_type{}, 0);
        else iterate(std::false_type{}, 0, std::true_type{}, 0);
    } else {
        int blk = 0;
        bool gave_up = false;
        for (;; blk++) {
            if (blk > 0) {
                const unsigned tg = tag0 | (unsigned)(blk - 1);
                const int par = (blk - 1) & 1;
                // ---- the neighbours' progress, then the ring's state (both published in
                // the previous block)
                int bad = 0;
                if (hb >= 0 && (unsigned)(pfh >> 32) != tg && !persist_wait_tag(pk.xbuf + hb + par * X::WORDS, tg, pk.xs))
                    bad = 3;
#pragma unroll
                for (int k = 0; k < NK; k++) {
                    if (xr[k] < 0 || bad) continue;
                    const unsigned long long *pw = pk.xbuf + xr[k] + par * X::WORDS;
                    float f[10];
                    bool have = false;
                    if constexpr (NK == 1) {
                        have = true;
#pragma unroll
                        for (int v = 0; v < 10; v++) have = have && (unsigned)(pf[v] >> 32) == tg;
#pragma unroll
                        for (int v = 0; v < 10; v++) f[v] = __uint_as_float((unsigned)pf[v]);
                    }
                    // else the ten words in two groups of five (registers: a 64-bit word each)
#pragma unroll
                    for (int g = 0; g < 2 && !have; g++) {
                        unsigned long long a[5];
                        unsigned long long t0 = 0;
                        for (int polls = 0;; polls++) {
                            bool ok = true;
#pragma unroll
                            for (int v = 0; v < 5; v++) a[v] = xload(pw + (g * 5 + v) * X::VP);
#pragma unroll
                            for (int v = 0; v < 5; v++) ok = ok && (unsigned)(a[v] >> 32) == tg;
                            if (ok) break;
                            if (polls == 0) t0 = rtime();
                            if ((polls & 63) == 63 && (rtime() - t0 > kPollTicks || xload32(&pk.xs->abort) != 0)) {
                                bad = 2;
                                break;
                            }
                        }
#pragma unroll
                        for (int v = 0; v < 5; v++) f[g * 5 + v] = __uint_as_float((unsigned)a[v]);
                    }
                    const int i = tid + k * NT;
                    o[k].u = f[0];
                    o[k].ub = f[1];
                    o[k].v = make_float2(f[2], f[3]);
                    o[k].p = make_float2(f[4], f[5]);
                    L.ub[i] = o[k].ub;
                    L.v[i] = o[k].v;
                    L.q[i] = make_float4(f[6], f[7], f[8], f[9]);
                }
                if (persist_any(bad, pk.s_word)) {
                    persist_give_up(pk.xs, pk.ntiles, bad, blk);
                    gave_up = true;
                    break;
                }
            }
            if (blk == pk.blocks - 1) break;
            // progress: this tile has read its neighbours' words of block blk - 1, which
            // they overwrite in block blk + 1
            if (tid == NT - 1) xstore(mine + (blk & 1) * X::WORDS + X::HB, 0.f, tag0 | (unsigned)blk);
            if constexpr (kFullPath) {
                if (full) iterate(std::true_type{}, blk, std::false_type{}, 0);
                else iterate(std::false_type{}, blk, std::false_type{}, 0);
            } else {
                iterate(std::false_type{}, blk, std::false_type{}, 0);
            }
        }
        if (!gave_up) {
            if constexpr (kFullPath) {
                if (full) iterate(std::true_type{}, blk, std::true_type{}, pk.skip);
                else iterate(std::false_type{}, blk, std::true_type{}, pk.skip);
            } else {
                iterate(std::false_type{}, blk, std::true_type{}, pk.skip);
            }
        }
        persist_finish(pk.xs, pk.ntiles, epoch);
    }
}

// Half workgroups (512 work-items, two or three positions each) are held to 128 registers, so that
// two of them share a compute unit.
template <typename T, int SH, int K, bool HALF>
__global__ __launch_bounds__((BlkK<SH, K, HALF>::NT)) __attribute__((amdgpu_waves_per_eu(HALF ? 4 : 1)))
void k_iterk(const PairDesc *__restrict__ D, const TickArgs A)
{
    using B = BlkK<SH, K, HALF>;
    __shared__ LdsK<SH, K> L;
    const int nb = gridDim.x, i = blockIdx.x, xcd = i & 7, q = nb >> 3, r = nb & 7;
    const int b = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + (i >> 3);
    int lo = 0;
    if (A.n > 1) {
        const int lane = __lane_id();
        const int ts = lane < A.n ? A.e[lane].tile_start : 0x7fffffff;
        lo = __popcll(__ballot(ts <= b)) - 1;
        lo = __builtin_amdgcn_readfirstlane(lo);
    }
    const TickEntry &E = A.e[lo];
    Tile t;
    t.d = D + E.pair;
    t.arena = E.arena;
    t.plane = E.plane;
    t.mask = E.mask;
    t.w = E.w;
    t.h = E.h;
    t.s = E.s;
    const int local = b - E.tile_start;
    const int ntx = (t.w + B::TW - 1) / B::TW;
    t.x0 = (local % ntx) * B::TW;
    t.y0 = (local / ntx) * B::TH;
    // step sizes of iterations it0 .. it0 + K - 1 of the warp (stereotgv.cu:449-477)
    const float4 *sched = t.d->sched + E.it0;
    IterParams ip0;
    ip0.k1 = ip0.k0 = ip0.tau = ip0.mu = 0.f;  // set per iteration
    ip0.lambda = A.lambda;
    ip0.alpha0 = A.alpha0;
    ip0.alpha1 = A.alpha1;
    ip0.first = (E.flags & EF_FIRST) != 0;
    ip0.write_tp = (E.flags & EF_WRITE_TP_LAST) != 0;  // the op's last iteration
    ip0.preload_tp = 0;
    ip0.deriv_level = (E.flags & EF_DERIV) ? (int)E.level : -1;
    if (E.masked) op_iterk<SH, K, HALF, true, T>(t, E.roles, sched, ip0, L);
    else op_iterk<SH, K, HALF, false, T>(t, E.roles, sched, ip0, L);
}

// The blocked persistent op (OP_ITERPK): one pair per launch, one workgroup per tile (tiles
// in XCD-contiguous runs as the other op kernels), an even number of iterations in blocks
// of K (the last one shortened to the rest).
template <typename T, int SH, int K>
__global__ __launch_bounds__((BlkK<SH, K, false>::NT)) void k_iterkp(const PairDesc *__restrict__ D, const TickArgs A)
{
    using B = BlkK<SH, K, false>;
    __shared__ LdsK<SH, K> L;
    __shared__ unsigned s_word[3];
    const int nb = gridDim.x, i = blockIdx.x, xcd = i & 7, q = nb >> 3, r = nb & 7;
    const int b = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + (i >> 3);
    const TickEntry &E = A.e[0];
    Tile t;
    t.d = D + E.pair;
    t.arena = E.arena;
    t.plane = E.plane;
    t.mask = E.mask;
    t.w = E.w;
    t.h = E.h;
    t.s = E.s;
    const int ntx = (t.w + B::TW - 1) / B::TW;
    t.x0 = (b % ntx) * B::TW;
    t.y0 = (b / ntx) * B::TH;
    const float4 *sched = t.d->sched + iterp_first(E.it0);
    IterParams ip0;
    ip0.k1 = ip0.k0 = ip0.tau = ip0.mu = 0.f;  // set per iteration
    ip0.lambda = A.lambda;
    ip0.alpha0 = A.alpha0;
    ip0.alpha1 = A.alpha1;
    ip0.first = (E.flags & EF_FIRST) != 0;
    ip0.write_tp = (E.flags & EF_WRITE_TP_LAST) != 0;  // the op's last iteration
    ip0.preload_tp = 0;
    ip0.deriv_level = (E.flags & EF_DERIV) ? (int)E.level : -1;
    PersistK pk;
    const int n = iterp_count(E.it0);
    pk.blocks = (n + K - 1) / K;
    pk.skip = pk.blocks * K - n;
    pk.tile = b;
    pk.ntiles = nb;
    pk.xs = t.d->xsync;
    pk.xbuf = t.d->xbuf;
    pk.s_word = s_word;
    if (E.masked) op_iterk<SH, K, false, true, T, true>(t, E.roles, sched, ip0, L, pk);
    else op_iterk<SH, K, false, false, T, true>(t, E.roles, sched, ip0, L, pk);
}

// ------------------------------------------------------------------ a whole warp per launch
//
// OP_ITERP runs an even number of inner iterations (normally all of a warp's) in one
// launch of persistent workgroups, one per tile, all resident at once.  A workgroup stages
// the tile grown by 2 pixels to the left / top and 1 to the right / bottom (K = 1 of the
// K-iteration op's geometry) once, and keeps it on chip: each work-item holds what only
// its positions read in registers, the values neighbours read in LDS.  Iteration j
// computes the duals over the tile grown by 1 to the left / top, then thresholding and
// the primal step over the tile.  The duals of the one-pixel ring are computed here from
// the ring's u_ and v_, p and q as the ring's own tile computes them (the same code, so
// the same values), so after each iteration the workgroups exchange only u_ and v_: each
// publishes those of its two outermost columns and rows to the exchange words, and
// before the next iteration each refreshes its ring from its neighbours' words.  A word
// is (value bits | tag << 32) with tag = (launch counter << 8) | iteration, written and
// read by agent-scope atomics, so a reader waits for the tag it needs without fences.
// Words are double-buffered by the parity of the iteration.  Besides the ring values every
// tile waits each iteration for a progress word of each of its (up to 8) neighbours, so that
// no tile runs more than one iteration ahead of a neighbour whatever the mask (a tile reads
// its upper-left neighbour's corner but not the other way round, and a ring outside the
// mask reads nothing): a tile writes iteration j + 2's words only after every neighbour has
// published j + 1, i.e. has read the tile's words of iteration j.
// Buffers, stores and out-of-mask values follow the K-iteration op with K = the op's
// iteration count: iteration n - 2 stores u, u_, v_ to the other buffers, iteration n - 1
// the whole state to the spare ones, and the host rotates the roles.  Every wait is
// bounded: a workgroup that finds not every workgroup resident, or waits too long for a
// neighbour, sets the pair's abort flag and leaves; the others see it and leave too, and
// the host runs the solve again without this op.
template <int SH>
struct BlkP {
    static constexpr int TW = kShape[SH].tw, TH = kShape[SH].th;
    static constexpr int RO = -2, RW = TW + 3, RH = TH + 3, NR = RW * RH;  // staged region
    // LDS rows of an odd pitch: a column of the region spans every bank
    static constexpr int PITCH = RW | 1, NL = RH * PITCH;
    static constexpr int NT = NR >= 1024 ? 1024 : (NR + 63) / 64 * 64;
    static constexpr int NK = (NR + NT - 1) / NT;  // staged positions per work-item
    static_assert(persist_shape(SH) && NK <= 3 && TW >= 3 && TH >= 3, "persistent tile shape");
};

template <int SH>
struct LdsP {
    static constexpr int NR = BlkP<SH>::NR;
    float ub[BlkP<SH>::NL];
    float2 v[BlkP<SH>::NL];
    float4 q[BlkP<SH>::NL];
    float2 tp[BlkP<SH>::NL];
    unsigned char mk[BlkP<SH>::NL];
};

// Exchange words of one tile and parity: u_, v_.x, v_.y (value-major) of column tw - 1
// (A), column tw - 2 (B), row th - 1 (C), row th - 2 (D), column 0 (E) and row 0 (F),
// then the tile's progress word (HB).  Every position of a tile's ring lies in one of its
// neighbours' strips.
template <int SH>
struct StripP {
    static constexpr int TW = kShape[SH].tw, TH = kShape[SH].th;
    static constexpr int A = 0, B = 3 * TH, C = 6 * TH, D = C + 3 * TW, E = D + 3 * TW, F = E + 3 * TH;
    static constexpr int HB = F + 3 * TW;
    static constexpr int WORDS = HB + 1;
    static_assert(WORDS == persist_words(SH), "strip layout");
};


template <int SH, bool MASKED, typename T>
__device__ __forceinline__ void op_iterp(const Tile &t, int roles, const float4 *__restrict__ sched,
                                         const IterParams &ip0, int n_iter, int tile, int ntiles,
                                         XSync *__restrict__ xs, unsigned long long *__restrict__ xbuf, LdsP<SH> &L,
                                         unsigned *s_word)
{
    using B = BlkP<SH>;
    using X = StripP<SH>;
    constexpr int NT = B::NT, NK = B::NK, TW = B::TW, TH = B::TH;
    const int w = t.w, h = t.h, s = t.s, x0 = t.x0, y0 = t.y0;
    const int tid = threadIdx.x;

    // ---- every workgroup of the launch must be resident
    persist_arrive(xs, ntiles, s_word);
    const unsigned epoch = s_word[0];

    if (s_word[1] == 0) {
        const float *__restrict__ mask = t.mask;
        const float *__restrict__ u_cur = slot(t, slot_u(rcur(roles, G_U)));
        const float *__restrict__ ub_cur = slot(t, slot_ub(rcur(roles, G_UB)));
        const float2 *__restrict__ v_cur = slotp<float2>(t, slot_v(rcur(roles, G_V)));
        const V2<T> *__restrict__ p_cur = slotp<V2<T>>(t, slot_p(rcur(roles, G_PQ)));
        const V4<T> *__restrict__ q_cur = slotp<V4<T>>(t, slot_q(rcur(roles, G_PQ)));
        const float *__restrict__ ub_oth = slot(t, slot_ub(roth(roles, G_UB)));
        const float2 *__restrict__ v_oth = slotp<float2>(t, slot_v(roth(roles, G_V)));
        const float2 *__restrict__ Tp = slotp<float2>(t, PL_TP);
        const T *__restrict__ at = (const T *)slot(t, PL_A);
        const T *__restrict__ bt = (const T *)slot(t, PL_B);
        const T *__restrict__ ct = (const T *)slot(t, PL_C);
        const T *__restrict__ Iu = (const T *)slot(t, PL_IU);
        const T *__restrict__ Iz = (const T *)slot(t, PL_IZ);

        // ---- the positions of this work-item: position q = tid + k NT in persist_rank's
        // order (the tile's border at k = 0, so that the work-items that compute it in the
        // primal step publish it), through a table in the LDS of q (staged over afterwards);
        // pi[k] = LDS index | column << 12 | row << 20 in the staged region
        static_assert(persist_border(SH) <= NT && B::NL <= 4096 && B::RW < 256 && X::WORDS < 2047, "positions");
        int pi[NK];
        {
            int *tab = reinterpret_cast<int *>(L.q);
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int i = tid + k * NT;
                if (i < B::NR) tab[persist_rank(SH, i)] = i;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int q = tid + k * NT;
                const int ri = q < B::NR ? tab[q] : 0, c = ri % B::RW, r = ri / B::RW;
                pi[k] = (r * B::PITCH + c) | c << 12 | r << 20;
            }
            __syncthreads();
        }
        auto staged = [&](int k) { return tid + k * NT < B::NR; };
        auto pos_i = [&](int k) { return pi[k] & 0xfff; };
        auto pos_x = [&](int k) { return x0 + B::RO + ((pi[k] >> 12) & 0xff); };
        auto pos_y = [&](int k) { return y0 + B::RO + (pi[k] >> 20); };

        // ---- the warp's derivatives (EF_DERIV), as the K-iteration op computes them
        if (ip0.deriv_level >= 0) {
#pragma unroll
            for (int k = 0; k < NK; k++) {
                if (!staged(k)) continue;
                const int i = pos_i(k), x = pos_x(k), y = pos_y(k);
                const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
                const bool own = xc == x && yc == y && x >= x0 && x < x0 + TW && y >= y0 && y < y0 + TH;
                L.tp[i] = staged_deriv<T>(t, ip0.deriv_level, xc, yc, own);
            }
        }

        // ---- stage (as the K-iteration op): own values into registers, shared ones into
        // LDS.  Outside the mask, q is never read, and its LDS slot holds the other buffers'
        // u_ and v_ instead (the values read there at odd iterations; the two buffers'
        // values trade places between the slots every iteration).
        OwnK o[NK];
        float omk[NK];
        float2 otp[NK];
        float4 oq[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int i = pos_i(k), x = pos_x(k), y = pos_y(k);
            if (staged(k)) {
                const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
                const long pos = (long)yc * s + xc;
                omk[k] = MASKED ? mask[pos] : 1.f;
                o[k].ub = ub_cur[pos];
                o[k].v = v_cur[pos];
                o[k].u = u_cur[pos];
                o[k].a = ldT(at, pos);
                o[k].b = ldT(bt, pos);
                o[k].c = ldT(ct, pos);
                if (ip0.deriv_level >= 0) {
                    const float2 d = L.tp[i];  // computed above
                    o[k].iu = d.x;
                    o[k].iz = d.y;
                } else {
                    o[k].iu = ldT(Iu, pos);
                    o[k].iz = ldT(Iz, pos);
                }
                if (ip0.first) {
                    o[k].p = make_float2(0.f, 0.f);
                    oq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                } else {
                    o[k].p = ld2T(p_cur, pos);
                    oq[k] = ld4T(q_cur, pos);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NK; k++) {
            otp[k] = make_float2(0.f, 0.f);
            if (MASKED && staged(k) && omk[k] == 0.f) {
                const int x = pos_x(k), y = pos_y(k);
                const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
                const long pos = (long)yc * s + xc;
                const float2 vo = v_oth[pos];
                oq[k] = make_float4(ub_oth[pos], vo.x, vo.y, 0.f);
                otp[k] = Tp[pos];
            }
        }
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int i = pos_i(k);
            if (staged(k)) {
                L.ub[i] = o[k].ub;
                L.v[i] = o[k].v;
                L.q[i] = oq[k];
                L.tp[i] = otp[k];
                L.mk[i] = omk[k] != 0.f ? 1 : 0;
            }
        }
        // row padding of the last tile column: the spare buffers take the current values
        if (x0 + TW >= w && s > w) {
            float *__restrict__ u_spr = slot(t, slot_u(rspare(roles, G_U)));
            float *__restrict__ ub_spr = slot(t, slot_ub(rspare(roles, G_UB)));
            float2 *__restrict__ v_spr = slotp<float2>(t, slot_v(rspare(roles, G_V)));
            const int pw = s - w, rows = min(TH, h - y0);
            for (int i = tid; i < pw * rows; i += NT) {
                const long pos = (long)(y0 + i / pw) * s + w + i % pw;
                u_spr[pos] = u_cur[pos];
                ub_spr[pos] = ub_cur[pos];
                v_spr[pos] = v_cur[pos];
            }
        }
        // Interior tiles (staged region inside the image and the mask / circle) run the
        // iterations without the boundary tests, except in 53x50 tiles, where every tile
        // takes the path with them: measured 4.38 against 4.53 us per iteration at
        // 848x800 (in the other shapes the split is the faster one).  Same values either way.
        constexpr bool kFullPath = SH != 18;
        bool inner = true;
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int x = pos_x(k), y = pos_y(k);
            if (staged(k))
                inner = inner && x >= 0 && y >= 0 && x < w && y < h && omk[k] != 0.f &&
                        (MASKED || !outside_circle(x, y, w, h));
        }
        const bool full = kFullPath ? __syncthreads_and(inner) != 0 : (__syncthreads(), false);
        PrimalCoef kc[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) kc[k] = primal_coef(ip0, o[k].a, o[k].b, o[k].c, o[k].iu);

        auto act = [&](int k) { return !MASKED || L.mk[k] != 0; };
        // Per own position, once per launch: the neighbour flags of the path with boundary
        // tests (as the K-iteration op) and the regions (duals: the tile grown by 1 to the
        // left / top; primal steps: the tile).
        enum : unsigned { F_IN = 1, F_ACT = 2, F_R = 4, F_D = 8, F_L = 16, F_U = 32, F_OUT = 64, F_DUAL = 128, F_TILE = 256 };
        unsigned fl[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int i = pos_i(k), x = pos_x(k), y = pos_y(k);
            const int dx = x - x0, dy = y - y0;
            fl[k] = 0;
            if (!staged(k)) continue;
            if (dx >= -1 && dy >= -1 && dx <= TW - 1 && dy <= TH - 1) fl[k] |= F_DUAL;
            if (dx >= 0 && dy >= 0 && dx <= TW - 1 && dy <= TH - 1) fl[k] |= F_TILE;
            if (!full && x >= 0 && y >= 0 && x < w && y < h) {
                const int col = (pi[k] >> 12) & 0xff, row = pi[k] >> 20;
                fl[k] |= F_IN | (act(i) ? F_ACT : 0u);
                if (x + 1 < w && col + 1 < B::RW && act(i + 1)) fl[k] |= F_R;
                if (y + 1 < h && row + 1 < B::RH && act(i + B::PITCH)) fl[k] |= F_D;
                if (x >= 1 && col >= 1 && act(i - 1)) fl[k] |= F_L;
                if (y >= 1 && row >= 1 && act(i - B::PITCH)) fl[k] |= F_U;
                if (!MASKED && outside_circle(x, y, w, h)) fl[k] |= F_OUT;
            }
        }
        // Exchange slots, once per launch.  The primal step that computes a position of the
        // tile's border (this work-item's k = 0 position when tid < persist_border) publishes
        // its u_ and v_ into every strip holding it, where it lies in the image and is
        // processed: pub_s = column strip word + 1 | row strip word + 1 << 11 (0: none).
        // Work-item `tid` refreshes ring position tid (those some dual reads, 3 (tw + th) + 5:
        // columns -2, -1 and tw, rows -2, -1 and th of the staged region), where that position
        // lies in the image and is processed; the next 8 work-items wait for the 8 neighbours'
        // progress words.  rcv_i: LDS index (-2: a progress word); rcv_w: the source tile's
        // word (parity 0) | 1 << 30 for a column strip (value stride th; rows: tw).
        constexpr int RS = 3 * (TW + TH) + 5;
        static_assert(B::NL <= 4096 && X::WORDS < (1 << 20) && RS + 8 <= NT, "exchange slots");
        const int ntx = (w + TW - 1) / TW, nty = (h + TH - 1) / TH;
        int pub_s = 0, rcv_i = -1, rcv_w = 0;
        if (tid < persist_border(SH) && pos_x(0) < w && pos_y(0) < h && act(pos_i(0))) {
            const int dx = pos_x(0) - x0, dy = pos_y(0) - y0;
            const int co = dx == TW - 1 ? X::A + dy : dx == TW - 2 ? X::B + dy : dx == 0 ? X::E + dy : -1;
            const int ro = dy == TH - 1 ? X::C + dx : dy == TH - 2 ? X::D + dx : dy == 0 ? X::F + dx : -1;
            pub_s = (co + 1) | ((ro + 1) << 11);
        }
        if (tid < 3 * (TW + TH) + 5) {
            int dx, dy;
            if (tid < TH + 1) dx = -2, dy = tid - 1;
            else if (tid < 2 * TH + 4) dx = -1, dy = tid - (TH + 1) - 2;
            else if (tid < 3 * TH + 5) dx = TW, dy = tid - (2 * TH + 4) - 1;
            else if (tid < 3 * TH + 5 + TW) dx = tid - (3 * TH + 5), dy = -2;
            else if (tid < 3 * TH + 5 + 2 * TW) dx = tid - (3 * TH + 5 + TW), dy = -1;
            else dx = tid - (3 * TH + 5 + 2 * TW), dy = TH;
            const int x = x0 + dx, y = y0 + dy, i = (dy - B::RO) * B::PITCH + (dx - B::RO);
            if (x >= 0 && y >= 0 && x < w && y < h && act(i)) {
                // its tile's strip holding it
                const int sx = x / TW, sy = y / TH, lx = x - sx * TW, ly = y - sy * TH;
                int off, col;
                if (lx == TW - 1) off = X::A + ly, col = 1;
                else if (lx == TW - 2) off = X::B + ly, col = 1;
                else if (ly == TH - 1) off = X::C + lx, col = 0;
                else if (ly == TH - 2) off = X::D + lx, col = 0;
                else if (lx == 0) off = X::E + ly, col = 1;
                else off = X::F + lx, col = 0;  // ly == 0
                rcv_i = i;
                rcv_w = ((sy * ntx + sx) * 2 * X::WORDS + off) | (col << 30);
            }
        } else if (tid < RS + 8) {
            const int d = tid - RS, ddx = d < 3 ? d - 1 : d < 5 ? (d == 3 ? -1 : 1) : d - 6;
            const int ddy = d < 3 ? -1 : d < 5 ? 0 : 1;
            const int nx = x0 / TW + ddx, ny = y0 / TH + ddy;
            if (nx >= 0 && ny >= 0 && nx < ntx && ny < nty) {
                rcv_i = -2;
                rcv_w = (ny * ntx + nx) * 2 * X::WORDS + X::HB;
            }
        }
        unsigned long long *__restrict__ mine = xbuf + (long)tile * 2 * X::WORDS;
        const unsigned tag0 = epoch << 8;
        // this work-item's exchange words, loaded at the end of each iteration
        const unsigned long long *rw = xbuf + (long)(rcv_w & 0x3fffffff);
        const int rst = (rcv_w >> 30) ? TH : TW;

        // One iteration.  MODE 0: an iteration before the op's last two; 1: iteration
        // n - 2 (u, u_, v_ to the other buffers); 2: the last one (the whole state to the
        // spare buffers).  Every iteration but the last ends by refreshing the ring for the
        // next one.  Returns false when a wait gave up.
        auto step = [&](int j, auto full_c, auto mode_c) -> bool {
            constexpr bool FULL = decltype(full_c)::value;
            constexpr int MODE = decltype(mode_c)::value;
            IterParams ip = ip0;
            const float4 sc = sched[j];
            ip.k1 = sc.w;
            ip.k0 = ip0.alpha0 * sc.x / 2.0f;
            ip.tau = sc.y;
            ip.mu = sc.z;
            ip.first = j == 0 && ip0.first;
            ip.write_tp = MODE == 2 && ip0.write_tp;

            // ---- duals over the tile grown by 1 to the left / top (u_ and v_ of every
            // position from LDS, where the ring's are refreshed)
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int i = pos_i(k);
                if (!(fl[k] & F_DUAL)) continue;
                if (!FULL && !(fl[k] & F_ACT && fl[k] & F_IN)) continue;
                const int x = pos_x(k), y = pos_y(k);
                const int kR = (FULL || x + 1 < w) ? i + 1 : i, kD = (FULL || y + 1 < h) ? i + B::PITCH : i;
                DualIn rr;
                rr.ok = true;
                rr.act = true;
                rr.hasR = FULL || (fl[k] & F_R) != 0;
                rr.hasD = FULL || (fl[k] & F_D) != 0;
                rr.a = o[k].a;
                rr.b = o[k].b;
                rr.c = o[k].c;
                rr.u0 = L.ub[i];
                rr.uR = L.ub[kR];
                rr.uD = L.ub[kD];
                rr.uL = (!FULL && x >= 1) ? L.ub[i - 1] : 0.f;
                rr.uU = (!FULL && y >= 1) ? L.ub[i - B::PITCH] : 0.f;
                rr.v0 = L.v[i];
                rr.vR = L.v[kR];
                rr.vD = L.v[kD];
                rr.vL = (!FULL && x >= 1) ? L.v[i - 1] : make_float2(0.f, 0.f);
                rr.vU = (!FULL && y >= 1) ? L.v[i - B::PITCH] : make_float2(0.f, 0.f);
                rr.po = make_float2(rT<T>(o[k].p.x), rT<T>(o[k].p.y));
                const float4 qo = L.q[i];
                rr.qo = make_float4(rT<T>(qo.x), rT<T>(qo.y), rT<T>(qo.z), rT<T>(qo.w));
                float2 pn, tpn;
                float4 qn;
                dual_step<MASKED>(rr, !FULL && (fl[k] & F_OUT) != 0, ip, pn, qn, tpn);
                o[k].p = pn;
                L.q[i] = qn;
                L.tp[i] = tpn;
                if (MODE == 2 && (fl[k] & F_TILE)) {
                    const long pos = (long)y * s + x;
                    st2T(slotp<V2<T>>(t, slot_p(rspare(roles, G_PQ))), pos, pn);
                    st4T(slotp<V4<T>>(t, slot_q(rspare(roles, G_PQ))), pos, qn);
                    if (ip.write_tp) slotp<float2>(t, PL_TP)[pos] = tpn;
                }
            }
            __syncthreads();

            // ---- thresholding + primal step over the tile; outside the mask, the stored
            // u_ and v_ switch to the buffer the next iteration reads there
#pragma unroll
            for (int k = 0; k < NK; k++) {
                const int i = pos_i(k);
                if (FULL && !(fl[k] & F_TILE)) continue;
                if (!FULL && !(fl[k] & F_IN)) continue;
                const int x = pos_x(k), y = pos_y(k);
                const long pos = (long)y * s + x;
                if (!FULL && MASKED && !(fl[k] & F_ACT)) {
                    // the two buffers' values trade places between the u_ / v_ slots (read
                    // by neighbours) and the q slot: the current buffers' after odd
                    // iterations, the other buffers' after even ones; the last iteration
                    // finds the current buffers' values in the q slot
                    const float4 sw = L.q[i];
                    if (MODE == 2) {
                        if (fl[k] & F_TILE) {
                            slot(t, slot_u(rspare(roles, G_U)))[pos] = o[k].u;
                            slot(t, slot_ub(rspare(roles, G_UB)))[pos] = sw.x;
                            slotp<float2>(t, slot_v(rspare(roles, G_V)))[pos] = make_float2(sw.y, sw.z);
                        }
                    } else {
                        const float2 vv = L.v[i];
                        L.q[i] = make_float4(L.ub[i], vv.x, vv.y, 0.f);
                        L.ub[i] = sw.x;
                        L.v[i] = make_float2(sw.y, sw.z);
                    }
                    continue;
                }
                if (!(fl[k] & F_TILE)) continue;
                float un, u_s;
                float2 vs;
                if (!FULL && !MASKED && (fl[k] & F_OUT)) {
                    un = 0.f;
                    u_s = 0.f;
                    vs = make_float2(0.f, 0.f);
                } else {
                    const bool hR = FULL || (fl[k] & F_R) != 0, hD = FULL || (fl[k] & F_D) != 0;
                    const bool hL = FULL || (fl[k] & F_L) != 0, hU = FULL || (fl[k] & F_U) != 0;
                    const float tpLx = (FULL || x >= 1) ? L.tp[i - 1].x : 0.f;
                    const float tpUy = (FULL || y >= 1) ? L.tp[i - B::PITCH].y : 0.f;
                    const float2 tp0 = L.tp[i];
                    const float divp = div_tp<MASKED>(x, y, w, h, hR, hD, hL, hU, tp0, tpLx, tpUy);
                    const float4 ql = hL ? L.q[i - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const float4 qu = hU ? L.q[i - B::PITCH] : make_float4(0.f, 0.f, 0.f, 0.f);
                    primal_core(ip, kc[k], L.ub[i], o[k].u, o[k].iu, o[k].iz, L.v[i], hR, hD, hL, hU, divp, L.q[i], ql,
                                qu, tp0, un, u_s, vs);
                }
                o[k].u = un;
                L.ub[i] = u_s;
                L.v[i] = vs;
                if (MODE != 2 && k == 0 && pub_s != 0) {
                    // a border position: its strip words of this iteration
                    const unsigned tg = tag0 | (unsigned)j;
                    unsigned long long *pw = mine + (j & 1) * X::WORDS;
                    const int co = (pub_s & 2047) - 1, ro = (pub_s >> 11) - 1;
                    if (co >= 0) {
                        xstore(pw + co, u_s, tg);
                        xstore(pw + co + TH, vs.x, tg);
                        xstore(pw + co + 2 * TH, vs.y, tg);
                    }
                    if (ro >= 0) {
                        xstore(pw + ro, u_s, tg);
                        xstore(pw + ro + TW, vs.x, tg);
                        xstore(pw + ro + 2 * TW, vs.y, tg);
                    }
                }
                if (MODE == 1) {
                    slot(t, slot_u(roth(roles, G_U)))[pos] = un;
                    slot(t, slot_ub(roth(roles, G_UB)))[pos] = u_s;
                    slotp<float2>(t, slot_v(roth(roles, G_V)))[pos] = vs;
                } else if (MODE == 2) {
                    slot(t, slot_u(rspare(roles, G_U)))[pos] = un;
                    slot(t, slot_ub(rspare(roles, G_UB)))[pos] = u_s;
                    slotp<float2>(t, slot_v(rspare(roles, G_V)))[pos] = vs;
                }
            }
            if constexpr (MODE != 2) {
                // ---- refresh the ring from the neighbours' strips of this iteration (their
                // primal steps publish them as this one's does); the ring positions are
                // not read between the barriers around the primal steps.  The barrier of
                // persist_any then also serves the next iteration's duals, which read this
                // one's u_ and v_ from LDS.
                int bad = 0;
                const unsigned tg = tag0 | (unsigned)j;
                const unsigned long long *pw = rw + (j & 1) * X::WORDS;
                if (rcv_i == -2) {
                    // a neighbour's progress word
                    if (!persist_wait_tag(pw, tg, xs)) bad = 3;
                } else if (rcv_i >= 0) {
                    unsigned long long a0, a1, a2, t0 = 0;
                    if constexpr (NK == 1) __builtin_amdgcn_s_sleep(kRingDelay);
                    for (int polls = 0;; polls++) {
                        a0 = xload(pw);
                        a1 = xload(pw + rst);
                        a2 = xload(pw + 2 * rst);
                        if ((unsigned)(a0 >> 32) == tg && (unsigned)(a1 >> 32) == tg && (unsigned)(a2 >> 32) == tg)
                            break;
                        if (polls == 0) t0 = rtime();
                        if ((polls & 63) == 63 && (rtime() - t0 > kPollTicks || xload32(&xs->abort) != 0)) {
                            bad = 2;
                            break;
                        }
                    }
                    L.ub[rcv_i] = __uint_as_float((unsigned)a0);
                    L.v[rcv_i] = make_float2(__uint_as_float((unsigned)a1), __uint_as_float((unsigned)a2));
                }
                if (persist_any(bad, s_word)) {
                    persist_give_up(xs, ntiles, bad, j + 1);
                    return false;
                }
                // progress: this tile has read its neighbours' words of iteration j, which
                // their primal steps of iteration j + 2 overwrite
                if (tid == NT - 1) xstore(mine + ((j + 1) & 1) * X::WORDS + X::HB, 0.f, tag0 | (unsigned)(j + 1));
            }
            return true;
        };
        auto run = [&](auto full_c) {
            using I0 = std::integral_constant<int, 0>;
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            for (int j = 0; j < n_iter - 2; j++)
                if (!step(j, full_c, I0{})) return;
            if (step(n_iter - 2, full_c, I1{})) step(n_iter - 1, full_c, I2{});
        };
        // progress of iteration 0: no words read yet (the staging covered the ring)
        if (tid == NT - 1) xstore(mine + X::HB, 0.f, tag0);
        if constexpr (kFullPath) {
            if (full) run(std::true_type{});
            else run(std::false_type{});
        } else {
            run(std::false_type{});
        }
    }

    persist_finish(xs, ntiles, epoch);
}

template <typename T, int SH>
__global__ __launch_bounds__(BlkP<SH>::NT) void k_iterp(const PairDesc *__restrict__ D, const TickArgs A)
{
    using B = BlkP<SH>;
    __shared__ LdsP<SH> L;
    __shared__ unsigned s_word[3];
    // tiles in XCD-contiguous runs (bands of rows), as the other op kernels
    const int nb = gridDim.x, i = blockIdx.x, xcd = i & 7, q = nb >> 3, r = nb & 7;
    const int b = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + (i >> 3);
    const TickEntry &E = A.e[0];
    Tile t;
    t.d = D + E.pair;
    t.arena = E.arena;
    t.plane = E.plane;
    t.mask = E.mask;
    t.w = E.w;
    t.h = E.h;
    t.s = E.s;
    const int ntx = (t.w + B::TW - 1) / B::TW;
    t.x0 = (b % ntx) * B::TW;
    t.y0 = (b / ntx) * B::TH;
    const float4 *sched = t.d->sched + iterp_first(E.it0);
    IterParams ip0;
    ip0.k1 = ip0.k0 = ip0.tau = ip0.mu = 0.f;  // set per iteration
    ip0.lambda = A.lambda;
    ip0.alpha0 = A.alpha0;
    ip0.alpha1 = A.alpha1;
    ip0.first = (E.flags & EF_FIRST) != 0;
    ip0.write_tp = (E.flags & EF_WRITE_TP_LAST) != 0;
    ip0.preload_tp = 0;
    ip0.deriv_level = (E.flags & EF_DERIV) ? (int)E.level : -1;
    XSync *xs = t.d->xsync;
    unsigned long long *xbuf = t.d->xbuf;
    if (E.masked) op_iterp<SH, true, T>(t, E.roles, sched, ip0, iterp_count(E.it0), b, nb, xs, xbuf, L, s_word);
    else op_iterp<SH, false, T>(t, E.roles, sched, ip0, iterp_count(E.it0), b, nb, xs, xbuf, L, s_word);
}

// 13th smallest of 25 (the value medianFilter.cu:307-316's partial selection sort leaves
// at position 12) by a comparator network: Batcher's odd-even merge sort of 32 inputs,
// the 7 beyond the window taken as +infinity, pruned to the 113 comparators that
// position 12 depends on (each a min and a max; the compiler drops the unused ones).  A
// network selects the k-th smallest of all inputs if it does so for every 0/1 input (the
// 0-1 principle); tests/test_median_network.py checks all 2^25 of them.  With NaN in the
// window the result may differ from the reference's selection sort (fminf / fmaxf drop
// NaN operands).
constexpr unsigned char kMedian25[113][2] = {
    {0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}, {4, 5}, {6, 7}, {4, 6},
    {5, 7}, {5, 6}, {0, 4}, {2, 6}, {2, 4}, {1, 5}, {3, 7}, {3, 5},
    {1, 2}, {3, 4}, {5, 6}, {8, 9}, {10, 11}, {8, 10}, {9, 11}, {9, 10},
    {12, 13}, {14, 15}, {12, 14}, {13, 15}, {13, 14}, {8, 12}, {10, 14}, {10, 12},
    {9, 13}, {11, 15}, {11, 13}, {9, 10}, {11, 12}, {13, 14}, {0, 8}, {4, 12},
    {4, 8}, {2, 10}, {6, 14}, {6, 10}, {2, 4}, {6, 8}, {10, 12}, {1, 9},
    {5, 13}, {5, 9}, {3, 11}, {7, 15}, {7, 11}, {3, 5}, {7, 9}, {11, 13},
    {1, 2}, {3, 4}, {5, 6}, {7, 8}, {9, 10}, {11, 12}, {13, 14}, {16, 17},
    {18, 19}, {16, 18}, {17, 19}, {17, 18}, {20, 21}, {22, 23}, {20, 22}, {21, 23},
    {21, 22}, {16, 20}, {18, 22}, {18, 20}, {17, 21}, {19, 23}, {19, 21}, {17, 18},
    {19, 20}, {21, 22}, {16, 24}, {20, 24}, {18, 20}, {22, 24}, {19, 21}, {17, 18},
    {19, 20}, {21, 22}, {23, 24}, {0, 16}, {8, 24}, {8, 16}, {4, 20}, {12, 20},
    {12, 16}, {2, 18}, {10, 18}, {6, 22}, {6, 10}, {10, 12}, {1, 17}, {9, 17},
    {5, 21}, {13, 21}, {13, 17}, {3, 19}, {11, 19}, {7, 23}, {7, 11}, {11, 13},
    {11, 12},
};

__device__ __forceinline__ float median25(float m[25])
{
#pragma unroll
    for (int i = 0; i < (int)(sizeof kMedian25 / sizeof kMedian25[0]); i++) {
        const float a = m[kMedian25[i][0]], b = m[kMedian25[i][1]];
        m[kMedian25[i][0]] = fminf(a, b);
        m[kMedian25[i][1]] = fmaxf(a, b);
    }
    return m[12];
}

// 5x5 window as medianFilter.cu:282-304: edge replicate when exactly one of row / column
// is outside the image, 0 when both are.
__device__ __forceinline__ float median5_at(const float *__restrict__ u, int x, int y, int w, int h, int s)
{
    float m[25];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int row = y + j - 2;
        const bool rin = row >= 0 && row < h;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const int col = x + i - 2;
            const bool cin = col >= 0 && col < w;
            float v = 0.f;
            if (rin || cin) {
                const int rr = rin ? row : (row < 0 ? 0 : h - 1);
                const int cc = cin ? col : (col < 0 ? 0 : w - 1);
                v = u[(size_t)rr * s + cc];
            }
            m[j * 5 + i] = v;
        }
    }
    return median25(m);
}

// End of a warp (stereotgv.cu:481-493 masked, :354-363 unmasked): [median of u
// (medianFilter.cu:264-319) into the other u slot,] du = min(u - u_last, limit)
// (scalar.cu:129-141, limitRange.cu:3-24), u = u_last + du (scalar.cu:77-89),
// u_ = u (copyClone.cu:3-12), warpUV += du * tv2 (warpImage.cu:136-150, scalar.cu:102-115).
template <class S, bool MASKED>
__device__ __forceinline__ void op_end(const Tile &t, int roles, float limit)
{
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *out = MASKED ? slot(t, slot_u(roth(roles, G_U))) : (float *)u;
    float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    const float *ulast = slot(t, PL_ULAST);
    const float2 *tv2 = slotp<float2>(t, PL_TV2);
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        const float uval = MASKED ? median5_at(u, x, y, t.w, t.h, t.s) : u[pos];
        const float ul = ulast[pos];
        float du = uval - ul;
        du = (du > limit) ? limit : du;
        const float un = ul + du;
        out[pos] = un;
        u_[pos] = un;
        const float2 tt = tv2[pos];
        const float2 o = uv[pos];
        uv[pos] = make_float2(o.x + du * tt.x, o.y + du * tt.y);
    }
}

// The end of one warp and the preparation of the next at the same level in one op: each
// pixel's end-of-warp values (op_end) are all the next warp's preparation (op_prep) reads
// at that pixel, so the two run per pixel without a launch between them.  Roles as for
// op_end; the preparation's u is the end's result.
template <class S, bool MASKED, typename T>
__device__ __forceinline__ void op_end_prep(const Tile &t, int level, int roles, float limit)
{
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *out = MASKED ? slot(t, slot_u(roth(roles, G_U))) : (float *)u;
    float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    float *ulast = slot(t, PL_ULAST);
    float2 *tv2 = slotp<float2>(t, PL_TV2);
    float2 *v_ = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    const float2 *tv = t.d->tv[level];
    const T *I1 = (const T *)t.d->I1[level];
    T *i1w = (T *)slot(t, PL_I1W);
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const long pos = (long)y * t.s + x;
        // op_end
        const float uval = MASKED ? median5_at(u, x, y, t.w, t.h, t.s) : u[pos];
        const float ul = ulast[pos];
        float du = uval - ul;
        du = (du > limit) ? limit : du;
        const float un = ul + du;
        out[pos] = un;
        u_[pos] = un;
        const float2 tt = tv2[pos];
        const float2 o = uv[pos];
        const float2 d = make_float2(o.x + du * tt.x, o.y + du * tt.y);
        uv[pos] = d;
        // op_prep of the next warp
        const float px = (float)x + d.x, py = (float)y + d.y;
        tv2[pos] = bil2(tv, t.w, t.h, t.s, px, py);
        stT(i1w, pos, bil(I1, t.w, t.h, t.s, px, py));
        ulast[pos] = un;
        v_[pos] = make_float2(0.f, 0.f);
    }
}

// upscale.cu:7-21, 46-61 for u, u_ and warpUV (stereotgv.cu:497-510): bilinear at the
// finer level's pixel centres, times fScale; writes the other slot of each pair.
template <class S>
__device__ __forceinline__ void op_up(const Tile &t, int level, int roles, float scale)
{  // the tile belongs to the finer level level - 1
    const int w = t.d->w[level], h = t.d->h[level], s = t.d->s[level];
    const float *u = slot(t, slot_u(rcur(roles, G_U)));
    float *ou = slot(t, slot_u(roth(roles, G_U)));
    const float *u_ = slot(t, slot_ub(rcur(roles, G_UB)));
    float *ou_ = slot(t, slot_ub(roth(roles, G_UB)));
    const float2 *uv = slotp<float2>(t, slot_uv(roles, true));
    float2 *ouv = slotp<float2>(t, slot_uv(roles, false));
    const float rx = (float)w / (float)t.w, ry = (float)h / (float)t.h;
    FOR_TILE_PIXELS(S, t, x, y)
    {
        const float px = ((float)x + 0.5f) * rx - 0.5f, py = ((float)y + 0.5f) * ry - 0.5f;
        const long pos = (long)y * t.s + x;
        ou[pos] = bil(u, w, h, s, px, py) * scale;
        ou_[pos] = bil(u_, w, h, s, px, py) * scale;
        const float2 v = bil2(uv, w, h, s, px, py);
        ouv[pos] = make_float2(v.x * scale, v.y * scale);
    }
}


// ------------------------------------------------------------------ energy monitor
//
// FP64 sums of the TGV-L1 energy and of the last iteration's change of u over the
// processed pixels of a pair (no reference counterpart; the monitor of the solve).  Every
// per-pixel term is evaluated in double from the stored values; the gradients are the dual
// step's (updateDualVariables.cu:152-212: forward differences, backward where the right /
// lower neighbour is outside the mask or the image).  The sums run in a fixed order: each
// OP_ENERGY workgroup takes kEnergyChunk consecutive pixels of the level in row-major
// order (4 per work-item, then a butterfly over the wave and the 4 waves in order), and
// OP_EFINAL adds the partials in index order, so the result does not depend on the launch
// shape or on the other pairs of a launch.

__device__ __forceinline__ void add_part(EnergyPart &a, const EnergyPart &b)
{
    a.energy += b.energy;
    a.data += b.data;
    a.residual += b.residual;
    a.nonfinite += b.nonfinite;
    a.pixels += b.pixels;
}

// Reduces each work-item's partial over the 256 work-items of a tick workgroup; the total
// is valid in work-item 0.
__device__ __forceinline__ EnergyPart reduce_part(EnergyPart p, int tid)
{
    __shared__ EnergyPart waves[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        p.energy += __shfl_xor(p.energy, o);
        p.data += __shfl_xor(p.data, o);
        p.residual += __shfl_xor(p.residual, o);
        p.nonfinite += __shfl_xor(p.nonfinite, o);
        p.pixels += __shfl_xor(p.pixels, o);
    }
    if ((tid & 63) == 0) waves[tid >> 6] = p;
    __syncthreads();
    EnergyPart r = waves[0];
#pragma unroll
    for (int k = 1; k < NT / 64; k++) add_part(r, waves[k]);
    return r;
}

// One chunk of the level: roles as after the warp's last inner iteration (u and v_ the
// current buffers, u_prev the other u buffer).
template <class S, bool MASKED, typename T>
__device__ __forceinline__ void op_energy(const Tile &t, int roles, int chunk, float lambda, float alpha0,
                                          float alpha1)
{
    const int tid = threadIdx.y * S::BX + threadIdx.x;
    const float *__restrict__ u = slot(t, slot_u(rcur(roles, G_U)));
    const float *__restrict__ up = slot(t, slot_u(roth(roles, G_U)));
    const float2 *__restrict__ v = slotp<float2>(t, slot_v(rcur(roles, G_V)));
    const float *__restrict__ ul = slot(t, PL_ULAST);
    const T *__restrict__ Iu = (const T *)slot(t, PL_IU);
    const T *__restrict__ Iz = (const T *)slot(t, PL_IZ);
    const T *__restrict__ at = (const T *)slot(t, PL_A);
    const T *__restrict__ bt = (const T *)slot(t, PL_B);
    const T *__restrict__ ct = (const T *)slot(t, PL_C);
    const float *__restrict__ mask = t.mask;
    const int w = t.w, h = t.h, s = t.s;
    const long n = (long)w * h;
    EnergyPart p{0.0, 0.0, 0.0, 0u, 0u};
#pragma unroll
    for (int k = 0; k < kEnergyChunk / NT; k++) {
        const long i = (long)chunk * kEnergyChunk + k * NT + tid;
        if (i >= n) break;
        const int y = (int)(i / w), x = (int)(i - (long)y * w);
        const long pos = (long)y * s + x;
        bool hasR, hasD;
        if (MASKED) {
            if (mask[pos] == 0.f) continue;
            hasR = x + 1 < w && mask[pos + 1] != 0.f;
            hasD = y + 1 < h && mask[pos + s] != 0.f;
        } else {
            if (outside_circle(x, y, w, h)) continue;
            hasR = x + 1 < w;
            hasD = y + 1 < h;
        }
        const double u0 = u[pos];
        const float2 v0 = v[pos];
        double ux, uy, g0, g1, g2, g3;
        if (hasR) {
            const float2 vr = v[pos + 1];
            ux = ((double)u[pos + 1] - u0) - (double)v0.x;
            g0 = (double)vr.x - (double)v0.x;
            g3 = (double)vr.y - (double)v0.y;
        } else {
            const double uL = pos >= 1 ? (double)u[pos - 1] : 0.0;
            const float2 vl = pos >= 1 ? v[pos - 1] : make_float2(0.f, 0.f);
            ux = (u0 - uL) - (double)v0.x;
            g0 = (double)v0.x - (double)vl.x;
            g3 = (double)v0.y - (double)vl.y;
        }
        if (hasD) {
            const float2 vd = v[pos + s];
            uy = ((double)u[pos + s] - u0) - (double)v0.y;
            g1 = (double)vd.y - (double)v0.y;
            g2 = (double)vd.x - (double)v0.x;
        } else {
            const double uU = pos >= s ? (double)u[pos - s] : 0.0;
            const float2 vu = pos >= s ? v[pos - s] : make_float2(0.f, 0.f);
            uy = (u0 - uU) - (double)v0.y;
            g1 = (double)v0.y - (double)vu.y;
            g2 = (double)v0.x - (double)vu.x;
        }
        const double a = ldT(at, pos), b = ldT(bt, pos), c = ldT(ct, pos);
        const double tx = a * ux + c * uy, ty = c * ux + b * uy;
        const double rho = (double)ldT(Iu, pos) * (u0 - (double)ul[pos]) + (double)ldT(Iz, pos);
        const double data = (double)lambda * fabs(rho);
        const double term = data + (double)alpha1 * sqrt(tx * tx + ty * ty) +
                            (double)alpha0 * sqrt(g0 * g0 + g1 * g1 + g2 * g2 + g3 * g3);
        const double res = fabs(u0 - (double)up[pos]);
        p.pixels++;
        if (__builtin_isfinite(term) && __builtin_isfinite(res)) {
            p.energy += term;
            p.data += data;
            p.residual += res;
        } else {
            p.nonfinite++;
        }
    }
    const EnergyPart r = reduce_part(p, tid);
    if (tid == 0) t.d->mon_part[chunk] = r;
}

// The partials of the level's OP_ENERGY, added in index order into record (level, warp).
template <class S>
__device__ __forceinline__ void op_efinal(const Tile &t, int level, int warp)
{
    const int tid = threadIdx.y * S::BX + threadIdx.x;
    const int n = energy_chunks(t.w, t.h);
    const EnergyPart *__restrict__ part = t.d->mon_part;
    EnergyPart p{0.0, 0.0, 0.0, 0u, 0u};
    for (int j = tid; j < n; j += NT) add_part(p, part[j]);
    const EnergyPart r = reduce_part(p, tid);
    if (tid == 0) t.d->mon_rec[mon_slot(level, warp)] = r;
}

// ------------------------------------------------------------------ the tick kernel
//
// One launch executes one op for each entry of `A`: entry i owns workgroups
// [A.e[i].tile_start, A.e[i+1].tile_start).  A batched solve passes the same op for every
// pair; the serving loop passes each active request's next op, so requests at different
// points of their schedules share launches.  Each pair's result depends only on its own
// entries, never on what runs beside it.
template <typename T, int SH>
__global__ __launch_bounds__(NT) void k_tick(const PairDesc *__restrict__ D, const TickArgs A)
{
    using S = Sh<SH>;
    __shared__ float2 sP[S::TH + 1][S::TW + 1];
    __shared__ float4 sQ[S::TH + 1][S::TW + 1];
    __shared__ float2 sT[S::TH + 1][S::TW + 1];
    // Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each XCD
    // owns one contiguous run of tiles (whole bands of rows), keeping the halo rows that
    // neighbouring tiles share in the same L2.
    const int nb = gridDim.x, i = blockIdx.x, xcd = i & 7, q = nb >> 3, r = nb & 7;
    const int b = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + (i >> 3);
    // Entry of this workgroup: the last one whose first tile is <= b.  Each lane of the
    // wave reads one entry's tile_start (kMaxEntries <= 64), so the lookup costs one load.
    int lo = 0;
    if (A.n > 1) {
        const int lane = __lane_id();
        const int ts = lane < A.n ? A.e[lane].tile_start : 0x7fffffff;
        lo = __popcll(__ballot(ts <= b)) - 1;
        lo = __builtin_amdgcn_readfirstlane(lo);
    }
    const TickEntry &E = A.e[lo];
    const int kind = E.kind, level = E.level;
    Tile t;
    t.d = D + E.pair;
    t.arena = E.arena;
    t.plane = E.plane;
    t.mask = E.mask;
    t.w = E.w;
    t.h = E.h;
    t.s = E.s;
    const int local = b - E.tile_start;
    const int ntx = (t.w + S::TW - 1) / S::TW;
    t.x0 = (local % ntx) * S::TW;
    t.y0 = (local / ntx) * S::TH;
    // workgroups beyond the op's tiles: the energy record's final sum of a warp end
    // (EF_EFINAL, one workgroup) or the spare buffers' copy of the derivatives' op (EF_SYNC)
    if ((E.flags & (EF_EFINAL | EF_SYNC)) && (kind == OP_END || kind == OP_ENDPREP || kind == OP_DERIV)) {
        const int ntiles = ntx * ((t.h + S::TH - 1) / S::TH);
        if (local >= ntiles) {
            if (kind == OP_DERIV) op_sync_chunk<S>(t, level, E.roles, local - ntiles);
            else op_efinal<S>(t, level, E.it0);
            return;
        }
    }
    switch (kind) {
    case OP_ZERO:
        op_zero<S>(t);
        break;
    case OP_CALIB:
        op_calib<S, T>(t);
        break;
    case OP_DOWN:
        op_down<S, T>(t, level);
        break;
    case OP_CINIT:
        op_cinit<S>(t, level, E.roles);
        break;
    case OP_WINIT:
        op_winit<S>(t, level, E.roles, A.scale);
        break;
    case OP_TENSOR:
        if (E.masked) op_tensor<S, T, true>(t, level, A.xblock, A.beta, A.gamma);
        else op_tensor<S, T, false>(t, level, A.xblock, A.beta, A.gamma);
        if (E.flags & EF_PREP) op_prep<S, T>(t, level, E.roles);  // disjoint reads and writes
        break;
    case OP_PREP:
        op_prep<S, T>(t, level, E.roles);
        break;
    case OP_DERIV:
        op_deriv<S, T>(t, level);
        break;
    case OP_ITER: {
        IterParams ip;
        ip.k1 = A.alpha1 * E.sigma / 3.0f;  // eta_p = 3 (stereotgv.cu:416)
        ip.k0 = A.alpha0 * E.sigma / 2.0f;  // eta_q = 2 (stereotgv.cu:417)
        ip.tau = E.tau;
        ip.mu = E.mu;
        ip.lambda = A.lambda;
        ip.alpha0 = A.alpha0;
        ip.alpha1 = A.alpha1;
        ip.first = E.flags & EF_FIRST;
        ip.write_tp = (E.flags & EF_WRITE_TP) != 0;
        ip.preload_tp = (E.flags & EF_PRELOAD_TP) != 0;
        ip.deriv_level = -1;
        if (E.masked) op_iter<S, true, T>(t, E.roles, ip, sP, sQ, sT);
        else op_iter<S, false, T>(t, E.roles, ip, sP, sQ, sT);
        break;
    }
    case OP_END:
        if (E.masked) op_end<S, true>(t, E.roles, A.limit);
        else op_end<S, false>(t, E.roles, A.limit);
        break;
    case OP_ENDPREP:
        if (E.masked) op_end_prep<S, true, T>(t, level, E.roles, A.limit);
        else op_end_prep<S, false, T>(t, level, E.roles, A.limit);
        break;
    case OP_UP:
        op_up<S>(t, level, E.roles, A.fscale);
        break;
    case OP_SYNC:
        op_sync<S>(t, level, E.roles);
        break;
    case OP_ENERGY:
        if (E.masked) op_energy<S, true, T>(t, E.roles, local, A.lambda, A.alpha0, A.alpha1);
        else op_energy<S, false, T>(t, E.roles, local, A.lambda, A.alpha0, A.alpha1);
        break;
    case OP_EFINAL:
        op_efinal<S>(t, level, E.it0);
        break;
    default:
        break;
    }
}


// ------------------------------------------------------------------ per-storage dispatch

template <typename T, int SH>
void tick_shape(const PairDesc *desc, const TickArgs &args, hipStream_t st)
{
    k_tick<T, SH><<<dim3((unsigned)args.total), dim3(Sh<SH>::BX, Sh<SH>::BY), 0, st>>>(desc, args);
}

template <typename T>
void tick_storage(const PairDesc *desc, const TickArgs &args, int shape, hipStream_t st)
{
    switch (shape) {
    case 1: tick_shape<T, 1>(desc, args, st); break;
    case 2: tick_shape<T, 2>(desc, args, st); break;
    case 3: tick_shape<T, 3>(desc, args, st); break;
    case 4: tick_shape<T, 4>(desc, args, st); break;
    case 5: tick_shape<T, 5>(desc, args, st); break;
    case 6: tick_shape<T, 6>(desc, args, st); break;
    default: tick_shape<T, 0>(desc, args, st); break;
    }
}
static_assert(kTickShapes == 7, "tick_storage dispatches every tick shape");

// K-iteration launches exist for the shapes of multi_iter_shape() (64x4, 32x16, 32x8,
// 16x16, 8x8, 4x4, 32x32, 20x20), whose staged regions fit the workgroup's LDS.
template <typename T, int SH, int K>
void iterk_shape(const PairDesc *desc, const TickArgs &args, bool half, hipStream_t st)
{
    if constexpr (iterk_ok(SH, K)) {  // the host asks for no other
        const dim3 grid((unsigned)args.total);
        if constexpr (half_workgroup_ok(SH, K)) {
            if (half) {
                k_iterk<T, SH, K, true><<<grid, dim3(BlkK<SH, K, true>::NT), 0, st>>>(desc, args);
                return;
            }
        }
        k_iterk<T, SH, K, false><<<grid, dim3(BlkK<SH, K, false>::NT), 0, st>>>(desc, args);
    }
}

template <typename T, int K>
void iterk_storage(const PairDesc *desc, const TickArgs &args, int launch_shape, hipStream_t st)
{
    const bool half = (launch_shape & kHalfWorkgroup) != 0;
    switch (shape_index(launch_shape)) {
    case 1: iterk_shape<T, 1, K>(desc, args, half, st); break;
    case 3: iterk_shape<T, 3, K>(desc, args, half, st); break;
    case 4: iterk_shape<T, 4, K>(desc, args, half, st); break;
    case 6: iterk_shape<T, 6, K>(desc, args, half, st); break;
    case 7: iterk_shape<T, 7, K>(desc, args, half, st); break;
    case 8: iterk_shape<T, 8, K>(desc, args, half, st); break;
    case 9: iterk_shape<T, 9, K>(desc, args, half, st); break;
    case 10: iterk_shape<T, 10, K>(desc, args, half, st); break;
    case 11: iterk_shape<T, 11, K>(desc, args, half, st); break;
    case 12: iterk_shape<T, 12, K>(desc, args, half, st); break;
    case 13: iterk_shape<T, 13, K>(desc, args, half, st); break;
    case 14: iterk_shape<T, 14, K>(desc, args, half, st); break;
    case 15: iterk_shape<T, 15, K>(desc, args, half, st); break;
    case 16: iterk_shape<T, 16, K>(desc, args, half, st); break;
    case 17: iterk_shape<T, 17, K>(desc, args, half, st); break;
    default: break;  // the host never asks for another shape
    }
}

template <typename T, int SH, int K>
void iterkp_shape(const PairDesc *desc, const TickArgs &args, hipStream_t st)
{
    if constexpr (persistk_shape(SH, K))
        k_iterkp<T, SH, K><<<dim3((unsigned)args.total), dim3(BlkK<SH, K, false>::NT), 0, st>>>(desc, args);
}

template <typename T, int SH, int K>
int iterkp_shape_capacity()
{
    if constexpr (persistk_shape(SH, K)) {
        int per = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_iterkp<T, SH, K>, BlkK<SH, K, false>::NT, 0) !=
            hipSuccess)
            return 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        return per * cus;
    }
    return 0;
}

// The blocked persistent op's shapes (persistk_shape): 16x16, 8x8, 4x4, 20x20, 27x25, 14x13.
#define VFS_PERSISTK_SHAPES(X) X(6) X(7) X(8) X(10) X(12) X(13)

template <typename T, int K>
void iterkp_storage(const PairDesc *desc, const TickArgs &args, int shape, hipStream_t st)
{
    switch (shape) {
#define VFS_CASE(sh) \
    case sh: iterkp_shape<T, sh, K>(desc, args, st); break;
        VFS_PERSISTK_SHAPES(VFS_CASE)
#undef VFS_CASE
    default: break;  // the host never asks for another shape
    }
}

template <typename T, int K>
int iterkp_capacity(int shape)
{
    switch (shape) {
#define VFS_CASE(sh) \
    case sh: return iterkp_shape_capacity<T, sh, K>();
        VFS_PERSISTK_SHAPES(VFS_CASE)
#undef VFS_CASE
    default: return 0;
    }
}

template <typename T, int SH>
void iterp_shape(const PairDesc *desc, const TickArgs &args, hipStream_t st)
{
    if constexpr (persist_shape(SH))
        k_iterp<T, SH><<<dim3((unsigned)args.total), dim3(BlkP<SH>::NT), 0, st>>>(desc, args);
}

template <typename T, int SH>
int iterp_shape_capacity()
{
    if constexpr (persist_shape(SH)) {
        int per = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_iterp<T, SH>, BlkP<SH>::NT, 0) != hipSuccess) return 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        return per * cus;
    }
    return 0;
}

// The persistent op's shapes (persist_shape): 8x8, 4x4, 36x25, 27x25, 14x13, 7x7, 53x25, 53x50.
#define VFS_PERSIST_SHAPES(X) X(7) X(8) X(11) X(12) X(13) X(14) X(17) X(18)

template <typename T>
void iterp_storage(const PairDesc *desc, const TickArgs &args, int shape, hipStream_t st)
{
    switch (shape) {
#define VFS_CASE(sh) \
    case sh: iterp_shape<T, sh>(desc, args, st); break;
        VFS_PERSIST_SHAPES(VFS_CASE)
#undef VFS_CASE
    default: break;  // the host never asks for another shape
    }
}

template <typename T>
int iterp_capacity(int shape)
{
    switch (shape) {
#define VFS_CASE(sh) \
    case sh: return iterp_shape_capacity<T, sh>();
        VFS_PERSIST_SHAPES(VFS_CASE)
#undef VFS_CASE
    default: return 0;
    }
}

}  // namespace

}  // namespace vfs
