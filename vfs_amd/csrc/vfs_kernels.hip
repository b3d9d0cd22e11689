// gfx950 kernels of the StereoTgv TGV-L1 fisheye-stereo solver.
//
// Layout: every plane is row-major and pitched (stride s elements, s = alignUp(w, 32)),
// float / float2 / float4 per pixel.  Buffers are allocated at level-0 size and reused
// at every level with that level's stride, as in the reference (stereotgv.cu:106-161).
//
// Launch shape: 64x4 work-items per workgroup: one wave64 per image row segment, so
// every plane access of a wave is one contiguous 256 B (float) / 512 B (float2) run.
//
// Texture fetches of the reference (linear filter, mirror addressing, normalized
// coordinates) are evaluated in registers: exact bilinear weights at pixel-index
// position p = x_norm * dim - 0.5 over the symmetric extension of the image.
#include "vfs_kernels.h"

#include <math.h>

namespace vfs {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;

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
__device__ __forceinline__ float bil(const float *__restrict__ img, int w, int h, int s, float px,
                                     float py)
{
    if (__builtin_isnan(px) || __builtin_isnan(py)) return __builtin_nanf("");
    px = clampc(px);
    py = clampc(py);
    const float fx = floorf(px), fy = floorf(py);
    const float ax = px - fx, ay = py - fy;
    const int i = (int)fx, j = (int)fy;
    const int i0 = mir(i, w), i1 = mir(i + 1, w);
    const int j0 = mir(j, h), j1 = mir(j + 1, h);
    const float *r0 = img + (size_t)j0 * s;
    const float *r1 = img + (size_t)j1 * s;
    const float top = (1.f - ax) * r0[i0] + ax * r0[i1];
    const float bot = (1.f - ax) * r1[i0] + ax * r1[i1];
    return (1.f - ay) * top + ay * bot;
}

__device__ __forceinline__ float2 bil2(const float2 *__restrict__ img, int w, int h, int s, float px,
                                       float py)
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

__device__ __forceinline__ float texel(const float *__restrict__ img, int w, int h, int s, int x, int y)
{
    return img[(size_t)mir(y, h) * s + mir(x, w)];
}

// Circle gate of the unmasked solve (updateDualVariables.cu:13-22): true when
// radius >= width / 2.2 around (w/2, h/2).  Evaluated in double on the float bound.
__device__ __forceinline__ bool outside_circle(int x, int y, int w, int h)
{
    const float desired = (float)w / 2.20f;
    const double hw = (double)((float)w / 2.0f), hh = (double)((float)h / 2.0f);
    const double dx = (double)x - hw, dy = (double)y - hh;
    return dx * dx + dy * dy >= (double)desired * (double)desired;
}

__device__ __forceinline__ bool pixel(int &x, int &y, int w, int h)
{
    x = blockIdx.x * BX + threadIdx.x;
    y = blockIdx.y * BY + threadIdx.y;
    return x < w && y < h;
}

inline dim3 grid_for(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }
inline dim3 block2d() { return dim3(BX, BY); }

// ------------------------------------------------------------------ kernels

__global__ void k_fill_zero(float4 *__restrict__ p, size_t n4)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// imageConverter.cu:3-16
__global__ void k_cv8u(const uint8_t *__restrict__ in, int ip, float *__restrict__ out, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    out[(size_t)y * s + x] = (float)in[(size_t)y * ip + x] / 256.0f;
}

// imageConverter.cu:26-42
__global__ void k_cv8uc3(const uint8_t *__restrict__ in, int ip, float *__restrict__ out, int w, int h,
                         int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const uint8_t *px = in + (size_t)y * ip + 3 * (size_t)x;
    const float g = ((float)px[0] + (float)px[1] + (float)px[2]) / 3;
    out[(size_t)y * s + x] = g / 256.0f;
}

// warpImage.cu:9-23
__global__ void k_warp(const float *__restrict__ src, const float2 *__restrict__ uv, float *__restrict__ out,
                       int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float2 d = uv[pos];
    out[pos] = bil(src, w, h, s, (float)x + d.x, (float)y + d.y);
}

// downscale.cu:10-30 for two images at once (I0 and I1 pyramids, stereotgv.cu:400-405)
__global__ void k_downscale2(const float *__restrict__ a, const float *__restrict__ b, int W, int H, int S,
                             float *__restrict__ oa, float *__restrict__ ob, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const float rx = (float)W / (float)w, ry = (float)H / (float)h;
    const float cx = ((float)x + 0.5f) * rx - 0.5f, cy = ((float)y + 0.5f) * ry - 0.5f;
    const float ox = 0.25f * rx, oy = 0.25f * ry;
    const size_t pos = (size_t)y * s + x;
    oa[pos] = 0.25f * (bil(a, W, H, S, cx - ox, cy) + bil(a, W, H, S, cx + ox, cy) +
                       bil(a, W, H, S, cx, cy - oy) + bil(a, W, H, S, cx, cy + oy));
    ob[pos] = 0.25f * (bil(b, W, H, S, cx - ox, cy) + bil(b, W, H, S, cx + ox, cy) +
                       bil(b, W, H, S, cx, cy - oy) + bil(b, W, H, S, cx, cy + oy));
}

// downscale.cu:98-122 (trajectory pyramid, stereotgv.cu:187-194)
__global__ void k_downscale_f2(const float2 *__restrict__ src, int W, int H, int S, float2 *__restrict__ out,
                               int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const float rx = (float)W / (float)w, ry = (float)H / (float)h;
    const float cx = ((float)x + 0.5f) * rx - 0.5f, cy = ((float)y + 0.5f) * ry - 0.5f;
    const float ox = 0.25f * rx, oy = 0.25f * ry;
    const float2 v0 = bil2(src, W, H, S, cx - ox, cy), v1 = bil2(src, W, H, S, cx + ox, cy);
    const float2 v2 = bil2(src, W, H, S, cx, cy - oy), v3 = bil2(src, W, H, S, cx, cy + oy);
    out[(size_t)y * s + x] = make_float2(0.25f * (v0.x + v1.x + v2.x + v3.x), 0.25f * (v0.y + v1.y + v2.y + v3.y));
}

// downscale.cu:52-73: point sample at the coarse pixel centre, texel floor((x+.5) W / w)
__global__ void k_downscale_nn(const float *__restrict__ src, int W, int H, int S, float *__restrict__ out,
                               int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const long fx = ((2L * x + 1) * W) / (2L * w);
    const long fy = ((2L * y + 1) * H) / (2L * h);
    out[(size_t)y * s + x] = texel(src, W, H, S, (int)fx, (int)fy);
}

// warpImage.cu:136-150 (coarsest-level init warpUV = u * tv, stereotgv.cu:419-421)
__global__ void k_flowvec(const float *__restrict__ u, const float2 *__restrict__ tv, float2 *__restrict__ uv,
                          int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float us = u[pos];
    const float2 t = tv[pos];
    uv[pos] = make_float2(us * t.x, us * t.y);
}

// 3x3 binomial of the level image at (x, y), mirror addressing (calcTensor.cu:245-273).
__device__ __forceinline__ float gauss3(const float *__restrict__ I, int w, int h, int s, int x, int y)
{
    float t = 0.25f * texel(I, w, h, s, x, y);
    t += 0.0625f * texel(I, w, h, s, x - 1, y - 1);
    t += 0.0625f * texel(I, w, h, s, x - 1, y + 1);
    t += 0.0625f * texel(I, w, h, s, x + 1, y - 1);
    t += 0.0625f * texel(I, w, h, s, x + 1, y + 1);
    t += 0.125f * texel(I, w, h, s, x - 1, y);
    t += 0.125f * texel(I, w, h, s, x + 1, y);
    t += 0.125f * texel(I, w, h, s, x, y - 1);
    t += 0.125f * texel(I, w, h, s, x, y + 1);
    return t;
}

// The smoothed image as the tensor kernel samples it: the masked blur writes only
// mask != 0 pixels of a zeroed plane, and the texture fetch mirrors at the level edge.
__device__ __forceinline__ float smoothed(const float *__restrict__ I, const float *__restrict__ mask, int w,
                                          int h, int s, int x, int y)
{
    const int mx = mir(x, w), my = mir(y, h);
    if (mask && mask[(size_t)my * s + mx] == 0.f) return 0.f;
    return gauss3(I, w, h, s, mx, my);
}

// Fused GaussianMasked + CalcTensorMasked + SolveEtaMasked (calcTensor.cu:154-273,
// preconditioning.cu:41-60), or the unmasked trio when mask == nullptr.  Writes a, b,
// c for every in-image pixel: the tensor where the reference computes it, 0 where the
// reference leaves the per-level memset (mask == 0, or columns x >= xlimit, the
// height-sized launch of calcTensor.cu:81,227).  eta is recomputed from (a, b, c) by
// the iteration kernels.
__global__ void k_tensor(const float *__restrict__ I0, const float *__restrict__ mask, float beta, float gamma,
                         int xlimit, float *__restrict__ at, float *__restrict__ bt, float *__restrict__ ct,
                         int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    float a = 0.f, b = 0.f, c = 0.f;
    const bool active = mask ? (mask[pos] != 0.f) : true;
    if (active && x < xlimit) {
        const float s11 = smoothed(I0, mask, w, h, s, x + 1, y + 1);
        const float s01 = smoothed(I0, mask, w, h, s, x, y + 1);
        const float s10 = smoothed(I0, mask, w, h, s, x + 1, y);
        const float gx = s11 - s01;
        const float gy = s11 - s10;
        const float absg = sqrtf(gx * gx + gy * gy);
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
    at[pos] = a;
    bt[pos] = b;
    ct[pos] = c;
}

// Per-warp preparation (stereotgv.cu:433-447): tv2 = bilinear(tv, x + warpUV)
// (warpImage.cu:96-111), I1w = bilinear(I1, x + warpUV) (warpImage.cu:9-23),
// u_last = u, v_ = 0 (Clone of the zeroed v).
__global__ void k_warp_prep(const float2 *__restrict__ uv, const float2 *__restrict__ tv,
                            const float *__restrict__ I1, const float *__restrict__ u, float2 *__restrict__ tv2,
                            float *__restrict__ i1w, float *__restrict__ u_last, float2 *__restrict__ v_, int w,
                            int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float2 d = uv[pos];
    const float px = (float)x + d.x, py = (float)y + d.y;
    tv2[pos] = bil2(tv, w, h, s, px, py);
    i1w[pos] = bil(I1, w, h, s, px, py);
    u_last[pos] = u[pos];
    v_[pos] = make_float2(0.f, 0.f);
}

// computeDerivatives.cu:88-129
__global__ void k_deriv(const float *__restrict__ I0, const float *__restrict__ I1w, const float2 *__restrict__ tv,
                        float *__restrict__ Iu, float *__restrict__ Iz, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float2 v = tv[pos];
    const float r = sqrtf(v.x * v.x + v.y * v.y);
    const float dx = v.x / r, dy = v.y / r;
    const float X = (float)x, Y = (float)y;
    float t0 = bil(I0, w, h, s, X - 2.f * dx, Y - 2.f * dy);
    t0 -= bil(I0, w, h, s, X - dx, Y - dy) * 8.f;
    t0 += bil(I0, w, h, s, X + dx, Y + dy) * 8.f;
    t0 -= bil(I0, w, h, s, X + 2.f * dx, Y + 2.f * dy);
    t0 /= 12.f;
    float t1 = bil(I1w, w, h, s, X - 2.f * dx, Y - 2.f * dy);
    t1 -= bil(I1w, w, h, s, X - dx, Y - dy) * 8.f;
    t1 += bil(I1w, w, h, s, X + dx, Y + dy) * 8.f;
    t1 -= bil(I1w, w, h, s, X + 2.f * dx, Y + 2.f * dy);
    t1 /= 12.f;
    Iu[pos] = (t0 + t1) * 0.5f;
    Iz[pos] = I1w[pos] - I0[pos];
}

__device__ __forceinline__ float ld0(const float *__restrict__ b, long i) { return i < 0 ? 0.f : b[i]; }
__device__ __forceinline__ float2 ld0(const float2 *__restrict__ b, long i)
{
    return i < 0 ? make_float2(0.f, 0.f) : b[i];
}

// Dual ascent on p, q with unit-ball projection (updateDualVariables.cu:122-237 masked,
// :4-101 unmasked) followed by Tp = T p (thresholdingL1.cu:234-244 / :104-115).
template <bool MASKED>
__global__ void k_dual_tp(const float *__restrict__ mask, const float *__restrict__ u_,
                          const float2 *__restrict__ v_, const float *__restrict__ at, const float *__restrict__ bt,
                          const float *__restrict__ ct, float k1, float k0, int first, float2 *__restrict__ p,
                          float4 *__restrict__ q, float2 *__restrict__ Tp, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const long pos = (long)y * s + x;
    float2 pn;
    if (MASKED) {
        if (mask[pos] == 0.f) return;
    } else if (outside_circle(x, y, w, h)) {
        pn = make_float2(0.f, 0.f);
        p[pos] = pn;
        q[pos] = make_float4(0.f, 0.f, 0.f, 0.f);
        const float a = at[pos], b = bt[pos], c = ct[pos];
        Tp[pos] = make_float2(a * pn.x + c * pn.y, c * pn.x + b * pn.y);
        return;
    }
    bool hasR, hasD;
    if (MASKED) {
        hasR = (x + 1 < w) && (mask[pos + 1] != 0.f);
        hasD = (y + 1 < h) && (mask[pos + s] != 0.f);
    } else {
        hasR = x + 1 < w;
        hasD = y + 1 < h;
    }
    const float u0 = u_[pos];
    const float2 v0 = v_[pos];
    const float ux = hasR ? (u_[pos + 1] - u0 - v0.x) : (u0 - ld0(u_, pos - 1) - v0.x);
    const float uy = hasD ? (u_[pos + s] - u0 - v0.y) : (u0 - ld0(u_, pos - s) - v0.y);
    const float a = at[pos], b = bt[pos], c = ct[pos];
    const float dtx = a * ux + c * uy;
    const float dty = c * ux + b * uy;
    const float2 po = first ? make_float2(0.f, 0.f) : p[pos];
    const float px = po.x + k1 * dtx;
    const float py = po.y + k1 * dty;
    float rp = sqrtf(px * px + py * py);
    if (rp < 1.0f) rp = 1.0f;
    pn = make_float2(px / rp, py / rp);
    p[pos] = pn;
    float g0, g1, g2, g3;
    if (hasR) {
        const float2 vr = v_[pos + 1];
        g0 = vr.x - v0.x;
        g3 = vr.y - v0.y;
    } else {
        const float2 vl = ld0(v_, pos - 1);
        g0 = v0.x - vl.x;
        g3 = v0.y - vl.y;
    }
    if (hasD) {
        const float2 vd = v_[pos + s];
        g1 = vd.y - v0.y;
        g2 = vd.x - v0.x;
    } else {
        const float2 vu = ld0(v_, pos - s);
        g1 = v0.y - vu.y;
        g2 = v0.x - vu.x;
    }
    const float4 qo = first ? make_float4(0.f, 0.f, 0.f, 0.f) : q[pos];
    const float q0 = qo.x + k0 * g0, q1 = qo.y + k0 * g1, q2 = qo.z + k0 * g2, q3 = qo.w + k0 * g3;
    float rq = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    if (rq < 1.0f) rq = 1.0f;
    q[pos] = make_float4(q0 / rq, q1 / rq, q2 / rq, q3 / rq);
    Tp[pos] = make_float2(a * pn.x + c * pn.y, c * pn.x + b * pn.y);
}

// L1 proximal step (thresholdingL1.cu:180-215), linearised at the current u.
__device__ __forceinline__ float thresh_core(float divp, float tau, float etau, float u_pos, float upos,
                                             float Iu, float Iz, float lambda)
{
    const float te = (etau == 0.f) ? tau : tau / etau;
    const float uhat = u_pos + te * divp;
    const float dun = uhat - upos;
    const float rho = Iu * dun + Iz;
    const float upper = lambda * te * (Iu * Iu);
    const float lower = -lambda * te * (Iu * Iu);
    float du = 0.f;
    if ((rho <= upper) && (rho >= lower)) {
        du = (Iu == 0.f) ? dun : dun - rho / Iu;
    } else if (rho < lower) {
        du = dun + lambda * te * Iu;
    } else if (rho > upper) {
        du = dun - lambda * te * Iu;
    }
    return upos + du;
}

// ThresholdingL1 (thresholdingL1.cu:128-216 masked, :3-86 unmasked) fused with
// UpdatePrimalVariables (updatePrimalVariables.cu:138-245 masked, :3-113 unmasked).
// The primal step's `u` is the thresholding result at the same pixel (the swap of
// stereotgv.cu:468 happens between the two launches in the reference).
// eta_u, eta_v1, eta_v2 are recomputed from (a, b, c) (preconditioning.cu:57-59).
template <bool MASKED>
__global__ void k_thresh_primal(const float *__restrict__ mask, const float2 *__restrict__ Tp,
                                const float *__restrict__ u_, const float2 *__restrict__ v_,
                                const float *__restrict__ Iu, const float *__restrict__ Iz,
                                const float *__restrict__ at, const float *__restrict__ bt,
                                const float *__restrict__ ct, const float2 *__restrict__ p,
                                const float4 *__restrict__ q, const float *__restrict__ u, float tau, float mu,
                                float lambda, float alpha0, float alpha1, float *__restrict__ us,
                                float *__restrict__ u_s, float2 *__restrict__ v_s, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const long pos = (long)y * s + x;
    bool hasR, hasL, hasD, hasU;
    if (MASKED) {
        if (mask[pos] == 0.f) return;
        hasR = (x + 1 < w) && (mask[pos + 1] != 0.f);
        hasL = (x >= 1) && (mask[pos - 1] != 0.f);
        hasD = (y + 1 < h) && (mask[pos + s] != 0.f);
        hasU = (y >= 1) && (mask[pos - s] != 0.f);
    } else {
        if (outside_circle(x, y, w, h)) {
            us[pos] = 0.f;
            u_s[pos] = 0.f;
            v_s[pos] = make_float2(0.f, 0.f);
            return;
        }
        hasR = x + 1 < w;
        hasL = x >= 1;
        hasD = y + 1 < h;
        hasU = y >= 1;
    }
    const float a = at[pos], b = bt[pos], c = ct[pos];
    const float aa = alpha1 * alpha1;
    // ---- thresholding
    const float2 tp0 = Tp[pos];
    float dx, dy;
    if (MASKED) {
        if (hasL && hasR) dx = tp0.x - Tp[pos - 1].x;
        else if (!hasR) dx = -ld0(Tp, pos - 1).x;
        else dx = tp0.x;
        if (hasU && hasD) dy = tp0.y - Tp[pos - s].y;
        else if (!hasD) dy = -ld0(Tp, pos - s).y;
        else dy = tp0.y;
    } else {
        if (hasL) dx = tp0.x - Tp[pos - 1].x;
        else if (x == w - 1) dx = -ld0(Tp, pos - 1).x;
        else dx = tp0.x;
        if (hasU) dy = tp0.y - Tp[pos - s].y;
        else if (y == h - 1) dy = -ld0(Tp, pos - s).y;
        else dy = tp0.y;
    }
    const float etau = (a * a + b * b + 2 * c * c + (a + c) * (a + c) + (b + c) * (b + c)) * aa;
    const float u_pos = u_[pos];
    const float un = thresh_core(dx + dy, tau, etau, u_pos, u[pos], Iu[pos], Iz[pos], lambda);
    us[pos] = un;
    // ---- primal
    const float4 qc4 = q[pos];
    float qcx = 0.f, qcw = 0.f, qcy = 0.f, qcz = 0.f;
    if (hasR) {
        qcx = qc4.x;
        qcw = qc4.w;
    }
    if (hasD) {
        qcy = qc4.y;
        qcz = qc4.z;
    }
    float qwx = 0.f, qww = 0.f, qny = 0.f, qnz = 0.f;
    if (hasL) {
        const float4 ql = q[pos - 1];
        qwx = ql.x;
        qww = ql.w;
    }
    if (hasU) {
        const float4 qu = q[pos - s];
        qny = qu.y;
        qnz = qu.z;
    }
    const float dqx = (qcx - qwx) + (qcz - qnz);
    const float dqy = (qcw - qww) + (qcy - qny);
    const float2 pp = p[pos];
    const float tx = a * pp.x + c * pp.y;
    const float ty = c * pp.x + b * pp.y;
    const float e1 = aa * (b * b + c * c) + 4 * alpha0 * alpha0;
    const float e2 = aa * (a * a + c * c) + 4 * alpha0 * alpha0;
    const float2 vo = v_[pos];
    // e1, e2 >= 4 alpha0^2 > 0 for alpha0 != 0; the reference's zero branches
    // (updatePrimalVariables.cu:228-239) are kept for alpha0 == 0.
    float vx, vy;
    vx = vo.x + ((e1 != 0.f) ? tau / e1 : tau) * (alpha1 * tx + alpha0 * dqx);
    if (e2 != 0.f) {
        vy = vo.y + (tau / e2) * (alpha1 * ty + alpha0 * dqy);
    } else {
        vx = vo.x + tau * (alpha1 * tx + alpha0 * dqx);
        vy = 0.f;  // unwritten component of the write-only v
    }
    u_s[pos] = un + mu * (un - u_pos);
    v_s[pos] = make_float2(vx + mu * (vx - vo.x), vy + mu * (vy - vo.y));
}

// 13th smallest of 25 by the partial selection sort of medianFilter.cu:307-316.
__device__ __forceinline__ float median25(float m[25])
{
#pragma unroll
    for (int j = 0; j < 13; j++) {
#pragma unroll
        for (int i = j + 1; i < 25; i++) {
            const float a = m[j], b = m[i];
            const bool sw = a > b;
            m[j] = sw ? b : a;
            m[i] = sw ? a : b;
        }
    }
    return m[12];
}

// 5x5 window as medianFilter.cu:282-304: edge replicate when exactly one of row /
// column is outside the image, 0 when both are.
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

// End of a warp in the masked solve (stereotgv.cu:481-493): median (medianFilter.cu:264-319)
// into out_u, then du = min(out_u - u_last, limit) (scalar.cu:129-141, limitRange.cu:3-24),
// out_u = u_last + du (scalar.cu:77-89), u_ = out_u (copyClone.cu:3-12),
// warpUV += du * tv2 (warpImage.cu:136-150, scalar.cu:102-115).
__global__ void k_median_update(const float *__restrict__ u, const float *__restrict__ u_last,
                                const float2 *__restrict__ tv2, float limit, float *__restrict__ out_u,
                                float *__restrict__ u_, float2 *__restrict__ uv, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float med = median5_at(u, x, y, w, h, s);
    const float ul = u_last[pos];
    float du = med - ul;
    du = (du > limit) ? limit : du;
    const float un = ul + du;
    out_u[pos] = un;
    u_[pos] = un;
    const float2 t = tv2[pos];
    const float2 o = uv[pos];
    uv[pos] = make_float2(o.x + du * t.x, o.y + du * t.y);
}

// Same without the median (unmasked solve, stereotgv.cu:354-363), u updated in place.
__global__ void k_update(float *__restrict__ u, const float *__restrict__ u_last, const float2 *__restrict__ tv2,
                         float limit, float *__restrict__ u_, float2 *__restrict__ uv, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const float ul = u_last[pos];
    float du = u[pos] - ul;
    du = (du > limit) ? limit : du;
    const float un = ul + du;
    u[pos] = un;
    u_[pos] = un;
    const float2 t = tv2[pos];
    const float2 o = uv[pos];
    uv[pos] = make_float2(o.x + du * t.x, o.y + du * t.y);
}

// upscale.cu:7-21, 46-61 for u, u_ and warpUV at once (stereotgv.cu:500-509)
__global__ void k_upscale3(const float *__restrict__ u, const float *__restrict__ u_, const float2 *__restrict__ uv,
                           int w, int h, int s, float scale, float *__restrict__ ou, float *__restrict__ ou_,
                           float2 *__restrict__ ouv, int W, int H, int S)
{
    int x, y;
    if (!pixel(x, y, W, H)) return;
    const float rx = (float)w / (float)W, ry = (float)h / (float)H;
    const float px = ((float)x + 0.5f) * rx - 0.5f, py = ((float)y + 0.5f) * ry - 0.5f;
    const size_t pos = (size_t)y * S + x;
    ou[pos] = bil(u, w, h, s, px, py) * scale;
    ou_[pos] = bil(u_, w, h, s, px, py) * scale;
    const float2 v = bil2(uv, w, h, s, px, py);
    ouv[pos] = make_float2(v.x * scale, v.y * scale);
}

__device__ double kb_theta(double r, double d1, double d2, double d3, double d4)
{
    double th = 0.0;
    for (int it = 0; it < 5; it++) {
        const double t2 = th * th, t3 = t2 * th, t4 = t2 * t2, t5 = t4 * th, t6 = t4 * t2, t7 = t6 * th;
        const double t8 = t4 * t4, t9 = t8 * th;
        const double td = th + d1 * t3 + d2 * t5 + d3 * t7 + d4 * t9;
        const double dtd = 1.0 + 3.0 * d1 * t2 + 5.0 * d2 * t4 + 7.0 * d3 * t6 + 9.0 * d4 * t8;
        th = th - (r - td) / (-dtd);
    }
    return th;
}

// conversion.cu:3-69 in FP64: Kannala-Brandt inversion of both frames, Z from the x
// components, X = (x0 Z, y0 Z, Z), depth = |X|.
__global__ void k_convert_kb(const float2 *__restrict__ uv, double fx, double fy, double cx, double cy, double d1,
                             double d2, double d3, double d4, double t1, double t2, double t3, float *__restrict__ X,
                             float *__restrict__ depth, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const size_t pos = (size_t)y * s + x;
    const double u0 = x, v0 = y;
    const double xp0 = (u0 - fx) / cx, yp0 = (v0 - fy) / cy;
    const float2 d = uv[pos];
    const double xp1 = (u0 + (double)d.x - fx) / cx, yp1 = (v0 + (double)d.y - fy) / cy;
    const double r0 = sqrt(xp0 * xp0 + yp0 * yp0);
    const double th0 = kb_theta(r0, d1, d2, d3, d4);
    const double x0o = tan(th0) * xp0 / r0, y0o = tan(th0) * yp0 / r0;
    const double r1 = sqrt(xp1 * xp1 + yp1 * yp1);
    const double th1 = kb_theta(r1, d1, d2, d3, d4);
    const double x1o = tan(th1) * xp1 / r1;
    (void)t2;
    const double Z = (t1 - x1o * t3) / (x1o - x0o);
    const double X0 = x0o * Z, X1 = y0o * Z, X2 = Z;
    X[3 * pos] = (float)X0;
    X[3 * pos + 1] = (float)X1;
    X[3 * pos + 2] = (float)X2;
    depth[pos] = (float)sqrt(X0 * X0 + X1 * X1 + X2 * X2);
}


// ------------------------------------------------------------ fused iteration
//
// One TGV-L1 inner iteration in one launch: dual ascent + projection
// (updateDualVariables.cu:122-237), Tp (thresholdingL1.cu:234-244), L1 thresholding
// (thresholdingL1.cu:128-216) and the primal step with over-relaxation
// (updatePrimalVariables.cu:138-245).  Each workgroup owns a 64x8 tile.  Phase A
// computes the new p, q and Tp on the tile plus its left column and top row (the
// threshold and primal of a pixel read Tp and q of the left and upper neighbours) into
// LDS; phase B finishes thresholding and the primal step of the tile.  Old p, q are
// read from one buffer and new ones written to the other, so every workgroup sees the
// previous iteration's values in its recomputed halo (the reference's Jacobi order).
constexpr int FTX = 64, FTY = 8;
constexpr int EXW = FTX + 1, EXH = FTY + 1;

template <bool MASKED>
__global__ __launch_bounds__(256) void k_tgv_iter(
    const float *__restrict__ mask, const float *__restrict__ u_, const float2 *__restrict__ v_,
    const float *__restrict__ at, const float *__restrict__ bt, const float *__restrict__ ct,
    const float *__restrict__ Iu, const float *__restrict__ Iz, const float *__restrict__ u,
    const float2 *__restrict__ p_in, const float4 *__restrict__ q_in, float2 *__restrict__ p_out,
    float4 *__restrict__ q_out, float2 *__restrict__ Tp, float *__restrict__ us, float *__restrict__ u_s,
    float2 *__restrict__ v_s, IterParams ip, int w, int h, int s)
{
    __shared__ float2 sP[EXH][EXW];
    __shared__ float4 sQ[EXH][EXW];
    __shared__ float2 sT[EXH][EXW];
    const int x0 = blockIdx.x * FTX, y0 = blockIdx.y * FTY;
    const int tid = threadIdx.y * BX + threadIdx.x;

    // ---- phase A: dual update on the tile + left column + top row
    for (int i = tid; i < EXH * EXW; i += BX * BY) {
        const int ey = i / EXW, ex = i - ey * EXW;
        const int x = x0 - 1 + ex, y = y0 - 1 + ey;
        if (x < 0 || y < 0 || x >= w || y >= h) continue;
        const long pos = (long)y * s + x;
        const bool interior = ex >= 1 && ey >= 1;
        const float a = at[pos], b = bt[pos], c = ct[pos];
        float2 pn;
        float4 qn;
        bool hasR, hasD;
        if (MASKED) {
            if (mask[pos] == 0.f) continue;
            hasR = (x + 1 < w) && (mask[pos + 1] != 0.f);
            hasD = (y + 1 < h) && (mask[pos + s] != 0.f);
        } else {
            hasR = x + 1 < w;
            hasD = y + 1 < h;
        }
        if (!MASKED && outside_circle(x, y, w, h)) {
            pn = make_float2(0.f, 0.f);
            qn = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            const float u0 = u_[pos];
            const float2 v0 = v_[pos];
            const float ux = hasR ? (u_[pos + 1] - u0 - v0.x) : (u0 - ld0(u_, pos - 1) - v0.x);
            const float uy = hasD ? (u_[pos + s] - u0 - v0.y) : (u0 - ld0(u_, pos - s) - v0.y);
            const float dtx = a * ux + c * uy;
            const float dty = c * ux + b * uy;
            const float2 po = ip.first ? make_float2(0.f, 0.f) : p_in[pos];
            const float px = po.x + ip.k1 * dtx;
            const float py = po.y + ip.k1 * dty;
            float rp = sqrtf(px * px + py * py);
            if (rp < 1.0f) rp = 1.0f;
            pn = make_float2(px / rp, py / rp);
            float g0, g1, g2, g3;
            if (hasR) {
                const float2 vr = v_[pos + 1];
                g0 = vr.x - v0.x;
                g3 = vr.y - v0.y;
            } else {
                const float2 vl = ld0(v_, pos - 1);
                g0 = v0.x - vl.x;
                g3 = v0.y - vl.y;
            }
            if (hasD) {
                const float2 vd = v_[pos + s];
                g1 = vd.y - v0.y;
                g2 = vd.x - v0.x;
            } else {
                const float2 vu = ld0(v_, pos - s);
                g1 = v0.y - vu.y;
                g2 = v0.x - vu.x;
            }
            const float4 qo = ip.first ? make_float4(0.f, 0.f, 0.f, 0.f) : q_in[pos];
            const float q0 = qo.x + ip.k0 * g0, q1 = qo.y + ip.k0 * g1, q2 = qo.z + ip.k0 * g2,
                        q3 = qo.w + ip.k0 * g3;
            float rq = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
            if (rq < 1.0f) rq = 1.0f;
            qn = make_float4(q0 / rq, q1 / rq, q2 / rq, q3 / rq);
        }
        const float2 tpn = make_float2(a * pn.x + c * pn.y, c * pn.x + b * pn.y);
        sP[ey][ex] = pn;
        sQ[ey][ex] = qn;
        sT[ey][ex] = tpn;
        if (interior) {
            p_out[pos] = pn;
            q_out[pos] = qn;
            if (ip.write_tp) Tp[pos] = tpn;
        }
    }
    __syncthreads();

    // ---- phase B: thresholding + primal step on the tile
    const int x = x0 + threadIdx.x;
    if (x >= w) return;
    for (int ty = threadIdx.y; ty < FTY; ty += BY) {
        const int y = y0 + ty;
        if (y >= h) break;
        const long pos = (long)y * s + x;
        const int ex = threadIdx.x + 1, ey = ty + 1;
        bool hasR, hasL, hasD, hasU;
        if (MASKED) {
            if (mask[pos] == 0.f) continue;
            hasR = (x + 1 < w) && (mask[pos + 1] != 0.f);
            hasL = (x >= 1) && (mask[pos - 1] != 0.f);
            hasD = (y + 1 < h) && (mask[pos + s] != 0.f);
            hasU = (y >= 1) && (mask[pos - s] != 0.f);
        } else {
            if (outside_circle(x, y, w, h)) {
                us[pos] = 0.f;
                u_s[pos] = 0.f;
                v_s[pos] = make_float2(0.f, 0.f);
                continue;
            }
            hasR = x + 1 < w;
            hasL = x >= 1;
            hasD = y + 1 < h;
            hasU = y >= 1;
        }
        const float a = at[pos], b = bt[pos], c = ct[pos];
        const float aa = ip.alpha1 * ip.alpha1;
        const float2 tp0 = sT[ey][ex];
        // a neighbour outside the mask (or the image) is read from the Tp buffer, which
        // holds what earlier levels left there, as the reference does
        const float tpLx = hasL ? sT[ey][ex - 1].x : ld0(Tp, pos - 1).x;
        const float tpUy = hasU ? sT[ey - 1][ex].y : ld0(Tp, pos - s).y;
        float dx, dy;
        if (MASKED) {
            if (hasL && hasR) dx = tp0.x - tpLx;
            else if (!hasR) dx = -tpLx;
            else dx = tp0.x;
            if (hasU && hasD) dy = tp0.y - tpUy;
            else if (!hasD) dy = -tpUy;
            else dy = tp0.y;
        } else {
            if (hasL) dx = tp0.x - tpLx;
            else if (x == w - 1) dx = -tpLx;
            else dx = tp0.x;
            if (hasU) dy = tp0.y - tpUy;
            else if (y == h - 1) dy = -tpUy;
            else dy = tp0.y;
        }
        const float etau = (a * a + b * b + 2 * c * c + (a + c) * (a + c) + (b + c) * (b + c)) * aa;
        const float u_pos = u_[pos];
        const float un = thresh_core(dx + dy, ip.tau, etau, u_pos, u[pos], Iu[pos], Iz[pos], ip.lambda);
        us[pos] = un;
        const float4 qc4 = sQ[ey][ex];
        float qcx = 0.f, qcw = 0.f, qcy = 0.f, qcz = 0.f;
        if (hasR) {
            qcx = qc4.x;
            qcw = qc4.w;
        }
        if (hasD) {
            qcy = qc4.y;
            qcz = qc4.z;
        }
        float qwx = 0.f, qww = 0.f, qny = 0.f, qnz = 0.f;
        if (hasL) {
            const float4 ql = sQ[ey][ex - 1];
            qwx = ql.x;
            qww = ql.w;
        }
        if (hasU) {
            const float4 qu = sQ[ey - 1][ex];
            qny = qu.y;
            qnz = qu.z;
        }
        const float dqx = (qcx - qwx) + (qcz - qnz);
        const float dqy = (qcw - qww) + (qcy - qny);
        const float2 pp = sP[ey][ex];
        const float tx = a * pp.x + c * pp.y;
        const float tyy = c * pp.x + b * pp.y;
        const float e1 = aa * (b * b + c * c) + 4 * ip.alpha0 * ip.alpha0;
        const float e2 = aa * (a * a + c * c) + 4 * ip.alpha0 * ip.alpha0;
        const float2 vo = v_[pos];
        float vx, vy;
        vx = vo.x + ((e1 != 0.f) ? ip.tau / e1 : ip.tau) * (ip.alpha1 * tx + ip.alpha0 * dqx);
        if (e2 != 0.f) {
            vy = vo.y + (ip.tau / e2) * (ip.alpha1 * tyy + ip.alpha0 * dqy);
        } else {
            vx = vo.x + ip.tau * (ip.alpha1 * tx + ip.alpha0 * dqx);
            vy = 0.f;
        }
        u_s[pos] = un + ip.mu * (un - u_pos);
        v_s[pos] = make_float2(vx + ip.mu * (vx - vo.x), vy + ip.mu * (vy - vo.y));
    }
}

}  // namespace

// ------------------------------------------------------------------ launchers

void launch_fill_zero(void *ptr, size_t bytes, hipStream_t st)
{
    const size_t n4 = bytes / 16;
    if (n4 == 0) return;
    const int threads = 256;
    size_t blocks = (n4 + threads - 1) / threads;
    if (blocks > 8192) blocks = 8192;
    k_fill_zero<<<dim3((unsigned)blocks), dim3(threads), 0, st>>>((float4 *)ptr, n4);
}

void launch_cv8u(const uint8_t *in, int ip, float *out, Dims d, hipStream_t st)
{
    k_cv8u<<<grid_for(d.w, d.h), block2d(), 0, st>>>(in, ip, out, d.w, d.h, d.s);
}

void launch_cv8uc3(const uint8_t *in, int ip, float *out, Dims d, hipStream_t st)
{
    k_cv8uc3<<<grid_for(d.w, d.h), block2d(), 0, st>>>(in, ip, out, d.w, d.h, d.s);
}

void launch_warp(const float *src, const float2 *uv, float *out, Dims d, hipStream_t st)
{
    k_warp<<<grid_for(d.w, d.h), block2d(), 0, st>>>(src, uv, out, d.w, d.h, d.s);
}

void launch_downscale2(const float *a, const float *b, Dims f, float *oa, float *ob, Dims c, hipStream_t st)
{
    k_downscale2<<<grid_for(c.w, c.h), block2d(), 0, st>>>(a, b, f.w, f.h, f.s, oa, ob, c.w, c.h, c.s);
}

void launch_downscale_f2(const float2 *src, Dims f, float2 *out, Dims c, hipStream_t st)
{
    k_downscale_f2<<<grid_for(c.w, c.h), block2d(), 0, st>>>(src, f.w, f.h, f.s, out, c.w, c.h, c.s);
}

void launch_downscale_nn(const float *src, Dims f, float *out, Dims c, hipStream_t st)
{
    k_downscale_nn<<<grid_for(c.w, c.h), block2d(), 0, st>>>(src, f.w, f.h, f.s, out, c.w, c.h, c.s);
}

void launch_flowvec(const float *u, const float2 *tv, float2 *uv, Dims d, hipStream_t st)
{
    k_flowvec<<<grid_for(d.w, d.h), block2d(), 0, st>>>(u, tv, uv, d.w, d.h, d.s);
}

void launch_tensor(const float *I0, const float *mask, float beta, float gamma, float alpha0, float alpha1,
                   int xlimit, float *a, float *b, float *c, Dims d, hipStream_t st)
{
    (void)alpha0;
    (void)alpha1;
    k_tensor<<<grid_for(d.w, d.h), block2d(), 0, st>>>(I0, mask, beta, gamma, xlimit, a, b, c, d.w, d.h, d.s);
}

void launch_warp_prep(const float2 *uv, const float2 *tv, const float *I1, const float *u, float2 *tv2,
                      float *i1w, float *u_last, float2 *v_, Dims d, hipStream_t st)
{
    k_warp_prep<<<grid_for(d.w, d.h), block2d(), 0, st>>>(uv, tv, I1, u, tv2, i1w, u_last, v_, d.w, d.h, d.s);
}

void launch_deriv(const float *I0, const float *I1w, const float2 *tv, float *Iu, float *Iz, Dims d,
                  hipStream_t st)
{
    k_deriv<<<grid_for(d.w, d.h), block2d(), 0, st>>>(I0, I1w, tv, Iu, Iz, d.w, d.h, d.s);
}

void launch_dual_tp(int masked, const float *mask, const float *u_, const float2 *v_, const float *a,
                    const float *b, const float *c, const IterParams &ip, float2 *p, float4 *q, float2 *Tp,
                    Dims d, hipStream_t st)
{
    if (masked)
        k_dual_tp<true><<<grid_for(d.w, d.h), block2d(), 0, st>>>(mask, u_, v_, a, b, c, ip.k1, ip.k0, ip.first, p,
                                                                 q, Tp, d.w, d.h, d.s);
    else
        k_dual_tp<false><<<grid_for(d.w, d.h), block2d(), 0, st>>>(mask, u_, v_, a, b, c, ip.k1, ip.k0, ip.first,
                                                                  p, q, Tp, d.w, d.h, d.s);
}

void launch_thresh_primal(int masked, const float *mask, const float2 *Tp, const float *u_, const float2 *v_,
                          const float *Iu, const float *Iz, const float *a, const float *b, const float *c,
                          const float2 *p, const float4 *q, const float *u, const IterParams &ip, float *us,
                          float *u_s, float2 *v_s, Dims d, hipStream_t st)
{
    if (masked)
        k_thresh_primal<true><<<grid_for(d.w, d.h), block2d(), 0, st>>>(
            mask, Tp, u_, v_, Iu, Iz, a, b, c, p, q, u, ip.tau, ip.mu, ip.lambda, ip.alpha0, ip.alpha1, us, u_s,
            v_s, d.w, d.h, d.s);
    else
        k_thresh_primal<false><<<grid_for(d.w, d.h), block2d(), 0, st>>>(
            mask, Tp, u_, v_, Iu, Iz, a, b, c, p, q, u, ip.tau, ip.mu, ip.lambda, ip.alpha0, ip.alpha1, us, u_s,
            v_s, d.w, d.h, d.s);
}

void launch_median_update(const float *u, const float *u_last, const float2 *tv2, float limit, float *out_u,
                          float *u_, float2 *warpUV, Dims d, hipStream_t st)
{
    k_median_update<<<grid_for(d.w, d.h), block2d(), 0, st>>>(u, u_last, tv2, limit, out_u, u_, warpUV, d.w, d.h,
                                                             d.s);
}

void launch_update(float *u, const float *u_last, const float2 *tv2, float limit, float *u_, float2 *warpUV,
                   Dims d, hipStream_t st)
{
    k_update<<<grid_for(d.w, d.h), block2d(), 0, st>>>(u, u_last, tv2, limit, u_, warpUV, d.w, d.h, d.s);
}

void launch_upscale3(const float *u, const float *u_, const float2 *uv, Dims c, float scale, float *ou,
                     float *ou_, float2 *ouv, Dims f, hipStream_t st)
{
    k_upscale3<<<grid_for(f.w, f.h), block2d(), 0, st>>>(u, u_, uv, c.w, c.h, c.s, scale, ou, ou_, ouv, f.w, f.h,
                                                        f.s);
}

void launch_convert_kb(const float2 *uv, double focalx, double focaly, double cx, double cy, double d1, double d2,
                       double d3, double d4, double t1, double t2, double t3, float *X, float *depth, Dims d,
                       hipStream_t st)
{
    k_convert_kb<<<grid_for(d.w, d.h), block2d(), 0, st>>>(uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3,
                                                          X, depth, d.w, d.h, d.s);
}


void launch_tgv_iter(int masked, const float *mask, const float *u_, const float2 *v_, const float *a,
                     const float *b, const float *c, const float *Iu, const float *Iz, const float *u,
                     const float2 *p_in, const float4 *q_in, float2 *p_out, float4 *q_out, float2 *Tp, float *us,
                     float *u_s, float2 *v_s, const IterParams &ip, Dims d, hipStream_t st)
{
    const dim3 grid((d.w + FTX - 1) / FTX, (d.h + FTY - 1) / FTY);
    if (masked)
        k_tgv_iter<true><<<grid, block2d(), 0, st>>>(mask, u_, v_, a, b, c, Iu, Iz, u, p_in, q_in, p_out, q_out, Tp,
                                                    us, u_s, v_s, ip, d.w, d.h, d.s);
    else
        k_tgv_iter<false><<<grid, block2d(), 0, st>>>(mask, u_, v_, a, b, c, Iu, Iz, u, p_in, q_in, p_out, q_out,
                                                     Tp, us, u_s, v_s, ip, d.w, d.h, d.s);
}

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) dst[i] = src[i];
}

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n4 = bytes / 16;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_copy<<<dim3((unsigned)blocks), dim3(256), 0, st>>>((const float4 *)src, (float4 *)dst, n4);
}

}  // namespace vfs
