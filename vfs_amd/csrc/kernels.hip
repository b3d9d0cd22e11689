// Host launchers of libvfs's kernels: launch_tick (the op kernels, instantiated per storage
// type and kernel family in tgv_inst.hip) and the single-plane kernels used at load time
// and for the outputs.
#include "tgv_device.h"

namespace vfs {

// Op launches, one per (storage type, kernel family): tgv_inst.hip built with
// VFS_INST_STORAGE (0 float, 1 __half, 2 __hip_bfloat16) and VFS_INST_K (1: k_tick, 0:
// k_iterp, else k_iterk with that many iterations).
#define VFS_INST_DECL(sfx) \
    void launch_ops_##sfx(const PairDesc *desc, const TickArgs &args, int shape, hipStream_t st);
VFS_INST_DECL(f32_k1)
VFS_INST_DECL(f32_k2)
VFS_INST_DECL(f32_k4)
VFS_INST_DECL(f32_k8)
VFS_INST_DECL(f16_k1)
VFS_INST_DECL(f16_k2)
VFS_INST_DECL(f16_k4)
VFS_INST_DECL(f16_k8)
VFS_INST_DECL(bf16_k1)
VFS_INST_DECL(bf16_k2)
VFS_INST_DECL(bf16_k4)
VFS_INST_DECL(bf16_k8)
VFS_INST_DECL(f32_k0)
VFS_INST_DECL(f16_k0)
VFS_INST_DECL(bf16_k0)
#define VFS_CAP_DECL(sfx) int persist_capacity_##sfx(int shape);
VFS_CAP_DECL(f32_k0)
VFS_CAP_DECL(f32_k2)
VFS_CAP_DECL(f32_k4)
VFS_CAP_DECL(f32_k8)
VFS_CAP_DECL(f16_k0)
VFS_CAP_DECL(f16_k2)
VFS_CAP_DECL(f16_k4)
VFS_CAP_DECL(f16_k8)
VFS_CAP_DECL(bf16_k0)
VFS_CAP_DECL(bf16_k2)
VFS_CAP_DECL(bf16_k4)
VFS_CAP_DECL(bf16_k8)
#undef VFS_CAP_DECL

namespace {

// ------------------------------------------------------------------ single-plane kernels

// 2-D launch layout of the single-plane kernels: 64 x 4 work-items per workgroup.
constexpr int PBX = 64, PBY = 4;

__device__ __forceinline__ bool pixel(int &x, int &y, int w, int h)
{
    x = blockIdx.x * PBX + threadIdx.x;
    y = blockIdx.y * PBY + threadIdx.y;
    return x < w && y < h;
}
inline dim3 grid_for(int w, int h) { return dim3((w + PBX - 1) / PBX, (h + PBY - 1) / PBY); }
inline dim3 block2d() { return dim3(PBX, PBY); }

__global__ void k_fill_zero(float4 *__restrict__ p, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// imageConverter.cu:3-16
__global__ void k_cv8u(const uint8_t *__restrict__ in, int ip, float *__restrict__ out, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    out[(size_t)y * s + x] = (float)in[(size_t)y * ip + x] / 256.0f;
}

// imageConverter.cu:26-42
__global__ void k_cv8uc3(const uint8_t *__restrict__ in, int ip, float *__restrict__ out, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const uint8_t *px = in + (size_t)y * ip + 3 * (size_t)x;
    const float g = ((float)px[0] + (float)px[1] + (float)px[2]) / 3;
    out[(size_t)y * s + x] = g / 256.0f;
}

// downscale.cu:98-122 (trajectory pyramid, stereotgv.cu:187-194)
__global__ void k_downscale_f2(const float2 *__restrict__ src, int W, int H, int S, float2 *__restrict__ out, int w,
                               int h, int s)
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
__global__ void k_downscale_nn(const float *__restrict__ src, int W, int H, int S, float *__restrict__ out, int w,
                               int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const long fx = ((2L * x + 1) * W) / (2L * w);
    const long fy = ((2L * y + 1) * H) / (2L * h);
    out[(size_t)y * s + x] = texel(src, W, H, S, (int)fx, (int)fy);
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

// conversion.cu:3-69 in FP64: Kannala-Brandt inversion of both frames (5 Newton steps
// from theta = 0), Z from the x components, X = (x0 Z, y0 Z, Z), depth = |X|.
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

// flowToHsv.cu:3-111 (ComputeColorKernel): colour coding of the flow.  The 55-entry
// Middlebury colour wheel is built from its segment lengths (red-yellow 15, yellow-green
// 6, green-cyan 4, cyan-blue 11, blue-magenta 13, magenta-red 6; entry i of a segment of
// length n ramps by floor(255 i / n)).  Positions are 1-based as in the reference, whose
// 55-entry table is read one past its end at position 55 (angle within 2.7e-6 rad of
// +pi): that position is taken as entry 0.  Output (B, G, R) per pixel.
__device__ __forceinline__ void wheel_rgb(int k, float &r, float &g, float &b)
{
    k = k % 55;
    int si = 0, start = 0, n = 15;
    if (k >= 15) { si = 1; start = 15; n = 6; }
    if (k >= 21) { si = 2; start = 21; n = 4; }
    if (k >= 25) { si = 3; start = 25; n = 11; }
    if (k >= 36) { si = 4; start = 36; n = 13; }
    if (k >= 49) { si = 5; start = 49; n = 6; }
    const int up = (255 * (k - start)) / n, down = 255 - up;
    int rr = 0, gg = 0, bb = 0;
    switch (si) {
    case 0: rr = 255; gg = up; break;
    case 1: rr = down; gg = 255; break;
    case 2: gg = 255; bb = up; break;
    case 3: gg = down; bb = 255; break;
    case 4: rr = up; bb = 255; break;
    default: rr = 255; bb = down; break;
    }
    r = (float)rr;
    g = (float)gg;
    b = (float)bb;
}

__global__ void k_flow_hsv(const float2 *__restrict__ uv, int w, int h, int s, float *__restrict__ bgr, float scale)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    const long pos = (long)y * s + x;
    const float2 f = uv[pos];
    const float du = f.x / scale, dv = f.y / scale;
    const float rad = sqrtf(du * du + dv * dv);
    const float a = atan2f(-dv, -du) / 3.14159f;
    const float fk = (a + 1.f) / 2.f * 54.f + 1.f;
    const int k0 = (int)floorf(fk);
    int k1 = k0 + 1;
    if (k1 == 56) k1 = 1;
    const float fr = fk - (float)k0;
    float r0, g0, b0, r1, g1, b1;
    wheel_rgb(k0, r0, g0, b0);
    wheel_rgb(k1, r1, g1, b1);
    float cr = (1.f - fr) * (r0 / 255.f) + fr * (r1 / 255.f);
    float cg = (1.f - fr) * (g0 / 255.f) + fr * (g1 / 255.f);
    float cb = (1.f - fr) * (b0 / 255.f) + fr * (b1 / 255.f);
    if (rad <= 1.f) {
        cr = 1.f - rad * (1.f - cr);
        cg = 1.f - rad * (1.f - cg);
        cb = 1.f - rad * (1.f - cb);
    } else {
        cr *= 0.75f;
        cg *= 0.75f;
        cb *= 0.75f;
    }
    bgr[3 * pos] = cb;
    bgr[3 * pos + 1] = cg;
    bgr[3 * pos + 2] = cr;
}

template <typename T>
__global__ void k_widen(const T *__restrict__ src, int nch, float *__restrict__ dst, int w, int h, int s)
{
    int x, y;
    if (!pixel(x, y, w, h)) return;
    for (int c = 0; c < nch; c++) {
        const size_t i = ((size_t)y * s + x) * nch + c;
        dst[i] = Cv<T>::f(src[i]);
    }
}

}  // namespace

// ------------------------------------------------------------------ launchers

#define STORAGE_DISPATCH(storage, KERNEL_CALL)                  \
    do {                                                        \
        if ((storage) == kF16) {                                \
            using T = __half;                                   \
            KERNEL_CALL;                                        \
        } else if ((storage) == kBF16) {                        \
            using T = __hip_bfloat16;                           \
            KERNEL_CALL;                                        \
        } else {                                                \
            using T = float;                                    \
            KERNEL_CALL;                                        \
        }                                                       \
    } while (0)

hipError_t launch_tick(const PairDesc *desc, const TickArgs &args, int storage, int shape, hipStream_t st)
{
    if (args.n < 1 || args.total < 1) return hipSuccess;
    if (args.e[0].kind == OP_ITERP) {
        const int n = iterp_count(args.e[0].it0);
        if (args.n != 1 || shape < 0 || shape >= kShapes || !persist_shape(shape) || n < 2 || n % 2 ||
            iterp_first(args.e[0].it0) + n > kMaxSched || args.total > kMaxPersistTiles)
            return hipErrorInvalidValue;
        static constexpr void (*kP[3])(const PairDesc *, const TickArgs &, int, hipStream_t) = {
            launch_ops_f32_k0, launch_ops_f16_k0, launch_ops_bf16_k0};
        kP[storage == kF16 ? 1 : storage == kBF16 ? 2 : 0](desc, args, shape, st);
        return hipSuccess;
    }
    if (args.e[0].kind == OP_ITERPK) {
        const int n = iterp_count(args.e[0].it0), K = iterp_block(args.e[0].it0);
        if (args.n != 1 || shape < 0 || shape >= kShapes || (K != 2 && K != 4 && K != 8) ||
            !persistk_shape(shape, K) || n < 2 || n % 2 || iterp_first(args.e[0].it0) + n > kMaxSched ||
            args.total > kMaxPersistTiles)
            return hipErrorInvalidValue;
        using Fn = void (*)(const PairDesc *, const TickArgs &, int, hipStream_t);
        static constexpr Fn kP[3][3] = {{launch_ops_f32_k2, launch_ops_f32_k4, launch_ops_f32_k8},
                                        {launch_ops_f16_k2, launch_ops_f16_k4, launch_ops_f16_k8},
                                        {launch_ops_bf16_k2, launch_ops_bf16_k4, launch_ops_bf16_k8}};
        kP[storage == kF16 ? 1 : storage == kBF16 ? 2 : 0][K == 2 ? 0 : K == 4 ? 1 : 2](desc, args, shape, st);
        return hipSuccess;
    }
    if (op_multi(args.e[0].kind)) {
        const int K = op_iterations(args.e[0].kind), sh = shape_index(shape);
        if (shape < 0 || sh >= kShapes || !iterk_ok(sh, K) ||
            ((shape & kHalfWorkgroup) && !half_workgroup_ok(sh, K)))
            return hipErrorInvalidValue;
    } else if (shape < 0 || shape >= kTickShapes) {
        return hipErrorInvalidValue;
    }
    const int K = op_iterations(args.e[0].kind);
    const int fam = op_multi(args.e[0].kind) ? (K == 2 ? 1 : K == 4 ? 2 : 3) : 0;
    using Fn = void (*)(const PairDesc *, const TickArgs &, int, hipStream_t);
    static constexpr Fn kFn[3][4] = {{launch_ops_f32_k1, launch_ops_f32_k2, launch_ops_f32_k4, launch_ops_f32_k8},
                                     {launch_ops_f16_k1, launch_ops_f16_k2, launch_ops_f16_k4, launch_ops_f16_k8},
                                     {launch_ops_bf16_k1, launch_ops_bf16_k2, launch_ops_bf16_k4, launch_ops_bf16_k8}};
    const int si = storage == kF16 ? 1 : storage == kBF16 ? 2 : 0;
    kFn[si][fam](desc, args, shape, st);
    return hipSuccess;
}

int persist_capacity(int storage, int shape, int K)
{
    if (shape < 0 || shape >= kShapes) return 0;
    const int si = storage == kF16 ? 1 : storage == kBF16 ? 2 : 0;
    if (K == 1) {
        if (!persist_shape(shape)) return 0;
        static constexpr int (*kC[3])(int) = {persist_capacity_f32_k0, persist_capacity_f16_k0, persist_capacity_bf16_k0};
        return kC[si](shape);
    }
    if ((K != 2 && K != 4 && K != 8) || !persistk_shape(shape, K)) return 0;
    static constexpr int (*kC[3][3])(int) = {{persist_capacity_f32_k2, persist_capacity_f32_k4, persist_capacity_f32_k8},
                                             {persist_capacity_f16_k2, persist_capacity_f16_k4, persist_capacity_f16_k8},
                                             {persist_capacity_bf16_k2, persist_capacity_bf16_k4, persist_capacity_bf16_k8}};
    return kC[si][K == 2 ? 0 : K == 4 ? 1 : 2](shape);
}

void launch_fill_zero(void *ptr, size_t bytes, hipStream_t st)
{
    const size_t n4 = bytes / 16;
    if (n4 == 0) return;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_fill_zero<<<dim3((unsigned)blocks), dim3(256), 0, st>>>((float4 *)ptr, n4);
}

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n4 = bytes / 16;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_copy<<<dim3((unsigned)blocks), dim3(256), 0, st>>>((const float4 *)src, (float4 *)dst, n4);
}

void launch_cv8u(const uint8_t *in, int ip, float *out, Dims d, hipStream_t st)
{
    k_cv8u<<<grid_for(d.w, d.h), block2d(), 0, st>>>(in, ip, out, d.w, d.h, d.s);
}

void launch_cv8uc3(const uint8_t *in, int ip, float *out, Dims d, hipStream_t st)
{
    k_cv8uc3<<<grid_for(d.w, d.h), block2d(), 0, st>>>(in, ip, out, d.w, d.h, d.s);
}

void launch_downscale_f2(const float2 *src, Dims f, float2 *out, Dims c, hipStream_t st)
{
    k_downscale_f2<<<grid_for(c.w, c.h), block2d(), 0, st>>>(src, f.w, f.h, f.s, out, c.w, c.h, c.s);
}

void launch_downscale_nn(const float *src, Dims f, float *out, Dims c, hipStream_t st)
{
    k_downscale_nn<<<grid_for(c.w, c.h), block2d(), 0, st>>>(src, f.w, f.h, f.s, out, c.w, c.h, c.s);
}

void launch_convert_kb(const float2 *uv, double focalx, double focaly, double cx, double cy, double d1, double d2,
                       double d3, double d4, double t1, double t2, double t3, float *X, float *depth, Dims d,
                       hipStream_t st)
{
    k_convert_kb<<<grid_for(d.w, d.h), block2d(), 0, st>>>(uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3, X,
                                                          depth, d.w, d.h, d.s);
}

void launch_flow_hsv(const float2 *uv, Dims d, float *bgr, float scale, hipStream_t st)
{
    k_flow_hsv<<<grid_for(d.w, d.h), block2d(), 0, st>>>(uv, d.w, d.h, d.s, bgr, scale);
}

void launch_widen(const void *src, int storage, int nch, float *dst, Dims d, hipStream_t st)
{
    STORAGE_DISPATCH(storage, (k_widen<T><<<grid_for(d.w, d.h), block2d(), 0, st>>>((const T *)src, nch, dst, d.w,
                                                                                     d.h, d.s)));
}

}  // namespace vfs
