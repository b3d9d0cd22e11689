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
#include <time.h>

#include <algorithm>
#include <array>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/vfs.h"
#include "vfs_engine.h"

using vfs::Dims;
using vfs::kMaxLevels;

namespace {

thread_local std::string g_create_error;

// Device-side launch set: pair descriptors (device copy + host copy for tile counts)
// and the graphs captured for the programs run on it.
struct GraphKey {
    int masked, warm_level, warps, iters, start_roles, monitor;
    float limit;
    std::array<unsigned int, kMaxLevels> tuning;  // Tuning::sig() of the program's launch choices
    int persist = 0;  // the program runs the tuning's persistent launches
    bool operator<(const GraphKey &o) const
    {
        if (persist != o.persist) return persist < o.persist;
        if (tuning != o.tuning) return tuning < o.tuning;
        if (masked != o.masked) return masked < o.masked;
        if (warm_level != o.warm_level) return warm_level < o.warm_level;
        if (warps != o.warps) return warps < o.warps;
        if (iters != o.iters) return iters < o.iters;
        if (start_roles != o.start_roles) return start_roles < o.start_roles;
        if (monitor != o.monitor) return monitor < o.monitor;
        return limit < o.limit;
    }
};

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    int final_roles = 0;
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    vfs::PairDesc *d_desc = nullptr;
    std::vector<vfs::PairDesc> h_desc;
    std::map<GraphKey, GraphEntry> graphs;
};

struct Params {
    int nLevels, nWarpIters, nSolverIters, blockWidth, storage;
    float beta, gamma, alpha0, alpha1, timestep_lambda, lambda, fScale, limitRange;
    // energy / residual records (OP_ENERGY): 0 none, 1 after the last warp of every level,
    // 2 after every warp
    int monitor = 1;
};

// One op of a solve program (see vfs::OpKind), with the scalars it needs and the tile
// shape of its launch.
struct ProgOp {
    unsigned char kind, level, masked, flags, shape;
    int roles;
    float sigma, tau, mu;  // OP_ITER
    int it0;               // OP_ITER2 / OP_ITER4 / OP_ITER8: first iteration's index in the warp
};

// Launch choices for one pyramid level: tile shape of the level's one-iteration launches
// (every op but the multi-iteration ones), whether the inner iteration fetches its
// Tp-plane neighbours up front (vfs::EF_PRELOAD_TP), how many inner iterations share one
// launch (1, 2 = vfs::OP_ITER2, 4 = vfs::OP_ITER4 or 8 = vfs::OP_ITER8) and the tile
// shape of those multi-iteration launches.
struct LevelTune {
    unsigned char shape = 0, preload = 0, per_launch = 1, mshape = 4;
    unsigned char mhalf = 0;  // multi-iteration launches in half workgroups (vfs::kHalfWorkgroup)
    // with four or eight iterations per launch: tile shape and half flag of the
    // two-iteration launches that finish an iteration count not divisible by four
    unsigned char m2shape = 4, m2half = 0;
    // a warp's iterations in one launch of persistent workgroups, in tile shape pshape,
    // wherever the solve allows it (a single pair's cold solve); otherwise the launches
    // above.  pblock 1: borders exchanged every iteration (vfs::OP_ITERP); 2, 4, 8: blocks
    // of that many iterations with recomputed halos, rings exchanged between blocks
    // (vfs::OP_ITERPK)
    unsigned char persist = 0, pshape = 18, pblock = 1;
    // tile shape of the inner iterations
    int iter_shape() const { return persist ? pshape : per_launch > 1 ? mshape : shape; }
};

struct Tuning {
    LevelTune lv[kMaxLevels];
    std::array<unsigned int, kMaxLevels> sig() const
    {
        std::array<unsigned int, kMaxLevels> k{};
        for (int l = 0; l < kMaxLevels; l++) {
            const int pl = lv[l].per_launch, code = pl == 8 ? 3 : pl == 4 ? 2 : pl == 2 ? 1 : 0;
            // tick shape 3 bits, preload 1, iterations-per-launch code 2, multi-iteration
            // shapes 5 bits each (kShapes <= 32) with their half flags
            k[l] = (unsigned int)lv[l].shape | (unsigned int)lv[l].preload << 3 | (unsigned int)code << 4 |
                   (unsigned int)lv[l].mshape << 6 | (unsigned int)lv[l].mhalf << 11 |
                   (unsigned int)lv[l].m2shape << 12 | (unsigned int)lv[l].m2half << 17 |
                   (unsigned int)lv[l].persist << 18 | (unsigned int)lv[l].pshape << 19 |
                   (unsigned int)(lv[l].pblock == 8 ? 3 : lv[l].pblock == 4 ? 2 : lv[l].pblock == 2 ? 1 : 0) << 24;
        }
        return k;
    }
};

bool multi_shape(int shape) { return shape >= 0 && shape < vfs::kShapes && vfs::multi_iter_shape(shape); }
bool valid_per_launch(int k) { return k == 1 || k == 2 || k == 4 || k == 8; }
// The op of `k` iterations per launch.
unsigned char iter_kind(int k)
{
    return (unsigned char)(k == 8 ? vfs::OP_ITER8 : k == 4 ? vfs::OP_ITER4 : k == 2 ? vfs::OP_ITER2 : vfs::OP_ITER);
}

// Step sizes (sigma, tau, mu) of the first n inner iterations of a warp
// (stereotgv.cu:449-477): tau = 1, sigma = 1 / tau at the start of every warp;
// mu = 1 / sqrt(1 + 0.7 tau timeStepLambda) while sigma < 1000 (else 1), then
// sigma /= mu and tau *= mu.
void iter_schedule(float timestep_lambda, int n, float4 *out)
{
    float tau = 1.0f;
    float sigma = 1.0f / tau;
    for (int it = 0; it < n; it++) {
        const float mu = sigma < 1000.0f ? 1.0f / sqrtf(1 + 0.7f * tau * timestep_lambda) : 1.0f;
        out[it] = make_float4(sigma, tau, mu, 0.0f);
        sigma = sigma / mu;
        tau = tau * mu;
    }
}

struct Program {
    std::vector<ProgOp> ops;
    int final_roles = 0;
    std::vector<int> mon_slots;  // energy records the program writes (vfs::mon_slot), in order
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
    int force_shape = -1;   // option "shape": -1 = per-level tuner
    bool autotune = true;   // option "autotune"
    int force_per_launch = -1;  // option "per_launch": -1 = tuner, else 1, 2, 4 or 8 wherever allowed
    int persist_opt = -1;   // option "persist": -1 = tuner, 0 = never, 1 = wherever possible
    int monitor = 1;        // option "energy": 0, 1 or 2 (Params::monitor)
    bool mask_col0[kMaxLevels] = {};  // a mask pixel lies in column 0 of the level
    unsigned long long mask_gen = 0;  // copyMaskToDevice calls since create (batches re-tune on change)
    Tuning tuning[2];       // [masked] launch choices per level
    bool tuned[2] = {false, false};
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
    // energy records (kMaxLevels x kMonWarps) followed by the partial sums of one OP_ENERGY
    vfs::EnergyPart *mon = nullptr;
    std::vector<int> mon_slots;  // records written by the last solve of this pair
    int numeric = 0;             // last solve: 0 not checked yet, 1 finite, -1 non-finite values

    // work arena
    float *arena = nullptr;
    long long plane = 0;
    // exchange words and launch bookkeeping of the persistent op (allocated on first use)
    unsigned long long *xbuf = nullptr;
    size_t xwords = 0;
    vfs::XSync *xsync = nullptr;
    // the last solve ran persistent launches whose abort flag the host has not read yet;
    // its solve form, to run it again without them if a launch gave up
    bool persist_check = false;
    int persist_masked = 0;
    Params persist_params{};  // the parameters of that solve (its re-run uses them)
    int persist_reruns = 0;  // solves run again without persistent launches
    unsigned persist_why = 0;  // XSync::why of the last launch that gave up
    int roles = vfs::kRolesInit;  // buffer roles after the last solve
    bool solved = false;  // a solve has run since initialize

    Engine eng;  // single-pair launch set
    bool desc_dirty = true;
    std::vector<void *> allocs;

    Params params() const
    {
        return Params{nLevels, nWarpIters, nSolverIters, blockWidth, storage, beta, gamma, alpha0, alpha1,
                      timestep_lambda, lambda, fScale, limitRange, monitor};
    }
    vfs::PairDesc desc() const
    {
        vfs::PairDesc d;
        memset(&d, 0, sizeof d);  // descriptors are compared bytewise
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
        iter_schedule(timestep_lambda, vfs::kMaxSched, d.sched);
        // k1 = alpha1 sigma / 3 of each iteration, rounded as the one-iteration kernel
        // forms it (float product, correctly rounded division)
        for (int i = 0; i < vfs::kMaxSched; i++) d.sched[i].w = alpha1 * d.sched[i].x / 3.0f;
        d.mon_rec = mon;
        d.mon_part = mon ? mon + kMaxLevels * vfs::kMonWarps : nullptr;
        d.xbuf = xbuf;
        d.xsync = xsync;
        return d;
    }
};

struct vfs_batch {
    Engine eng;
    std::vector<vfs_solver *> members;
    std::string err;
    Tuning tuning[2];  // [masked] launch choices of the batch's launches
    bool tuned[2] = {false, false};
    // members' mask generations when the launch choices were made: a new mask can forbid
    // multi-iteration launches at a level (a mask pixel in column 0), so it re-tunes
    std::vector<unsigned long long> mask_gen;
    // launch choices fixed by vfs_batch_set_level_tuning ([masked][level]); they replace
    // the tuner's wherever every member allows them
    bool fixed[2][kMaxLevels] = {};
    LevelTune fixed_lv[2][kMaxLevels];
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
    for (auto &kv : e.graphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    e.graphs.clear();
}

void engine_release(Engine &e)
{
    engine_drop_graphs(e);
    if (e.d_desc) (void)hipFree(e.d_desc);
    e.d_desc = nullptr;
    e.h_desc.clear();
}

// Uploads the pair descriptors of the launch set.
hipError_t engine_setup(Engine &e, const std::vector<vfs::PairDesc> &descs)
{
    engine_release(e);
    hipError_t r = hipMalloc(&e.d_desc, sizeof(vfs::PairDesc) * descs.size());
    if (r != hipSuccess) return r;
    r = hipMemcpyAsync(e.d_desc, descs.data(), sizeof(vfs::PairDesc) * descs.size(), hipMemcpyHostToDevice,
                       e.stream);
    if (r != hipSuccess) return r;
    e.h_desc = descs;
    return hipStreamSynchronize(e.stream);
}

void push_op(Program &pr, int kind, int level, int masked, int roles)
{
    ProgOp o{};
    o.kind = (unsigned char)kind;
    o.level = (unsigned char)level;
    o.masked = (unsigned char)masked;
    o.roles = roles;
    pr.ops.push_back(o);
}

// One level of the solve from the tensor to the end of the last warp
// (stereotgv.cu:409-494, masked; :275-364 unmasked).  With `sync_spare`, each warp's
// first multi-iteration op is preceded by OP_SYNC (warm solves above level 0).  With
// `pblock` 1, a warp's iterations run in OP_ITERP ops (an even count each, within the
// schedule table), with 2, 4 or 8 in OP_ITERPK ops (the same counts, in blocks of pblock),
// and the rest as without it.
void push_level(Program &pr, const Params &P, int masked, int l, int warps, int iters, int &roles, int per_launch,
                bool sync_spare = false, int pblock = 0)
{
    std::vector<float4> sched((size_t)std::max(iters, 1));
    iter_schedule(P.timestep_lambda, iters, sched.data());
    // the tensor and the first warp's preparation run in one op (EF_PREP); later warps'
    // preparations run with the previous warp's end (OP_ENDPREP)
    push_op(pr, vfs::OP_TENSOR, l, masked, roles);
    pr.ops.back().flags = vfs::EF_PREP;
    for (int wi = 0; wi < warps; wi++) {
        push_op(pr, vfs::OP_DERIV, l, masked, roles);
        const size_t deriv = pr.ops.size() - 1;
        for (int it = 0; it < iters;) {
            ProgOp o{};
            o.level = (unsigned char)l;
            o.masked = (unsigned char)masked;
            o.roles = roles;
            // K iterations in one op while they fit the iteration count and the
            // descriptors' schedule table; the rest one or two at a time
            int k = 1;
            for (int c : {8, 4, 2})
                if (k == 1 && c <= per_launch && it + c <= iters && it + c <= vfs::kMaxSched) k = c;
            const int kp = std::min(iters - it, vfs::kMaxSched - it) & ~1;
            if (pblock > 0 && !sync_spare && kp >= 2) {
                o.kind = pblock > 1 ? vfs::OP_ITERPK : vfs::OP_ITERP;
                o.it0 = (kp << 16) | (pblock > 1 ? pblock << 8 : 0) | it;
                o.flags = (unsigned char)((it == 0 ? vfs::EF_FIRST : 0) | (it + kp == iters ? vfs::EF_WRITE_TP_LAST : 0));
                roles = vfs::rrotate_iter(roles);
                pr.ops.push_back(o);
                it += kp;
                continue;
            }
            if (k > 1) {
                // the spare buffers' copy beyond the level's extent runs with the
                // derivatives (EF_SYNC), which read none of u, u_, v_
                if (sync_spare && it == 0) pr.ops[deriv].flags |= vfs::EF_SYNC;
                o.kind = iter_kind(k);
                o.it0 = it;
                o.flags = (unsigned char)((it == 0 ? vfs::EF_FIRST : 0) | (it + k == iters ? vfs::EF_WRITE_TP_LAST : 0));
                roles = vfs::rrotate_iter(roles);
            } else {
                o.kind = vfs::OP_ITER;
                o.sigma = sched[it].x;  // the kernels form alpha1 sigma / 3 and alpha0 sigma / 2
                o.tau = sched[it].y;
                o.mu = sched[it].z;
                o.flags = (unsigned char)((it == 0 ? vfs::EF_FIRST : 0) | (it == iters - 1 ? vfs::EF_WRITE_TP : 0));
                roles = vfs::rswap_iter(roles);
            }
            pr.ops.push_back(o);
            it += k;
        }
        // energy / residual record of the warp (before its end, which overwrites the other
        // u buffer and clears v_)
        // (the end's extra workgroup adds the partials into the record: EF_EFINAL)
        const bool mon = P.monitor == 2 || (P.monitor == 1 && wi + 1 == warps);
        if (mon) {
            push_op(pr, vfs::OP_ENERGY, l, masked, roles);
            pr.mon_slots.push_back(vfs::mon_slot(l, wi));
        }
        push_op(pr, wi + 1 < warps ? vfs::OP_ENDPREP : vfs::OP_END, l, masked, roles);
        if (mon) {
            pr.ops.back().flags |= vfs::EF_EFINAL;
            pr.ops.back().it0 = wi;
        }
        if (masked) roles = vfs::rswap(roles, vfs::G_U);
    }
}

// Sets every op's tile shape, and the Tp preload of the inner iterations, per level.
void apply_tuning(Program &pr, const Tuning &tn)
{
    for (ProgOp &o : pr.ops) {
        const LevelTune &lt = tn.lv[vfs::op_tile_level(o.kind, o.level)];
        if (o.kind == vfs::OP_ITERP || o.kind == vfs::OP_ITERPK) {
            o.shape = lt.pshape;
            continue;
        }
        const bool multi = vfs::op_multi(o.kind);
        const bool rest = o.kind == vfs::OP_ITER2 && lt.per_launch >= 4;  // the remainder of a warp
        o.shape = multi ? (rest ? lt.m2shape : lt.mshape) : lt.shape;
        o.flags &= (unsigned char)~vfs::EF_PRELOAD_TP;
        if (o.kind == vfs::OP_ITER && lt.preload) o.flags |= vfs::EF_PRELOAD_TP;
        // multi-iteration ops exist for some shapes only
        if (multi && !vfs::iterk_ok(o.shape, vfs::op_iterations(o.kind))) o.shape = 4;
        if (multi && (rest ? lt.m2half : lt.mhalf) && vfs::half_workgroup_ok(o.shape, vfs::op_iterations(o.kind)))
            o.shape = (unsigned char)(o.shape | vfs::kHalfWorkgroup);
    }
    // A warp's first inner op computes the warp's derivatives at its staged positions
    // (EF_DERIV) in place of the OP_DERIV op before it, where that op stages a halo (a
    // multi-iteration or persistent op) with at most two positions per work-item (three in
    // the whole-warp persistent op, which stages once per warp), and the OP_DERIV op does
    // not also sync the spare buffers (EF_SYNC).
    for (size_t i = 0; i + 1 < pr.ops.size(); i++) {
        const ProgOp &d = pr.ops[i];
        ProgOp &f = pr.ops[i + 1];
        if (d.kind != vfs::OP_DERIV || (d.flags & vfs::EF_SYNC)) continue;
        int K = 0;
        if (vfs::op_multi(f.kind)) K = vfs::op_iterations(f.kind);
        else if (f.kind == vfs::OP_ITERP) K = 1;
        else if (f.kind == vfs::OP_ITERPK) K = vfs::iterp_block(f.it0);
        if (K == 0 || vfs::staged_per_item(vfs::shape_index(f.shape), K, (f.shape & vfs::kHalfWorkgroup) != 0) >
                          (f.kind == vfs::OP_ITERP ? 3 : 2))
            continue;
        f.flags |= vfs::EF_DERIV;
        pr.ops.erase(pr.ops.begin() + (long)i);
    }
}

// The solve program.  Cold (warm_level < 0): stereotgv.cu:389-515 (masked) or :256-387,
// from zeroed buffers, all levels with the configured warps and iterations.
// Warm (warm_level >= 0): the previous solve's u, u_ and warpUV are downscaled (x 1/fScale
// per level) to warm_level and the solve runs from there to level 0 with `warps` and
// `iters`; nothing is zeroed.  `persist`: the levels whose tuning says so run their
// warps' iterations in persistent launches (single-pair cold solves only).
Program build_program(const Params &P, int masked, int warm_level, int warps, int iters, int start_roles,
                      const Tuning &tn, bool persist = false)
{
    Program pr;
    int roles = vfs::kRolesInit;
    if (warm_level < 0) {
        push_op(pr, vfs::OP_ZERO, 0, masked, 0);
        push_op(pr, vfs::OP_CALIB, 0, masked, 0);
        for (int l = 1; l < P.nLevels; l++) push_op(pr, vfs::OP_DOWN, l, masked, 0);
        for (int l = P.nLevels - 1; l >= 0; l--) {
            if (l == P.nLevels - 1) push_op(pr, vfs::OP_CINIT, l, masked, roles);
            push_level(pr, P, masked, l, P.nWarpIters, P.nSolverIters, roles, tn.lv[l].per_launch, false,
                       persist && tn.lv[l].persist ? tn.lv[l].pblock : 0);
            if (l > 0) {
                push_op(pr, vfs::OP_UP, l, masked, roles);
                roles = vfs::rswap(vfs::rswap(roles, vfs::G_U), vfs::G_UB) ^ vfs::RB_UV;
            }
        }
    } else {
        roles = start_roles;
        push_op(pr, vfs::OP_CALIB, 0, masked, roles);
        for (int l = 1; l <= warm_level; l++) push_op(pr, vfs::OP_DOWN, l, masked, roles);
        for (int l = 1; l <= warm_level; l++) {
            push_op(pr, vfs::OP_WINIT, l, masked, roles);
            roles = vfs::rswap(vfs::rswap(roles, vfs::G_U), vfs::G_UB) ^ vfs::RB_UV;
        }
        for (int l = warm_level; l >= 0; l--) {
            // Above level 0 the spare buffers first take the current values beyond the
            // level's extent (OP_SYNC): a multi-iteration launch makes the spare buffers
            // current there, where a warm solve's buffers still hold the previous solve's
            // level-0 data (a cold solve's are all zero).
            push_level(pr, P, masked, l, warps, iters, roles, tn.lv[l].per_launch, l > 0);
            if (l > 0) {
                push_op(pr, vfs::OP_UP, l, masked, roles);
                roles = vfs::rswap(vfs::rswap(roles, vfs::G_U), vfs::G_UB) ^ vfs::RB_UV;
            }
        }
    }
    pr.final_roles = roles;
    apply_tuning(pr, tn);
    return pr;
}

vfs::TickArgs tick_header(const Params &P)
{
    vfs::TickArgs a;
    memset(&a, 0, sizeof a);
    a.lambda = P.lambda;
    a.alpha0 = P.alpha0;
    a.alpha1 = P.alpha1;
    a.beta = P.beta;
    a.gamma = P.gamma;
    a.fscale = P.fScale;
    a.scale = 1.0f / P.fScale;
    a.limit = P.limitRange;
    a.xblock = P.blockWidth;
    return a;
}

void set_entry(vfs::TickEntry &e, const ProgOp &o, int pair, const vfs::PairDesc &d, int tile_start)
{
    const int tl = vfs::op_tile_level(o.kind, o.level);
    e.arena = d.arena;
    e.mask = d.mask[tl];
    e.plane = (int)d.plane;  // < 2^31 (checked at initialize)
    e.w = (unsigned short)(vfs::op_flat(o.kind) ? d.s[0] : d.w[tl]);
    e.h = (unsigned short)d.h[tl];
    e.s = (unsigned short)d.s[tl];
    e.pair = (unsigned short)pair;
    e.tile_start = tile_start;
    e.kind = o.kind;
    e.level = o.level;
    e.masked = o.masked;
    e.flags = o.flags;
    e.roles = (unsigned int)o.roles;
    e.sigma = o.sigma;
    e.tau = o.tau;
    e.mu = o.mu;
    e.it0 = o.it0;
}

// Issues every op of `pr` for every pair of the launch set: one k_tick launch per op per
// group of up to kMaxEntries pairs.  Stops at the first launch that has no kernel.
hipError_t issue_program(const Engine &e, const Params &P, const Program &pr)
{
    vfs::TickArgs a = tick_header(P);
    const int n = (int)e.h_desc.size();
    for (const ProgOp &o : pr.ops) {
        for (int p0 = 0; p0 < n; p0 += vfs::kMaxEntries) {
            const int cnt = std::min(vfs::kMaxEntries, n - p0);
            int acc = 0;
            for (int k = 0; k < cnt; k++) {
                const vfs::PairDesc &d = e.h_desc[p0 + k];
                set_entry(a.e[k], o, p0 + k, d, acc);
                acc += vfs::op_tiles(o.kind, o.level, d.w, d.h, d.s, o.shape, o.flags);
            }
            a.n = cnt;
            a.total = acc;
            const hipError_t r = vfs::launch_tick(e.d_desc, a, P.storage, o.shape, e.stream);
            if (r != hipSuccess) return r;
        }
    }
    return hipSuccess;
}

// Runs `pr` on the engine's stream: replays (capturing first if needed) the graph cached
// under `key`, or launches directly.
int engine_run(Engine &e, const Params &P, const Program &pr, const GraphKey &key, bool use_graph, std::string *err)
{
    auto ferr = [&](const char *what, hipError_t r) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(r));
        *err = buf;
        return VFS_ERR_DEVICE;
    };
    if (!use_graph) {
        hipError_t r = issue_program(e, P, pr);
        if (r == hipSuccess) r = hipGetLastError();
        return r == hipSuccess ? VFS_OK : ferr("kernel launch", r);
    }
    auto it = e.graphs.find(key);
    if (it == e.graphs.end()) {
        hipError_t r = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal);
        if (r != hipSuccess) return ferr("hipStreamBeginCapture", r);
        const hipError_t ri = issue_program(e, P, pr);
        hipGraph_t g = nullptr;
        r = hipStreamEndCapture(e.stream, &g);
        if (ri != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return ferr("kernel launch", ri);
        }
        if (r != hipSuccess) return ferr("hipStreamEndCapture", r);
        GraphEntry ge;
        r = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (r != hipSuccess) return ferr("hipGraphInstantiate", r);
        ge.final_roles = pr.final_roles;
        it = e.graphs.emplace(key, ge).first;
    }
    hipError_t r = hipGraphLaunch(it->second.exec, e.stream);
    if (r != hipSuccess) return ferr("hipGraphLaunch", r);
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
    s->xbuf = nullptr;
    s->xwords = 0;
    s->xsync = nullptr;
    s->persist_check = false;
    s->mon = nullptr;
    s->mon_slots.clear();
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
    hipError_t r = engine_setup(s->eng, d);
    if (r != hipSuccess) return fail(s, VFS_ERR_DEVICE, "engine setup: %s", hipGetErrorString(r));
    s->desc_dirty = false;
    return VFS_OK;
}

// ---- persistent launches (vfs::OP_ITERP, and vfs::OP_ITERPK in blocks of K = 2, 4, 8)

// Workgroups of the persistent op that the device surely holds at once, per (device,
// storage, shape, K): one per compute unit (where a compute unit holds one at all).  The
// occupancy calculation can promise more for small workgroups, but whether the hardware
// places that many depends on how it spreads their waves over the SIMDs; one per compute
// unit does not, and the faster tilings have at most that many tiles anyway.
int persist_cap(int device, int storage, int shape, int K)
{
    static std::mutex mu;
    static std::map<std::array<int, 4>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    const std::array<int, 4> k{device, storage, shape, K};
    auto it = cache.find(k);
    if (it != cache.end()) return it->second;
    const int c = vfs::persist_capacity(storage, shape, K);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 0;
    const int r = c > 0 ? std::min(c, cus) : 0;
    cache[k] = r;
    return r;
}

// Tiles of the persistent op (K = 1: OP_ITERP, else OP_ITERPK in blocks of K) at `level`
// in tile shape `sh`, 0 where it cannot run: a shape without the op, a mask pixel in
// column 0 (as for every multi-iteration op), or more tiles than the device holds
// workgroups at once (or than kMaxPersistTiles).
int persist_tiles(const vfs_solver *s, int masked, int level, int sh, int K = 1)
{
    if (sh < 0 || sh >= vfs::kShapes) return 0;
    if (K == 1 ? !vfs::persist_shape(sh) : !vfs::persistk_shape(sh, K)) return 0;
    if (s->lv[level].w < 2 || (masked && s->mask_col0[level])) return 0;
    const vfs::ShapeDim d = vfs::kShape[sh];
    const int n = ((s->lv[level].w + d.tw - 1) / d.tw) * ((s->lv[level].h + d.th - 1) / d.th);
    if (n > vfs::kMaxPersistTiles || n > persist_cap(s->device, s->storage, sh, K)) return 0;
    return n;
}

// The persistent op's shape at `level` when the option forces it: the one with the most
// tiles the device holds at once; -1 if none.
int persist_default_shape(const vfs_solver *s, int masked, int level)
{
    int best = -1, best_n = 0;
    for (int sh = 0; sh < vfs::kShapes; sh++) {
        const int n = persist_tiles(s, masked, level, sh);
        if (n > best_n) best = sh, best_n = n;
    }
    return best;
}

// Exchange words (both parities) of a persistent launch of `tiles` tiles in shape `sh`.
size_t persist_words_of(int tiles, int sh, int K = 1)
{
    return (size_t)tiles * 2 * (size_t)(K == 1 ? vfs::persist_words(sh) : vfs::persistk_words(sh));
}

// Allocates the launch bookkeeping and at least `words` exchange words.  The tags in the
// words start at launch counter 1, so zeroed words never match.
int ensure_xbuf(vfs_solver *s, size_t words)
{
    if (!s->xsync) {
        int rc = dev_alloc(s, (void **)&s->xsync, sizeof(vfs::XSync));
        if (rc) return rc;
        const vfs::XSync init{1u, 0u, 0u, 0u, 0u, {0u, 0u, 0u}};
        HIPCHK(s, hipMemcpyAsync(s->xsync, &init, sizeof init, hipMemcpyHostToDevice, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        s->desc_dirty = true;
    }
    if (words > s->xwords) {
        HIPCHK(s, hipStreamSynchronize(s->stream));
        if (s->xbuf) {
            s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)s->xbuf));
            HIPCHK(s, hipFree(s->xbuf));
            s->xbuf = nullptr;
            s->xwords = 0;
        }
        int rc = dev_alloc(s, (void **)&s->xbuf, words * sizeof(unsigned long long));
        if (rc) return rc;
        s->xwords = words;
        s->desc_dirty = true;
    }
    return VFS_OK;
}

// Reads the abort flag of the persistent launches (after the stream's work) and clears
// it; true if a launch gave up.
int persist_aborted(vfs_solver *s, bool *aborted)
{
    *aborted = false;
    if (!s->xsync) return VFS_OK;
    vfs::XSync x;
    HIPCHK(s, hipMemcpyAsync(&x, s->xsync, sizeof x, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (!x.abort) return VFS_OK;
    *aborted = true;
    s->persist_why = x.why;
    const unsigned zero = 0;
    HIPCHK(s, hipMemcpyAsync(&s->xsync->abort, &zero, sizeof zero, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return VFS_OK;
}

// ---- per-level launch tuner
//
// The inner iteration (the dominant op) of each level is timed in every tile shape, with
// and without the Tp preload, on a zeroed scratch arena (the pair's own state is not
// touched), and the fastest choice is kept.  Every op of the level then launches in that
// shape.  Results are shared by all solvers of the process on the same device with the
// same storage, solve form and level geometry.  Results do not depend on the choice: the
// kernels compute every pixel identically in every shape.
struct TuneKey {
    int device, storage, masked, w, h, s, multi_ok;
    int persist_iters;  // iterations of the persistent candidates (0: none)
    bool operator<(const TuneKey &o) const
    {
        const int a[8] = {device, storage, masked, w, h, s, multi_ok, persist_iters};
        const int b[8] = {o.device, o.storage, o.masked, o.w, o.h, o.s, o.multi_ok, o.persist_iters};
        return std::lexicographical_compare(a, a + 8, b, b + 8);
    }
};

std::mutex g_tune_mu;
std::map<TuneKey, LevelTune> g_tune;

// Without tuning: 64x8 tiles; the Tp preload below level 0, where launches are latency-bound.
LevelTune default_tune(int level)
{
    LevelTune t;
    t.shape = 0;
    t.preload = level > 0 ? 1 : 0;
    return t;
}

// Persistent candidates of a single pair's tuning: the iterations of a warp (within the
// schedule table; a launch takes the even count below it), per level, form (K = 1, 2, 4, 8)
// and shape the tiles of the launch (0: not possible there), and the pair's launch
// bookkeeping.
struct PersistCands {
    int iters = 0;
    std::vector<std::array<std::array<int, vfs::kShapes>, 4>> tiles;  // [level index][log2 K][shape]
    vfs::XSync *xsync = nullptr;
};
constexpr int kPersistK[4] = {1, 2, 4, 8};
// Iterations of one persistent launch (OP_ITERP, or OP_ITERPK in blocks of K with a
// shorter last block) out of `iters`: the even count.
int persist_launch_iters(int iters) { return iters & ~1; }

// Times the inner iteration of every level in `levels` in each (tile shape, Tp preload)
// on the launch set `descs` (their arenas point at scratch memory; `d_desc` is any valid
// descriptor array of at least descs.size() entries) and stores the fastest choices in tn.
// With `pc` (a single pair), the persistent op in each of its shapes is a candidate too;
// if one of its launches gave up (not every workgroup resident), none is taken.
hipError_t time_launch_choices(hipStream_t st, const vfs::PairDesc *d_desc, const std::vector<vfs::PairDesc> &descs,
                               const Params &P, int masked, const std::vector<int> &levels,
                               const std::vector<bool> &allow_multi, Tuning &tn, const PersistCands *pc = nullptr)
{
    // a short timing of every candidate, then a longer one of those within 10 % of the
    // fastest of their kind (one-iteration, multi-iteration, two-iteration launches)
    constexpr int kWarm = 3, kRounds = 2, kReps[kRounds] = {6, 20};
    constexpr float kRefine = 1.10f;
    // candidates: (shape, preload) of one iteration per launch, then two-, four- and
    // eight-iteration launches for the shapes that have them, in full or half workgroups
    // (then the persistent ops, OP_ITERP and OP_ITERPK in blocks of 2, 4 and 8, in each shape)
    constexpr int kSingle = vfs::kTickShapes * 2, kMulti = kSingle + 3 * vfs::kShapes * 2;
    constexpr int kCand = kMulti + 4 * vfs::kShapes;
    auto cand = [&](int c, int &shape, int &pre, int &per_launch, int &half) {
        if (c >= kMulti) {
            shape = (c - kMulti) % vfs::kShapes;
            pre = 0;
            half = 0;
            per_launch = pc ? persist_launch_iters(pc->iters) : 0;
        } else if (c < kSingle) {
            shape = c / 2;
            pre = c % 2;
            per_launch = 1;
            half = 0;
        } else {
            const int m = c - kSingle;
            per_launch = 2 << (m / (vfs::kShapes * 2));
            shape = (m % (vfs::kShapes * 2)) / 2;
            half = m % 2;
            pre = 0;
        }
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t r = hipEventCreate(&e0);
    if (r == hipSuccess) r = hipEventCreate(&e1);
    std::vector<float> best(levels.size() * kCand, 1e30f);  // ms per iteration
    const int n = std::min((int)descs.size(), vfs::kMaxEntries);
    for (int round = 0; round < kRounds && r == hipSuccess; round++) {
        for (size_t li = 0; li < levels.size() && r == hipSuccess; li++) {
            float fast[3] = {1e30f, 1e30f, 1e30f};  // one-iteration, multi-iteration, two-iteration
            for (int c = 0; round > 0 && c < kCand; c++) {
                const float b = best[li * kCand + c];
                fast[c < kSingle ? 0 : 1] = std::min(fast[c < kSingle ? 0 : 1], b);
                if (c >= kSingle && c < kSingle + vfs::kShapes * 2) fast[2] = std::min(fast[2], b);
            }
            for (int c = 0; c < kCand && r == hipSuccess; c++) {
                int shape, pre, per_launch, half;
                cand(c, shape, pre, per_launch, half);
                const bool persist = c >= kMulti;
                const int pk = persist ? kPersistK[(c - kMulti) / vfs::kShapes] : 0;
                if (persist && (!pc || per_launch < 2 ||
                                pc->tiles[li][(c - kMulti) / vfs::kShapes][shape] == 0 || n != 1))
                    continue;
                if (!persist && per_launch > 1 &&
                    (!vfs::iterk_ok(shape, per_launch) || !allow_multi[li] ||
                     (half && !vfs::half_workgroup_ok(shape, per_launch))))
                    continue;
                if (round > 0) {
                    const float b = best[li * kCand + c];
                    const bool near = b < kRefine * fast[c < kSingle ? 0 : 1] ||
                                      (per_launch == 2 && b < kRefine * fast[2]);
                    if (!near) continue;
                }
                ProgOp o{};
                o.kind = persist ? (unsigned char)(pk > 1 ? vfs::OP_ITERPK : vfs::OP_ITERP) : iter_kind(per_launch);
                o.level = (unsigned char)levels[li];
                o.masked = (unsigned char)masked;
                o.flags = (unsigned char)(pre ? vfs::EF_PRELOAD_TP : 0);
                o.shape = (unsigned char)(shape | (half ? vfs::kHalfWorkgroup : 0));
                o.roles = vfs::kRolesInit;
                o.sigma = 1.0f;
                o.tau = 1.0f;
                o.mu = 0.8f;
                o.it0 = persist ? (per_launch << 16) | (pk > 1 ? pk << 8 : 0) : 0;
                vfs::TickArgs a = tick_header(P);
                int acc = 0;
                for (int k = 0; k < n; k++) {
                    const vfs::PairDesc &d = descs[k];
                    set_entry(a.e[k], o, k, d, acc);
                    acc += vfs::op_tiles(o.kind, o.level, d.w, d.h, d.s, shape);
                }
                a.n = n;
                a.total = acc;
                auto launch = [&]() { return vfs::launch_tick(d_desc, a, P.storage, o.shape, st); };
                for (int k = 0; k < kWarm && r == hipSuccess; k++) r = launch();
                if (r == hipSuccess) r = hipEventRecord(e0, st);
                for (int k = 0; k < kReps[round] && r == hipSuccess; k++) r = launch();
                if (r == hipSuccess) r = hipEventRecord(e1, st);
                if (r == hipSuccess) r = hipEventSynchronize(e1);
                float ms = 0;
                if (r == hipSuccess) r = hipEventElapsedTime(&ms, e0, e1);
                if (r == hipSuccess) r = hipGetLastError();
                // ms per iteration; the longer timing replaces the short one
                best[li * kCand + c] = ms / (float)(kReps[round] * per_launch);
            }
        }
    }
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (r != hipSuccess) return r;
    // a persistent launch that gave up timed nothing: take none
    if (pc && pc->xsync) {
        vfs::XSync x;
        r = hipMemcpyAsync(&x, pc->xsync, sizeof x, hipMemcpyDeviceToHost, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        if (r != hipSuccess) return r;
        if (x.abort) {
            const unsigned zero = 0;
            r = hipMemcpyAsync(&pc->xsync->abort, &zero, sizeof zero, hipMemcpyHostToDevice, st);
            if (r == hipSuccess) r = hipStreamSynchronize(st);
            if (r != hipSuccess) return r;
            for (size_t li = 0; li < levels.size(); li++)
                for (int c = kMulti; c < kCand; c++) best[li * kCand + c] = 1e30f;
        }
    }
    for (size_t li = 0; li < levels.size(); li++) {
        const int l = levels[li];
        const LevelTune def = default_tune(l);
        LevelTune pick = def;
        // the one-iteration launch (also the shape of the level's other ops), leaving the
        // default only for a clear gain; then the fastest multi-iteration launch where it
        // is clearly faster still
        float pick_ms = best[li * kCand + def.shape * 2 + def.preload];
        for (int c = 0; c < kSingle; c++) {
            const float ms = best[li * kCand + c];
            if (ms < 0.98f * pick_ms) {
                pick.shape = (unsigned char)(c / 2);
                pick.preload = (unsigned char)(c % 2);
                pick_ms = ms;
            }
        }
        float best2 = 1e30f;  // the fastest two-iteration launch, for remainders
        int bestm = -1;       // the fastest multi-iteration launch
        for (int c = kSingle; c < kMulti; c++) {
            const float ms = best[li * kCand + c];
            int shape, pre, per_launch, half;
            cand(c, shape, pre, per_launch, half);
            if (per_launch == 2 && ms < best2) {
                best2 = ms;
                pick.m2shape = (unsigned char)shape;
                pick.m2half = (unsigned char)half;
            }
            if (bestm < 0 || ms < best[li * kCand + bestm]) bestm = c;
        }
        if (bestm >= 0 && best[li * kCand + bestm] < 0.98f * pick_ms) {
            int shape, pre, per_launch, half;
            cand(bestm, shape, pre, per_launch, half);
            pick.mshape = (unsigned char)shape;
            pick.per_launch = (unsigned char)per_launch;
            pick.mhalf = (unsigned char)half;
            pick_ms = best[li * kCand + bestm];
        }
        // the persistent op where it is clearly faster than both over a warp's iterations:
        // it runs all of them, where the launches above leave a remainder to two-iteration
        // or one-iteration launches (those stay the level's choice wherever it cannot run)
        int bestp = -1;
        for (int c = kMulti; c < kCand; c++)
            if (best[li * kCand + c] < 1e29f && (bestp < 0 || best[li * kCand + c] < best[li * kCand + bestp])) bestp = c;
        float warp_ms = pick_ms;  // per iteration, the remainder included
        if (pc && pc->iters >= 2 && pick.per_launch > 1) {
            const int w = pc->iters, rem = w % pick.per_launch;
            warp_ms = (pick_ms * (float)(w - rem) + (rem >= 2 ? best2 * (float)(rem & ~1) : 0.f) +
                       ((rem & 1) ? best[li * kCand + pick.shape * 2 + pick.preload] : 0.f)) / (float)w;
        }
        if (bestp >= 0 && best[li * kCand + bestp] < 0.98f * warp_ms) {
            pick.persist = 1;
            pick.pshape = (unsigned char)((bestp - kMulti) % vfs::kShapes);
            pick.pblock = (unsigned char)kPersistK[(bestp - kMulti) / vfs::kShapes];
        }
        tn.lv[l] = pick;
    }
    return hipSuccess;
}

// Levels where several iterations may share a launch: at least 2 columns and, in the
// masked form, no mask pixel in column 0.
std::vector<bool> multi_allowed(const vfs_solver *s, int masked, const std::vector<int> &levels)
{
    std::vector<bool> ok;
    for (int l : levels) ok.push_back(s->lv[l].w >= 2 && !(masked && s->mask_col0[l]));
    return ok;
}

int measure_levels(vfs_solver *s, int masked, const std::vector<int> &levels, Tuning &tn)
{
    // persistent candidates (a warp's iterations per launch, within the schedule table)
    PersistCands pc;
    if (s->persist_opt != 0) {
        pc.iters = std::min(s->nSolverIters, vfs::kMaxSched) & ~1;
        size_t words = 0;
        for (int l : levels) {
            std::array<std::array<int, vfs::kShapes>, 4> t{};
            for (int kb = 0; kb < 4; kb++)
                for (int sh = 0; sh < vfs::kShapes; sh++) {
                    t[kb][sh] = persist_tiles(s, masked, l, sh, kPersistK[kb]);
                    words = std::max(words, persist_words_of(t[kb][sh], sh, kPersistK[kb]));
                }
            pc.tiles.push_back(t);
        }
        if (pc.iters >= 2 && words > 0) {
            int rc = ensure_xbuf(s, words);
            if (!rc) rc = ensure_engine(s);
            if (rc) return rc;
            pc.xsync = s->xsync;
        } else {
            pc.iters = 0;
        }
    }
    const size_t bytes = sizeof(float) * (size_t)s->plane * vfs::PL_COUNT;
    float *scratch = nullptr;
    HIPCHK(s, hipMalloc(&scratch, bytes));
    vfs::launch_fill_zero(scratch, bytes, s->stream);
    std::vector<vfs::PairDesc> d{s->desc()};
    d[0].arena = scratch;
    const hipError_t r = time_launch_choices(s->stream, s->eng.d_desc, d, s->params(), masked, levels,
                                             multi_allowed(s, masked, levels), tn, pc.iters >= 2 ? &pc : nullptr);
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(scratch);
    if (r != hipSuccess) return fail(s, VFS_ERR_DEVICE, "launch tuner: %s", hipGetErrorString(r));
    return VFS_OK;
}

// Fills s->tuning[masked] (once per initialize / option change).
int tune_solver(vfs_solver *s, int masked)
{
    const int m = masked ? 1 : 0;
    if (s->tuned[m]) return VFS_OK;
    Tuning &tn = s->tuning[m];
    for (int l = 0; l < kMaxLevels; l++) tn.lv[l] = default_tune(l);
    if (s->force_shape >= 0) {
        for (int l = 0; l < s->nLevels; l++) {
            tn.lv[l].shape = (unsigned char)s->force_shape;
            tn.lv[l].mshape = tn.lv[l].m2shape = (unsigned char)(multi_shape(s->force_shape) ? s->force_shape : 4);
        }
    } else if (s->autotune) {
        std::vector<int> todo;
        const int piters = s->persist_opt != 0 ? std::min(s->nSolverIters, vfs::kMaxSched) & ~1 : 0;
        auto key = [&](int l) {
            const int ok = multi_allowed(s, m, std::vector<int>{l})[0] ? 1 : 0;
            return TuneKey{s->device, s->storage, m, s->lv[l].w, s->lv[l].h, s->lv[l].s, ok, piters};
        };
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            for (int l = 0; l < s->nLevels; l++) {
                auto it = g_tune.find(key(l));
                if (it != g_tune.end()) tn.lv[l] = it->second;
                else todo.push_back(l);
            }
        }
        if (!todo.empty()) {
            int rc = ensure_engine(s);
            if (!rc) rc = measure_levels(s, m, todo, tn);
            if (rc) return rc;
            std::lock_guard<std::mutex> g(g_tune_mu);
            for (int l : todo) g_tune[key(l)] = tn.lv[l];
        }
    }
    // option "per_launch": iterations per launch wherever allowed (1 everywhere else), no
    // persistent launches
    if (s->force_per_launch >= 0) {
        std::vector<int> all;
        for (int l = 0; l < s->nLevels; l++) all.push_back(l);
        const std::vector<bool> ok = multi_allowed(s, m, all);
        for (int l = 0; l < s->nLevels; l++) {
            tn.lv[l].per_launch = (unsigned char)(ok[l] ? s->force_per_launch : 1);
            tn.lv[l].persist = 0;
            if (!multi_shape(tn.lv[l].mshape)) tn.lv[l].mshape = 4;
            if (!multi_shape(tn.lv[l].m2shape)) tn.lv[l].m2shape = 4;
        }
    }
    // option "persist": 0 never, 1 wherever the op can run (in the shape with the most tiles)
    for (int l = 0; l < s->nLevels; l++) {
        if (s->persist_opt == 0) tn.lv[l].persist = 0;
        if (s->persist_opt == 1) {
            const int sh = persist_default_shape(s, m, l);
            tn.lv[l].persist = sh >= 0 ? 1 : 0;
            if (sh >= 0) tn.lv[l].pshape = (unsigned char)sh;
            tn.lv[l].pblock = 1;
        }
        // (a fixed or cached choice the device cannot hold at once is not taken)
        if (tn.lv[l].persist && persist_tiles(s, m, l, tn.lv[l].pshape, tn.lv[l].pblock) == 0) tn.lv[l].persist = 0;
    }
    s->tuned[m] = true;
    return VFS_OK;
}

// Whether a cold solve with these launch choices runs persistent launches, and the
// exchange words they need.
bool persist_program(const vfs_solver *s, int masked, const Tuning &tn, size_t *words)
{
    bool any = false;
    *words = 0;
    if (s->persist_opt == 0 || (std::min(s->nSolverIters, vfs::kMaxSched) & ~1) < 2) return false;
    for (int l = 0; l < s->nLevels; l++) {
        if (!tn.lv[l].persist) continue;
        const int K = tn.lv[l].pblock;
        const int n = persist_tiles(s, masked, l, tn.lv[l].pshape, K);
        if (n == 0 || persist_launch_iters(std::min(s->nSolverIters, vfs::kMaxSched)) < 2) continue;
        any = true;
        *words = std::max(*words, persist_words_of(n, tn.lv[l].pshape, K));
    }
    return any;
}

// Builds and runs the solve program on the solver's stream and records its end.  A cold
// solve runs the persistent launches of its tuning when `allow_persist`.
int run_solve(vfs_solver *s, int masked, int warm_level, int warps, int iters, bool allow_persist,
              const Params *params = nullptr)
{
    const Params P = params ? *params : s->params();
    const Tuning &tn = s->tuning[masked ? 1 : 0];
    size_t words = 0;
    const bool persist = allow_persist && warm_level < 0 && persist_program(s, masked, tn, &words);
    int rc;
    if (persist) {
        if ((rc = ensure_xbuf(s, words))) return rc;
        if ((rc = ensure_engine(s))) return rc;
    }
    const int start_roles = warm_level < 0 ? 0 : s->roles;
    const Program pr = build_program(P, masked, warm_level, warps, iters, start_roles, tn, persist);
    GraphKey key{masked, warm_level, warps, iters, start_roles, P.monitor, P.limitRange, tn.sig()};
    key.persist = persist ? 1 : 0;
    rc = engine_run(s->eng, P, pr, key, s->use_graph, &s->err);
    if (rc) return rc;
    s->roles = pr.final_roles;
    s->raw_i1_fresh = false;
    s->solved = true;
    s->mon_slots = pr.mon_slots;
    s->numeric = 0;
    s->persist_check = persist;
    s->persist_masked = masked;
    s->persist_params = P;
    HIPCHK(s, hipEventRecord(s->done, s->stream));
    return VFS_OK;
}

// After a solve with persistent launches: if one gave up (not every workgroup was
// resident, e.g. beside other kernels on the device), the solve runs again without them,
// from the same inputs (the calibrated I1 has not replaced the raw one yet).  Waits for the
// solve when there is something to check.
int resolve_persist(vfs_solver *s)
{
    if (!s->persist_check) return VFS_OK;
    s->persist_check = false;
    bool aborted = false;
    int rc = persist_aborted(s, &aborted);
    if (rc || !aborted) return rc;
    s->persist_reruns++;
    return run_solve(s, s->persist_masked, -1, 0, 0, false, &s->persist_params);
}

// warm_level < 0: cold solve; otherwise a warm start at warm_level with warps/iters.
int solve_common(vfs_solver *s, int masked, int warm_level = -1, int warps = 0, int iters = 0)
{
    if (!s) return VFS_ERR_ARG;
    int rc = ready_to_solve(s, masked, &s->err);
    if (rc) return rc;
    if (warm_level >= 0 && !s->solved)
        return fail(s, VFS_ERR_STATE, "warm-started solve before any solve of this pair");
    if (warm_level >= s->nLevels) return fail(s, VFS_ERR_ARG, "warm start level %d >= nLevels", warm_level);
    if ((rc = set_device(s))) return rc;
    // A warm start reads the previous solve's result: its persistent launches are settled
    // first.  A cold solve replaces it unread; if one of them gave up, the flag they left
    // makes this solve's persistent launches give up too, and this solve runs again at the
    // next wait.
    if (warm_level >= 0 && (rc = resolve_persist(s))) return rc;
    if ((rc = ensure_engine(s))) return rc;
    HIPCHK(s, refresh_raw_i1(s, s->stream));
    if ((rc = tune_solver(s, masked))) return rc;
    return run_solve(s, masked, warm_level, warps, iters, true);
}

// Waits (on the solver's stream) for the last solve of this pair, wherever it ran.
int wait_done(vfs_solver *s)
{
    int rc = resolve_persist(s);
    if (rc) return rc;
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

// After the last solve of the pair has finished: VFS_ERR_NUMERIC if one of its energy
// records counted non-finite values (a zero trajectory vector inside the mask makes the
// derivative NaN there, computeDerivatives.cu:99-104), else VFS_OK.  Reads the records
// once per solve, a few at a time into a buffer on the stack (no allocation).
int check_numeric(vfs_solver *s)
{
    if (s->mon_slots.empty()) return VFS_OK;
    if (s->numeric == 0) {
        constexpr size_t kBuf = 16;
        vfs::EnergyPart rec[kBuf];
        s->numeric = 1;
        for (size_t i0 = 0; i0 < s->mon_slots.size() && s->numeric > 0; i0 += kBuf) {
            const size_t n = std::min(kBuf, s->mon_slots.size() - i0);
            for (size_t k = 0; k < n; k++)
                HIPCHK(s, hipMemcpyAsync(&rec[k], s->mon + s->mon_slots[i0 + k], sizeof(vfs::EnergyPart),
                                         hipMemcpyDeviceToHost, s->stream));
            HIPCHK(s, hipStreamSynchronize(s->stream));
            for (size_t k = 0; k < n; k++) {
                if (rec[k].nonfinite == 0) continue;
                const int slot = s->mon_slots[i0 + k];
                s->numeric = -1;
                fail(s, VFS_ERR_NUMERIC,
                     "the solve produced %u non-finite values at level %d (warp %d); a zero trajectory vector "
                     "inside the processed region gives NaN derivatives",
                     rec[k].nonfinite, slot / vfs::kMonWarps, slot % vfs::kMonWarps);
                break;
            }
        }
    }
    return s->numeric < 0 ? VFS_ERR_NUMERIC : VFS_OK;
}

// copy_out of a result plane, then the numeric check of the solve that produced it.
int copy_result(vfs_solver *s, const void *dev, int nch, int level, float *out, int pitch_elems)
{
    const int rc = copy_out(s, dev, nch, level, out, pitch_elems);
    return rc ? rc : check_numeric(s);
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
    if (width < 1 || height < 1 || width > 32767 || height > 32767)
        return fail(s, VFS_ERR_ARG, "bad size %dx%d (1 .. 32767 per side)", width, height);
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
    s->tuned[0] = s->tuned[1] = false;
    // pyramid dims (stereotgv.cu:74-95)
    int nh = height, nw = width;
    for (int l = 0; l < nLevels; l++) {
        if (nw < 1 || nh < 1) return fail(s, VFS_ERR_ARG, "level %d collapses to %dx%d", l, nw, nh);
        s->lv[l] = Dims{nw, nh, align_up(nw, s->strideAlign)};
        nh = (int)((float)nh / fScale);
        nw = (int)((float)nw / fScale);
    }
    const size_t N = (size_t)s->stride * height;
    // launch entries carry strides as 16-bit and plane sizes as 31-bit integers
    if (s->stride > 65535 || N + 63 > (size_t)0x7fffffff)
        return fail(s, VFS_ERR_ARG, "row stride %d x height %d too large", s->stride, height);
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
    // energy records and the partial sums of the largest level
    const size_t nmon = (size_t)kMaxLevels * vfs::kMonWarps + (size_t)vfs::energy_chunks(s->stride, height);
    if ((rc = dev_alloc(s, (void **)&s->mon, nmon * sizeof(vfs::EnergyPart)))) return rc;
    s->mon_slots.clear();
    s->numeric = 0;
    // work arena: PL_COUNT slots of >= N floats, each 256 B aligned
    s->plane = (long long)((N + 63) & ~(size_t)63);
    if ((rc = dev_alloc(s, (void **)&s->arena, sizeof(float) * (size_t)vfs::PL_COUNT * s->plane))) return rc;
    s->roles = vfs::kRolesInit;
    s->solved = false;
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
    // column 0 of every mask level (multi-iteration launches need it empty, see op_iterk)
    for (int l = 0; l < s->nLevels; l++) {
        std::vector<float> col(s->lv[l].h);
        HIPCHK(s, hipMemcpy2DAsync(col.data(), sizeof(float), s->mask[l], sizeof(float) * (size_t)s->lv[l].s,
                                   sizeof(float), s->lv[l].h, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        s->mask_col0[l] = std::any_of(col.begin(), col.end(), [](float v) { return v != 0.f; });
    }
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->has_mask = true;
    s->tuned[1] = false;
    s->mask_gen++;
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

int vfs_solve_warm(vfs_solver *s, int masked, int start_level, int warps, int iters)
{
    if (!s) return VFS_ERR_ARG;
    if (start_level < 0 || warps < 0 || iters < 0) return fail(s, VFS_ERR_ARG, "bad warm schedule");
    return solve_common(s, masked, start_level, warps, iters);
}

static float *slot_ptr(vfs_solver *s, int k) { return s->arena + (long long)k * s->plane; }

int vfs_copy_disparity_1d(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    // a solve that runs again (persistent launches that gave up) may end with other roles
    if ((rc = resolve_persist(s))) return rc;
    return copy_result(s, slot_ptr(s, vfs::slot_u(vfs::rcur(s->roles, vfs::G_U))), 1, 0, out, pitch);
}

int vfs_copy_flow(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < 2 * s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = resolve_persist(s))) return rc;  // (roles, as above)
    return copy_result(s, slot_ptr(s, vfs::slot_uv(s->roles, true)), 2, 0, out, pitch);
}

int vfs_copy_warped_image(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    rc = copy_out_st(s, slot_ptr(s, vfs::PL_I1W), 1, 0, out, pitch);
    return rc ? rc : check_numeric(s);
}

}  // extern "C"

// ---- the warped image without allocation
//
// vfs_copy_warped_image in 16-bit storage widens the plane into a temporary float plane
// that it allocates per call.  vfs_copy_warped_image_noalloc stages it in `scratch`, a
// file-scope device array of kNoAllocFloats floats instead (every device the library runs
// on holds its own copy until the process ends).  In host code `scratch` names only a
// host-side stand-in: its device address, looked up once per device with
// hipGetSymbolAddress, is what kernels and copies take.  Nothing in it must start at zero
// (the widen kernel writes every element the copy reads).  Every call on a device shares
// that device's array, so calls that overlap on one device race.
namespace {

constexpr size_t kNoAllocFloats = (size_t)4 << 20;  // 16 MiB: planes up to 4 Mi floats
static __device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per device

// The device address of `scratch` on s->device (current); looked up on the first call there.
int scratch_address(vfs_solver *s, float **out)
{
    if (s->device < 0 || s->device >= kMaxDevices)
        return fail(s, VFS_ERR_DEVICE, "device %d beyond the %d devices of the staging table", s->device, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[s->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(s, VFS_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[s->device] = (float *)p;
    }
    *out = g_scratch_dev[s->device];
    return VFS_OK;
}

}  // namespace

extern "C" {

int vfs_copy_warped_image_noalloc(vfs_solver *s, float *out, int pitch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    const void *src = slot_ptr(s, vfs::PL_I1W);
    if (s->storage == vfs::kF32) return copy_result(s, src, 1, 0, out, pitch);  // a float plane: no staging
    const Dims d0 = s->lv[0];
    const size_t need = (size_t)d0.s * d0.h;
    if (need > kNoAllocFloats)
        return fail(s, VFS_ERR_ARG, "warped image of %zu floats (stride x height) exceeds the %zu-float staging "
                                    "of the no-allocation copy", need, kNoAllocFloats);
    float *stage = nullptr;
    if ((rc = scratch_address(s, &stage))) return rc;
    if ((rc = wait_done(s))) return rc;
    vfs::launch_widen(src, s->storage, 1, stage, d0, s->stream);
    if ((rc = check_launch(s))) return rc;
    return copy_result(s, stage, 1, 0, out, pitch);
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
    const float2 *uv = (const float2 *)slot_ptr(s, vfs::slot_uv(s->roles, true));
    vfs::launch_convert_kb(uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3, s->outX, s->outDepth, d0,
                           s->stream);
    if ((rc = check_launch(s))) return rc;
    if (depth && (rc = copy_out(s, s->outDepth, 1, 0, depth, depth_pitch))) return rc;
    if (X && (rc = copy_out(s, s->outX, 3, 0, X, X_pitch))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return check_numeric(s);
}

int vfs_copy_disparity_vis(vfs_solver *s, float *out, int pitch, float flowScale)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "copy before initialize");
    if (!out || pitch < 3 * s->width) return fail(s, VFS_ERR_ARG, "bad output pointer or pitch");
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = wait_done(s))) return rc;
    const float2 *uv = (const float2 *)slot_ptr(s, vfs::slot_uv(s->roles, true));
    vfs::launch_flow_hsv(uv, s->lv[0], s->outX, flowScale, s->stream);  // outX: 3-float staging plane
    if ((rc = check_launch(s))) return rc;
    return copy_result(s, s->outX, 3, 0, out, pitch);
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
    int rc = set_device(s);
    if (rc) return rc;
    // a solve that runs again (persistent launches that gave up) may end with other roles
    if ((rc = resolve_persist(s))) return rc;
    const int r = s->roles;
    const std::string n(name);
    const void *src = nullptr;
    int nch = 1;
    bool typed = false;  // stored in the storage type
    using namespace vfs;
    if (n == "u") src = slot_ptr(s, slot_u(rcur(r, G_U)));
    else if (n == "u_") src = slot_ptr(s, slot_ub(rcur(r, G_UB)));
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
    else if (n == "warpUV") { src = slot_ptr(s, slot_uv(r, true)); nch = 2; }
    else if (n == "tv2") { src = slot_ptr(s, PL_TV2); nch = 2; }
    else if (n == "tv") { src = s->tv[level]; nch = 2; }
    else if (n == "p") { src = slot_ptr(s, slot_p(rcur(r, G_PQ))); nch = 2; typed = true; }
    else if (n == "v_") { src = slot_ptr(s, slot_v(rcur(r, G_V))); nch = 2; }
    else if (n == "Tp") { src = slot_ptr(s, PL_TP); nch = 2; }
    else if (n == "u.other") src = slot_ptr(s, slot_u(roth(r, G_U)));
    else if (n == "u_.other") src = slot_ptr(s, slot_ub(roth(r, G_UB)));
    else if (n == "v_.other") { src = slot_ptr(s, slot_v(roth(r, G_V))); nch = 2; }
    else if (n == "q") { src = slot_ptr(s, slot_q(rcur(r, G_PQ))); nch = 4; typed = true; }
    if (!src) return fail(s, VFS_ERR_ARG, "unknown plane '%s'", name);
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
    return check_numeric(s);
}

int vfs_get_energy(vfs_solver *s, int level, int warp, vfs_energy_t *out)
{
    if (!s || !out) return VFS_ERR_ARG;
    if (!s->initialized) return fail(s, VFS_ERR_STATE, "energy before initialize");
    if (level < 0 || level >= s->nLevels || warp < 0) return fail(s, VFS_ERR_ARG, "bad level or warp");
    const int k = vfs::mon_slot(level, warp);
    if (warp >= vfs::kMonWarps || std::find(s->mon_slots.begin(), s->mon_slots.end(), k) == s->mon_slots.end())
        return fail(s, VFS_ERR_STATE, "the last solve wrote no energy record for level %d, warp %d", level, warp);
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = wait_done(s))) return rc;
    vfs::EnergyPart r;
    HIPCHK(s, hipMemcpyAsync(&r, s->mon + k, sizeof r, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    out->energy = r.energy;
    out->data = r.data;
    out->residual = r.residual;
    out->nonfinite = (int)r.nonfinite;
    out->pixels = (int)r.pixels;
    return VFS_OK;
}

int vfs_set_option(vfs_solver *s, const char *key, int value)
{
    if (!s || !key) return VFS_ERR_ARG;
    if (s->persist_check) {
        // a solve whose persistent launches may have given up runs again as it was set up
        int rc = set_device(s);
        if (!rc) rc = resolve_persist(s);
        if (rc) return rc;
    }
    if (!strcmp(key, "graph")) {
        s->use_graph = value != 0;
        return VFS_OK;
    }
    if (!strcmp(key, "storage")) {
        if (value < vfs::kF32 || value > vfs::kBF16) return fail(s, VFS_ERR_ARG, "storage must be 0, 1 or 2");
        if (value != s->storage) {
            s->storage = value;
            s->desc_dirty = true;
            s->tuned[0] = s->tuned[1] = false;
        }
        return VFS_OK;
    }
    if (!strcmp(key, "shape")) {
        if (value < -1 || value >= vfs::kTickShapes)
            return fail(s, VFS_ERR_ARG, "shape must be -1 .. %d", vfs::kTickShapes - 1);
        s->force_shape = value;
        s->tuned[0] = s->tuned[1] = false;
        return VFS_OK;
    }
    if (!strcmp(key, "autotune")) {
        s->autotune = value != 0;
        s->tuned[0] = s->tuned[1] = false;
        return VFS_OK;
    }
    if (!strcmp(key, "energy")) {
        if (value < 0 || value > 2) return fail(s, VFS_ERR_ARG, "energy must be 0, 1 or 2");
        s->monitor = value;
        return VFS_OK;
    }
    if (!strcmp(key, "per_launch") || !strcmp(key, "pair")) {
        // "pair": -1 = tuner, 0 = one iteration per launch, 1 = two
        if (!strcmp(key, "pair")) {
            if (value < -1 || value > 1) return fail(s, VFS_ERR_ARG, "pair must be -1, 0 or 1");
            value = value < 0 ? -1 : value + 1;
        }
        if (value != -1 && !valid_per_launch(value))
            return fail(s, VFS_ERR_ARG, "per_launch must be -1, 1, 2, 4 or 8");
        s->force_per_launch = value;
        s->tuned[0] = s->tuned[1] = false;
        return VFS_OK;
    }
    if (!strcmp(key, "persist")) {
        if (value < -1 || value > 1) return fail(s, VFS_ERR_ARG, "persist must be -1, 0 or 1");
        s->persist_opt = value;
        s->tuned[0] = s->tuned[1] = false;
        return VFS_OK;
    }
    if (!strcmp(key, "persist_give_up")) {
        // the next solve's persistent launches find the abort flag set and leave at once;
        // the solve then runs again without them (the re-run path, for tests)
        if (value != 1) return fail(s, VFS_ERR_ARG, "persist_give_up takes 1");
        if (!s->initialized) return fail(s, VFS_ERR_STATE, "persist_give_up before initialize");
        int rc = set_device(s);
        if (!rc) rc = resolve_persist(s);  // settle the last solve first
        if (!rc) rc = ensure_xbuf(s, 0);
        if (rc) return rc;
        const unsigned one = 1;
        HIPCHK(s, hipMemcpyAsync(&s->xsync->abort, &one, sizeof one, hipMemcpyHostToDevice, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return VFS_OK;
    }
    return fail(s, VFS_ERR_ARG, "unknown option '%s'", key);
}

int vfs_get_option(const vfs_solver *s, const char *key, int *value)
{
    if (!s || !key || !value) return VFS_ERR_ARG;
    if (!strcmp(key, "graph")) *value = s->use_graph ? 1 : 0;
    else if (!strcmp(key, "storage")) *value = s->storage;
    else if (!strcmp(key, "shape")) *value = s->force_shape;
    else if (!strcmp(key, "autotune")) *value = s->autotune ? 1 : 0;
    else if (!strcmp(key, "energy")) *value = s->monitor;
    else if (!strcmp(key, "per_launch")) *value = s->force_per_launch;
    else if (!strcmp(key, "persist")) *value = s->persist_opt;
    else if (!strcmp(key, "persist_reruns")) *value = s->persist_reruns;
    else if (!strcmp(key, "persist_why")) *value = (int)s->persist_why;
    else if (!strcmp(key, "pair")) *value = s->force_per_launch < 0 ? -1 : s->force_per_launch > 1 ? 1 : 0;
    else return VFS_ERR_ARG;
    return VFS_OK;
}

int vfs_get_level_tuning(vfs_solver *s, int masked, int level, int *shape, int *preload, int *per_launch)
{
    if (!s || !shape || !preload) return VFS_ERR_ARG;
    if (!s->initialized || level < 0 || level >= s->nLevels) return fail(s, VFS_ERR_ARG, "bad level");
    int rc = set_device(s);
    if (!rc) rc = tune_solver(s, masked);
    if (rc) return rc;
    const LevelTune t = s->tuning[masked ? 1 : 0].lv[level];
    *shape = t.iter_shape();
    *preload = t.preload;
    if (per_launch) *per_launch = t.persist ? VFS_PERSISTENT | (t.pblock > 1 ? t.pblock : 0) : t.per_launch;
    return VFS_OK;
}

int vfs_set_level_tuning(vfs_solver *s, int masked, int level, int shape, int preload, int per_launch)
{
    if (!s) return VFS_ERR_ARG;
    if (!s->initialized || level < 0 || level >= s->nLevels) return fail(s, VFS_ERR_ARG, "bad level");
    if (per_launch & VFS_PERSISTENT) {
        const int K = std::max(per_launch & ~VFS_PERSISTENT, 1);
        if (K != 1 && K != 2 && K != 4 && K != 8)
            return fail(s, VFS_ERR_ARG, "VFS_PERSISTENT takes 2, 4 or 8 iterations per block, or none");
        const int m = masked ? 1 : 0;
        int rc = set_device(s);
        if (!rc && !s->tuned[m]) rc = tune_solver(s, m);
        if (rc) return rc;
        if (s->persist_opt == 0) return fail(s, VFS_ERR_ARG, "persistent launches are switched off (option persist)");
        if (persist_tiles(s, m, level, shape, K) == 0)
            return fail(s, VFS_ERR_ARG, "no persistent launch in shape %d with blocks of %d at level %d (shape, mask "
                                        "column 0, or more tiles than the device holds at once)", shape, K, level);
        LevelTune &t = s->tuning[m].lv[level];
        t.persist = 1;
        t.pshape = (unsigned char)shape;
        t.pblock = (unsigned char)K;
        return VFS_OK;
    }
    const bool half = (per_launch & VFS_HALF_WORKGROUP) != 0;
    per_launch &= ~VFS_HALF_WORKGROUP;
    if (!valid_per_launch(per_launch) || (half && (per_launch == 1 || !multi_shape(shape) ||
                                                   !vfs::half_workgroup_ok(shape, per_launch))))
        return fail(s, VFS_ERR_ARG, "per_launch must be 1, 2, 4 or 8 (half workgroups where the staged region "
                                    "holds 513 .. 1536 positions)");
    if (per_launch > 1 ? !(multi_shape(shape) && vfs::iterk_ok(shape, per_launch))
                       : (shape < 0 || shape >= vfs::kTickShapes))
        return fail(s, VFS_ERR_ARG, "bad shape %d for %d iteration(s) per launch", shape, per_launch);
    const int m = masked ? 1 : 0;
    int rc = set_device(s);
    if (!rc && !s->tuned[m]) rc = tune_solver(s, m);  // the other levels keep their choices
    if (rc) return rc;
    if (per_launch > 1 && !multi_allowed(s, m, std::vector<int>{level})[0])
        return fail(s, VFS_ERR_ARG, "several iterations per launch are not allowed at level %d", level);
    LevelTune &t = s->tuning[m].lv[level];
    if (per_launch > 1) {
        t.mshape = t.m2shape = (unsigned char)shape;
        t.mhalf = t.m2half = (unsigned char)half;
    } else {
        t.shape = (unsigned char)shape;
        t.preload = (unsigned char)(preload != 0);
    }
    t.per_launch = (unsigned char)per_launch;
    t.persist = 0;
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
    // persistent launches that gave up did not do their work: the timing does not count
    // (the last solve runs again, so the solver's result stays right)
    const int reruns = s->persist_reruns;
    if ((rc = resolve_persist(s))) return rc;
    if (s->persist_reruns != reruns)
        return fail(s, VFS_ERR_DEVICE, "persistent launches gave up during the timed solves (not every workgroup "
                                       "resident); time with option \"persist\" 0");
    *ms_total = ms;
    return VFS_OK;
}

int vfs_time_inner_iterations(vfs_solver *s, int level, int reps, float *ms_total, int *iters_per_launch)
{
    const int reps0 = reps;
    if (!s || !ms_total || reps < 1) return VFS_ERR_ARG;
    if (!s->initialized || level < 0 || level >= s->nLevels) return fail(s, VFS_ERR_ARG, "bad level");
    int rc = set_device(s);
    if (rc) return rc;
    if ((rc = ensure_engine(s))) return rc;
    if ((rc = tune_solver(s, 1))) return rc;
    if ((rc = wait_done(s))) return rc;
    const Params P = s->params();
    Program pr;
    int roles = vfs::kRolesInit;
    int per_launch = s->tuning[1].lv[level].per_launch;
    // a persistent level: launches of a warp's iterations, as the solve runs them
    size_t words = 0;
    const LevelTune &lt = s->tuning[1].lv[level];
    const int pb = lt.pblock;
    const bool persist = persist_program(s, 1, s->tuning[1], &words) && lt.persist &&
                         persist_tiles(s, 1, level, lt.pshape, pb) > 0 &&
                         persist_launch_iters(std::min(s->nSolverIters, vfs::kMaxSched)) >= 2;
    const int sched_n = persist ? persist_launch_iters(std::min(s->nSolverIters, vfs::kMaxSched)) : 48;
    if (persist) {
        per_launch = sched_n;
        if ((rc = ensure_xbuf(s, words))) return rc;
        if ((rc = ensure_engine(s))) return rc;
    }
    // at least `reps` iterations of the sched_n-iteration schedule, restarted every
    // sched_n, in whole ops
    {
        Params Q = P;
        Q.nSolverIters = sched_n;
        Program w;
        int done = 0;
        while (done < reps) {
            w.ops.clear();
            push_level(w, Q, 1, level, 1, sched_n, roles, per_launch, false, persist ? pb : 0);
            for (const ProgOp &o : w.ops) {
                if (done >= reps) break;
                const int k = o.kind == vfs::OP_ITERP || o.kind == vfs::OP_ITERPK ? vfs::iterp_count(o.it0)
                                                                                   : vfs::op_iterations(o.kind);
                if (k > 0) {
                    pr.ops.push_back(o);
                    done += k;
                }
            }
        }
        reps = done;
        apply_tuning(pr, s->tuning[1]);
    }
    // the launches run on a copy of the arena (same data, so the same branches), which
    // leaves the solver's state as the last solve left it
    const size_t bytes = sizeof(float) * (size_t)s->plane * vfs::PL_COUNT;
    float *scratch = nullptr;
    HIPCHK(s, hipMalloc(&scratch, bytes));
    Engine tmp;
    tmp.device = s->device;
    tmp.stream = s->stream;
    tmp.d_desc = s->eng.d_desc;
    tmp.h_desc = s->eng.h_desc;
    tmp.h_desc[0].arena = scratch;
    hipError_t ri = hipMemcpyAsync(scratch, s->arena, bytes, hipMemcpyDeviceToDevice, s->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ri == hipSuccess) ri = hipEventCreate(&e0);
    if (ri == hipSuccess) ri = hipEventCreate(&e1);
    if (ri == hipSuccess) ri = hipEventRecord(e0, s->stream);
    if (ri == hipSuccess) ri = issue_program(tmp, P, pr);
    if (ri == hipSuccess) ri = hipEventRecord(e1, s->stream);
    if (ri == hipSuccess) ri = hipEventSynchronize(e1);
    float ms = 0;
    if (ri == hipSuccess) ri = hipEventElapsedTime(&ms, e0, e1);
    if (ri == hipSuccess) ri = hipGetLastError();
    (void)hipStreamSynchronize(s->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    if (ri != hipSuccess) return fail(s, VFS_ERR_DEVICE, "timing launches: %s", hipGetErrorString(ri));
    if (persist) {
        bool aborted = false;
        if ((rc = persist_aborted(s, &aborted))) return rc;
        if (aborted)
            return fail(s, VFS_ERR_DEVICE, "persistent launches gave up while timed (not every workgroup resident)");
    }
    *ms_total = ms * (float)reps0 / (float)reps;
    if (iters_per_launch) *iters_per_launch = per_launch;
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

// Launch choices of a batch, timed on scratch copies of its (up to) 8 largest members: a
// batched launch covers many pairs, so its fastest tile shape can differ from a single
// pair's.  The first member's "shape" / "autotune" options apply.
static int tune_batch(vfs_batch *b, int masked)
{
    const int m = masked ? 1 : 0;
    if (b->tuned[m]) return VFS_OK;
    const vfs_solver *f = b->members[0];
    Tuning &tn = b->tuning[m];
    for (int l = 0; l < kMaxLevels; l++) tn.lv[l] = default_tune(l);
    if (f->force_shape >= 0) {
        for (int l = 0; l < f->nLevels; l++) {
            tn.lv[l].shape = (unsigned char)f->force_shape;
            tn.lv[l].mshape = tn.lv[l].m2shape = (unsigned char)(multi_shape(f->force_shape) ? f->force_shape : 4);
        }
    } else if (f->autotune) {
        std::vector<const vfs_solver *> order(b->members.begin(), b->members.end());
        std::stable_sort(order.begin(), order.end(), [](const vfs_solver *x, const vfs_solver *y) {
            return (long long)x->lv[0].w * x->lv[0].h > (long long)y->lv[0].w * y->lv[0].h;
        });
        const int n = std::min((int)order.size(), 8);
        std::vector<vfs::PairDesc> descs;
        std::vector<float *> scratch;
        hipError_t r = hipSuccess;
        for (int i = 0; i < n && r == hipSuccess; i++) {
            const size_t bytes = sizeof(float) * (size_t)order[i]->plane * vfs::PL_COUNT;
            float *p = nullptr;
            r = hipMalloc(&p, bytes);
            if (r != hipSuccess) break;
            scratch.push_back(p);
            vfs::launch_fill_zero(p, bytes, b->eng.stream);
            descs.push_back(order[i]->desc());
            descs.back().arena = p;
        }
        std::vector<int> levels;
        for (int l = 0; l < f->nLevels; l++) levels.push_back(l);
        // multi-iteration launches only where every member allows them
        std::vector<bool> ok(levels.size(), true);
        for (const vfs_solver *s : b->members) {
            const std::vector<bool> o = multi_allowed(s, m, levels);
            for (size_t i = 0; i < ok.size(); i++) ok[i] = ok[i] && o[i];
        }
        if (r == hipSuccess)
            r = time_launch_choices(b->eng.stream, b->eng.d_desc, descs, f->params(), m, levels, ok, tn);
        (void)hipStreamSynchronize(b->eng.stream);
        for (float *p : scratch) (void)hipFree(p);
        if (r != hipSuccess) return failb(b, VFS_ERR_DEVICE, "batch launch tuner: %s", hipGetErrorString(r));
    }
    std::vector<int> all;
    for (int l = 0; l < f->nLevels; l++) all.push_back(l);
    std::vector<bool> ok(all.size(), true);
    for (const vfs_solver *s : b->members) {
        const std::vector<bool> o = multi_allowed(s, m, all);
        for (size_t i = 0; i < ok.size(); i++) ok[i] = ok[i] && o[i];
    }
    if (f->force_per_launch >= 0) {
        for (int l = 0; l < f->nLevels; l++) {
            tn.lv[l].per_launch = (unsigned char)(ok[l] ? f->force_per_launch : 1);
            if (!multi_shape(tn.lv[l].mshape)) tn.lv[l].mshape = 4;
            if (!multi_shape(tn.lv[l].m2shape)) tn.lv[l].m2shape = 4;
        }
    }
    for (int l = 0; l < f->nLevels; l++)
        if (b->fixed[m][l]) {
            tn.lv[l] = b->fixed_lv[m][l];
            if (!ok[l]) tn.lv[l].per_launch = 1;  // a member's mask forbids several iterations per launch here
        }
    b->tuned[m] = true;
    return VFS_OK;
}

static int batch_solve(vfs_batch *b, int masked, int warm_level, int warps, int iters)
{
    if (!b) return VFS_ERR_ARG;
    HIPCHKB(b, hipSetDevice(b->eng.device));
    const vfs_solver *f = b->members[0];
    Params P = f->params();
    bool rebuild = b->eng.h_desc.size() != b->members.size();
    for (size_t i = 0; i < b->members.size(); i++) {
        vfs_solver *s = b->members[i];
        std::string why;
        if (ready_to_solve(s, masked, &why)) return failb(b, VFS_ERR_STATE, "member %zu: %s", i, why.c_str());
        // a member's own last solve is settled before the batch overwrites its state (if
        // its persistent launches gave up it runs again first, from the same inputs)
        if (resolve_persist(s)) return failb(b, VFS_ERR_DEVICE, "member %zu: %s", i, s->err.c_str());
        if (s->storage != f->storage || s->limitRange != f->limitRange || s->monitor != f->monitor)
            return failb(b, VFS_ERR_ARG, "member %zu differs in storage, limitRange or the energy option", i);
        if (warm_level >= 0 && (!s->solved || s->roles != f->roles))
            return failb(b, VFS_ERR_STATE, "member %zu: warm start needs every member solved with the same program",
                         i);
        // the batch keeps its own copy of every member's descriptor: re-upload (and
        // re-tune) only when one of them changed
        if (!rebuild) {
            const vfs::PairDesc d = s->desc();
            if (memcmp(&b->eng.h_desc[i], &d, sizeof d) != 0) rebuild = true;
            if (b->mask_gen.size() != b->members.size() || b->mask_gen[i] != s->mask_gen) rebuild = true;
        }
    }
    if (warm_level >= P.nLevels) return failb(b, VFS_ERR_ARG, "warm start level %d >= nLevels", warm_level);
    if (rebuild) {
        std::vector<vfs::PairDesc> descs;
        for (vfs_solver *s : b->members) descs.push_back(s->desc());
        hipError_t r = engine_setup(b->eng, descs);
        if (r != hipSuccess) return failb(b, VFS_ERR_DEVICE, "engine setup: %s", hipGetErrorString(r));
        b->tuned[0] = b->tuned[1] = false;
        b->mask_gen.clear();
        for (vfs_solver *s : b->members) b->mask_gen.push_back(s->mask_gen);
    }
    // order after each member's pending work, then refresh re-used raw inputs
    for (vfs_solver *s : b->members) {
        HIPCHKB(b, hipStreamWaitEvent(b->eng.stream, s->done, 0));
        HIPCHKB(b, refresh_raw_i1(s, b->eng.stream));
    }
    int rc0 = tune_batch(b, masked);
    if (rc0) return rc0;
    const Tuning &tn = b->tuning[masked ? 1 : 0];
    const int start_roles = warm_level < 0 ? 0 : f->roles;
    const Program pr = build_program(P, masked, warm_level, warps, iters, start_roles, tn);
    const GraphKey key{masked, warm_level, warps, iters, start_roles, P.monitor, P.limitRange, tn.sig()};
    int rc = engine_run(b->eng, P, pr, key, f->use_graph, &b->err);
    if (rc) return rc;
    for (vfs_solver *s : b->members) {
        s->roles = pr.final_roles;
        s->raw_i1_fresh = false;
        s->solved = true;
        s->mon_slots = pr.mon_slots;
        s->numeric = 0;
        HIPCHKB(b, hipEventRecord(s->done, b->eng.stream));
    }
    return VFS_OK;
}

int vfs_batch_solve(vfs_batch *b, int masked) { return batch_solve(b, masked, -1, 0, 0); }

int vfs_batch_solve_warm(vfs_batch *b, int masked, int start_level, int warps, int iters)
{
    if (start_level < 0 || warps < 0 || iters < 0) return failb(b, VFS_ERR_ARG, "bad warm schedule");
    return batch_solve(b, masked, start_level, warps, iters);
}

int vfs_batch_get_level_tuning(vfs_batch *b, int masked, int level, int *shape, int *preload, int *per_launch)
{
    if (!b || !shape || !preload) return VFS_ERR_ARG;
    if (level < 0 || level >= b->members[0]->nLevels) return failb(b, VFS_ERR_ARG, "bad level");
    if (!b->tuned[masked ? 1 : 0]) return failb(b, VFS_ERR_STATE, "the batch has not been solved in this form yet");
    const LevelTune t = b->tuning[masked ? 1 : 0].lv[level];
    *shape = t.iter_shape();
    *preload = t.preload;
    if (per_launch) *per_launch = t.per_launch;
    return VFS_OK;
}

int vfs_batch_set_level_tuning(vfs_batch *b, int masked, int level, int shape, int preload, int per_launch)
{
    if (!b) return VFS_ERR_ARG;
    if (level < 0 || level >= b->members[0]->nLevels) return failb(b, VFS_ERR_ARG, "bad level");
    const bool half = (per_launch & VFS_HALF_WORKGROUP) != 0;
    per_launch &= ~VFS_HALF_WORKGROUP;
    if (!valid_per_launch(per_launch) || (half && (per_launch == 1 || !multi_shape(shape) ||
                                                   !vfs::half_workgroup_ok(shape, per_launch))))
        return failb(b, VFS_ERR_ARG, "per_launch must be 1, 2, 4 or 8 (half workgroups where the staged region "
                                     "holds 513 .. 1536 positions)");
    if (per_launch > 1 ? !(multi_shape(shape) && vfs::iterk_ok(shape, per_launch))
                       : (shape < 0 || shape >= vfs::kTickShapes))
        return failb(b, VFS_ERR_ARG, "bad shape %d for %d iteration(s) per launch", shape, per_launch);
    const int m = masked ? 1 : 0;
    LevelTune t = b->fixed[m][level] ? b->fixed_lv[m][level] : default_tune(level);
    if (per_launch > 1) {
        t.mshape = t.m2shape = (unsigned char)shape;
        t.mhalf = t.m2half = (unsigned char)half;
    } else {
        t.shape = (unsigned char)shape;
        t.preload = (unsigned char)(preload != 0);
    }
    t.per_launch = (unsigned char)per_launch;
    b->fixed_lv[m][level] = t;
    b->fixed[m][level] = true;
    b->tuned[m] = false;  // applied at the next solve
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

// ------------------------------------------------------------------ serving
//
// A continuous-batching server for many camera streams on one device.  Each stream owns a
// resident solver (pyramids, duals, previous disparity), materialized when one of its
// frames is admitted and the HBM budget allows, and evicted (least recently used, idle
// only) when another stream needs the memory.  A frame is admitted as soon as its
// stream is idle, an entry slot is free and its stream is resident; it then runs its own
// program (cold: the full schedule; warm: from the stream's previous result).  Every
// scheduling step launches ONE k_tick carrying the next op of every active frame, so
// frames that arrived at different times share launches while each result depends only on
// its own ops.  Everything is ordered on the server's stream; completion is observed
// through events.

namespace {

struct SrvStream {
    int width = 0, height = 0, nLevels = 0, nWarpIters = 0, nSolverIters = 0;
    int warm_level = 1, warm_warps = 2, warm_iters = 20;
    std::vector<float> mask, tv, cv;  // host copies for (re)materialization
    vfs_solver *solver = nullptr;     // resident state, or null
    int slot = -1;                    // descriptor slot while resident
    size_t bytes = 0;                 // device footprint while resident
    bool busy = false;                // a frame of this stream is admitted and unfinished
    bool warm_ok = false;             // the resident state holds a previous result
    long long last_use = 0;
    std::deque<struct SrvRequest *> queue;  // submitted, not yet admitted (in order)
};

struct SrvRequest {
    long long seq = 0;  // submission order
    long long tag;
    int stream;
    float *out = nullptr;  // caller buffer (width*height floats), may be null
    const float *input = nullptr;  // pinned images (2*w*h floats: i0, then i1)
    float *own_input = nullptr;    // the server's pooled copy behind `input` (vfs_server_submit), or null
    float *result = nullptr;       // pinned: the result (w*h) and the last energy record (kPinnedExtra)
    double t_submit = 0, t_done = 0;
    int warm = 0;
    Program prog;
    size_t pc = 0;
    int waited = 0;                 // steps since this frame's last op was launched
    bool wait_input = false;        // the next launch carrying this frame waits for in_ready
    hipEvent_t in_ready = nullptr;  // copy stream: the frame's images are on the device
    hipEvent_t ops_done = nullptr;  // compute stream: the frame's last op
    hipEvent_t done = nullptr;      // copy stream: the result is in `result`
    int nonfinite = 0;    // of the frame's last energy record (level 0)
    double energy = 0.0;
};

// Floats of a frame's pinned result buffer after its w h result floats: the last energy record.
constexpr size_t kPinnedExtra = sizeof(vfs::EnergyPart) / sizeof(float);

double now_ms()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

}  // namespace

struct vfs_server {
    int device = 0;
    hipStream_t stream = nullptr;       // the op launches
    hipStream_t copy_stream = nullptr;  // frames in and results out, overlapping the launches
    Params P{};
    size_t budget = 0, used = 0;
    int max_active = vfs::kMaxEntries;
    int max_slots = 0;
    vfs::PairDesc *d_desc = nullptr;
    std::vector<vfs::PairDesc> h_desc;  // host copy of every slot's descriptor
    std::vector<int> free_slots;
    std::vector<SrvStream> streams;
    std::vector<SrvRequest *> active, finishing;
    long long queued = 0, seq = 0;
    std::vector<SrvRequest *> done;
    std::vector<hipEvent_t> tick_events;  // in flight, oldest first
    // recycled host resources: pinned frame buffers (by size in floats) and events, so
    // that a frame costs no page locking and no event creation once the server is warm
    std::vector<std::pair<size_t, float *>> pinned_pool;
    std::vector<hipEvent_t> event_pool;
    long long clock = 0;
    long long ticks = 0, launched_entries = 0;
    std::string err;
};

namespace {

float *take_pinned(vfs_server *v, size_t floats)
{
    for (size_t i = 0; i < v->pinned_pool.size(); i++)
        if (v->pinned_pool[i].first == floats) {
            float *p = v->pinned_pool[i].second;
            v->pinned_pool.erase(v->pinned_pool.begin() + i);
            return p;
        }
    float *p = nullptr;
    return hipHostMalloc((void **)&p, sizeof(float) * floats, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

// Keeps at most 512 buffers (inputs and results) for reuse.
void give_pinned(vfs_server *v, float *p, size_t floats)
{
    if (!p) return;
    if (v->pinned_pool.size() < 512) v->pinned_pool.emplace_back(floats, p);
    else (void)hipHostFree(p);
}

hipError_t take_event(vfs_server *v, hipEvent_t *e)
{
    if (!v->event_pool.empty()) {
        *e = v->event_pool.back();
        v->event_pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreateWithFlags(e, hipEventDisableTiming);
}

void give_event(vfs_server *v, hipEvent_t e)
{
    if (e) v->event_pool.push_back(e);
}

int fails(vfs_server *v, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int fails(vfs_server *v, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = fail_str(v ? &v->err : nullptr, code, fmt, ap);
    va_end(ap);
    return r;
}

#define HIPCHKS(v, expr)                                                                                     \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return fails((v), VFS_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

// Device bytes a resident stream needs (what vfs_initialize allocates).
size_t stream_bytes(const SrvStream &st, float fScale)
{
    size_t total = 0;
    int w = st.width, h = st.height;
    const size_t s0 = (size_t)align_up(w, 32);
    for (int l = 0; l < st.nLevels; l++) {
        const size_t s = (size_t)align_up(w, 32);
        total += s * h * sizeof(float) * 5;  // I0, I1, mask, tv (2)
        h = (int)((float)h / fScale);
        w = (int)((float)w / fScale);
    }
    const size_t N = s0 * st.height;
    const size_t plane = (N + 63) & ~(size_t)63;
    total += N * sizeof(float) * 8;  // inputs, cv, X, depth
    total += plane * sizeof(float) * vfs::PL_COUNT;
    return total;
}

int materialize(vfs_server *v, SrvStream &st)
{
    vfs_solver *s = nullptr;
    int rc = vfs_create(&s);
    if (rc) return fails(v, rc, "stream solver: %s", vfs_last_error(nullptr));
    s->storage = v->P.storage;
    s->persist_opt = 0;  // the server's launches cover many frames: no persistent launches
    rc = vfs_initialize(s, st.width, st.height, v->P.beta, v->P.gamma, v->P.alpha0, v->P.alpha1,
                        v->P.timestep_lambda, v->P.lambda, st.nLevels, v->P.fScale, st.nWarpIters, st.nSolverIters);
    if (!rc) rc = vfs_set_limit_range(s, v->P.limitRange);
    if (!rc) rc = vfs_copy_mask(s, st.mask.data(), st.width);
    if (!rc) rc = vfs_load_vector_fields(s, st.tv.data(), 2 * st.width, st.cv.data(), 2 * st.width);
    if (rc) {
        fails(v, rc, "stream materialization: %s", s->err.c_str());
        vfs_destroy(s);
        return rc;
    }
    s->has_images = true;
    if ((rc = tune_solver(s, 1))) {
        fails(v, rc, "stream tuning: %s", s->err.c_str());
        vfs_destroy(s);
        return rc;
    }
    st.solver = s;
    st.slot = v->free_slots.back();
    v->free_slots.pop_back();
    st.warm_ok = false;
    v->used += st.bytes;
    v->h_desc[st.slot] = s->desc();
    HIPCHKS(v, hipMemcpyAsync(v->d_desc + st.slot, &v->h_desc[st.slot], sizeof(vfs::PairDesc),
                              hipMemcpyHostToDevice, v->stream));
    HIPCHKS(v, hipStreamSynchronize(v->stream));
    return VFS_OK;
}

void evict(vfs_server *v, SrvStream &st)
{
    if (!st.solver) return;
    vfs_destroy(st.solver);
    st.solver = nullptr;
    v->free_slots.push_back(st.slot);
    st.slot = -1;
    st.warm_ok = false;
    v->used -= st.bytes;
}

// Makes room for `need` bytes (and a descriptor slot) by evicting idle resident streams:
// streams with no queued frame first, least recently used first.
bool make_room(vfs_server *v, size_t need, int keep)
{
    while (v->used + need > v->budget || v->free_slots.empty()) {
        int victim = -1;
        bool victim_queued = true;
        for (int i = 0; i < (int)v->streams.size(); i++) {
            const SrvStream &c = v->streams[i];
            if (i == keep || !c.solver || c.busy) continue;
            const bool queued = !c.queue.empty();
            if (victim < 0 || (victim_queued && !queued) ||
                (queued == victim_queued && c.last_use < v->streams[victim].last_use)) {
                victim = i;
                victim_queued = queued;
            }
        }
        if (victim < 0) return false;
        evict(v, v->streams[victim]);
    }
    return true;
}

int retire_finished(vfs_server *v)
{
    for (size_t i = 0; i < v->finishing.size();) {
        SrvRequest *r = v->finishing[i];
        hipError_t q = hipEventQuery(r->done);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        if (q != hipSuccess) return fails(v, VFS_ERR_DEVICE, "request event: %s", hipGetErrorString(q));
        SrvStream &st = v->streams[r->stream];
        if (r->out) memcpy(r->out, r->result, sizeof(float) * st.width * st.height);
        if (!r->prog.mon_slots.empty()) {
            vfs::EnergyPart rec;
            memcpy(&rec, r->result + (size_t)st.width * st.height, sizeof rec);
            r->nonfinite = (int)rec.nonfinite;
            r->energy = rec.energy;
        }
        r->t_done = now_ms();
        st.busy = false;
        st.warm_ok = true;
        v->done.push_back(r);
        v->finishing.erase(v->finishing.begin() + i);
    }
    while (!v->tick_events.empty() && hipEventQuery(v->tick_events.front()) == hipSuccess) {
        give_event(v, v->tick_events.front());
        v->tick_events.erase(v->tick_events.begin());
    }
    return VFS_OK;
}

// Admits queued frames, oldest first, while active slots are free: a frame is
// admissible when it is the oldest queued frame of an idle stream that is (or can be
// made) resident.
int admit(vfs_server *v)
{
    while ((int)v->active.size() < v->max_active && v->queued > 0) {
        int pick = -1;
        for (int i = 0; i < (int)v->streams.size(); i++) {
            const SrvStream &c = v->streams[i];
            if (c.busy || c.queue.empty()) continue;
            if (pick < 0 || c.queue.front()->seq < v->streams[pick].queue.front()->seq) pick = i;
        }
        if (pick < 0) break;
        SrvStream &st = v->streams[pick];
        if (!st.solver) {
            if (!make_room(v, st.bytes, pick)) {
                // not resident and no room now: try the other streams, retry next step
                bool other = false;
                for (int i = 0; i < (int)v->streams.size() && !other; i++) {
                    const SrvStream &c = v->streams[i];
                    other = i != pick && !c.busy && !c.queue.empty() && c.solver;
                }
                if (!other) break;
                // admit the oldest frame among resident idle streams instead
                int alt = -1;
                for (int i = 0; i < (int)v->streams.size(); i++) {
                    const SrvStream &c = v->streams[i];
                    if (c.busy || c.queue.empty() || !c.solver) continue;
                    if (alt < 0 || c.queue.front()->seq < v->streams[alt].queue.front()->seq) alt = i;
                }
                pick = alt;
            } else {
                int rc = materialize(v, st);
                if (rc) return rc;
            }
        }
        SrvStream &sp = v->streams[pick];
        SrvRequest *r = sp.queue.front();
        sp.queue.pop_front();
        v->queued--;
        vfs_solver *s = sp.solver;
        const size_t n = (size_t)sp.width * sp.height;
        const Dims d0 = s->lv[0];
        // the images go in on the copy stream (the stream's previous frame has retired, so
        // nothing reads its inputs any more); the frame's first launch waits for them
        HIPCHKS(v, hipMemcpy2DAsync(s->in_i0, sizeof(float) * d0.s, r->input, sizeof(float) * sp.width,
                                    sizeof(float) * sp.width, sp.height, hipMemcpyHostToDevice, v->copy_stream));
        HIPCHKS(v, hipMemcpy2DAsync(s->in_i1, sizeof(float) * d0.s, r->input + n, sizeof(float) * sp.width,
                                    sizeof(float) * sp.width, sp.height, hipMemcpyHostToDevice, v->copy_stream));
        HIPCHKS(v, take_event(v, &r->in_ready));
        HIPCHKS(v, hipEventRecord(r->in_ready, v->copy_stream));
        r->wait_input = true;
        Params P = v->P;
        P.nLevels = sp.nLevels;
        P.nWarpIters = sp.nWarpIters;
        P.nSolverIters = sp.nSolverIters;
        r->warm = sp.warm_ok ? 1 : 0;
        const Tuning &tn = s->tuning[1];
        r->prog = sp.warm_ok ? build_program(P, 1, sp.warm_level, sp.warm_warps, sp.warm_iters, s->roles, tn)
                             : build_program(P, 1, -1, 0, 0, 0, tn);
        r->pc = 0;
        s->roles = r->prog.final_roles;
        sp.busy = true;
        sp.last_use = ++v->clock;
        v->active.push_back(r);
    }
    return VFS_OK;
}

// One launch of the next op of every frame in `grp` (ops of one kernel: tick ops, or
// multi-iteration ops of one iteration count).
int launch_group(vfs_server *v, const std::vector<SrvRequest *> &grp)
{
    vfs::TickArgs a = tick_header(v->P);
    // one tile shape per launch: the one chosen for the entry with the most pixels
    int shape = 0;
    long long most = -1;
    for (SrvRequest *r : grp) {
        const ProgOp &o = r->prog.ops[r->pc];
        const Dims dl = v->streams[r->stream].solver->lv[vfs::op_tile_level(o.kind, o.level)];
        if ((long long)dl.w * dl.h > most) {
            most = (long long)dl.w * dl.h;
            shape = o.shape;
        }
    }
    int acc = 0, k = 0;
    for (SrvRequest *r : grp) {
        const SrvStream &st = v->streams[r->stream];
        const ProgOp &o = r->prog.ops[r->pc];
        const vfs::PairDesc &d = v->h_desc[st.slot];
        set_entry(a.e[k], o, st.slot, d, acc);
        acc += vfs::op_tiles(o.kind, o.level, d.w, d.h, d.s, shape, o.flags);
        k++;
    }
    a.n = k;
    a.total = acc;
    // frames whose images are still on their way in: the launch waits for them
    for (SrvRequest *r : grp)
        if (r->wait_input) {
            HIPCHKS(v, hipStreamWaitEvent(v->stream, r->in_ready, 0));
            r->wait_input = false;
        }
    HIPCHKS(v, vfs::launch_tick(v->d_desc, a, v->P.storage, shape, v->stream));
    HIPCHKS(v, hipGetLastError());
    v->ticks++;
    v->launched_entries += k;
    return VFS_OK;
}

// One step: one launch advances by one op every active frame whose next op runs on the
// step's kernel: the tick kernel, or the multi-iteration kernel of one iteration count.
// The step takes the kernel with the most such frames, unless a frame has waited
// kMaxWait steps, whose kernel then goes first; the other frames keep their op.  (Frames
// that wait catch up with the others, so the frames in flight gather at the same ops and
// share launches: launching every kernel's frames in each step instead spreads them out,
// 12 ops per launch at saturation against 14, and served 6 % fewer frames per second.)
constexpr int kMaxWait = 4;

int launch_step(vfs_server *v)
{
    if (v->active.empty()) return VFS_OK;
    constexpr int kGroups = 4;
    std::vector<SrvRequest *> groups[kGroups];  // tick ops, then ops of 2, 4 and 8 iterations
    int starving = -1;
    for (SrvRequest *r : v->active) {
        const int kind = r->prog.ops[r->pc].kind;
        const int g = kind == vfs::OP_ITER8 ? 3 : kind == vfs::OP_ITER4 ? 2 : kind == vfs::OP_ITER2 ? 1 : 0;
        groups[g].push_back(r);
        if (r->waited >= kMaxWait && starving < 0) starving = g;
    }
    int pick = starving;
    if (pick < 0) {
        pick = 0;
        for (int g = 1; g < kGroups; g++)
            if (groups[g].size() > groups[pick].size()) pick = g;
    }
    const int rc = launch_group(v, groups[pick]);
    if (rc) return rc;
    for (int g = 0; g < kGroups; g++)
        for (SrvRequest *r : groups[g]) r->waited = g == pick ? 0 : r->waited + 1;
    for (size_t i = 0; i < v->active.size();) {
        SrvRequest *r = v->active[i];
        if (r->waited > 0 || ++r->pc < r->prog.ops.size()) {
            i++;
            continue;
        }
        // program done: the result (u at level 0) and the last energy record go out on the
        // copy stream after the frame's last op; `done` follows them
        SrvStream &st = v->streams[r->stream];
        vfs_solver *s = st.solver;
        const float *u = s->arena + (long long)vfs::slot_u(vfs::rcur(s->roles, vfs::G_U)) * s->plane;
        HIPCHKS(v, take_event(v, &r->ops_done));
        HIPCHKS(v, hipEventRecord(r->ops_done, v->stream));
        HIPCHKS(v, hipStreamWaitEvent(v->copy_stream, r->ops_done, 0));
        HIPCHKS(v, hipMemcpy2DAsync(r->result, sizeof(float) * st.width, u, sizeof(float) * s->lv[0].s,
                                    sizeof(float) * st.width, st.height, hipMemcpyDeviceToHost, v->copy_stream));
        if (!r->prog.mon_slots.empty())
            HIPCHKS(v, hipMemcpyAsync(r->result + (size_t)st.width * st.height, s->mon + r->prog.mon_slots.back(),
                                      sizeof(vfs::EnergyPart), hipMemcpyDeviceToHost, v->copy_stream));
        HIPCHKS(v, take_event(v, &r->done));
        HIPCHKS(v, hipEventRecord(r->done, v->copy_stream));
        v->finishing.push_back(r);
        v->active.erase(v->active.begin() + i);
    }
    hipEvent_t ev;
    HIPCHKS(v, take_event(v, &ev));
    HIPCHKS(v, hipEventRecord(ev, v->stream));
    v->tick_events.push_back(ev);
    return VFS_OK;
}

}  // namespace

extern "C" {

int vfs_server_create(vfs_server **out, const vfs_server_params *p)
{
    if (!out || !p) return fails(nullptr, VFS_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->max_streams < 1 || p->max_active < 1 || p->max_active > vfs::kMaxEntries)
        return fails(nullptr, VFS_ERR_ARG, "max_streams >= 1 and 1 <= max_active <= %d required", vfs::kMaxEntries);
    if (p->storage < vfs::kF32 || p->storage > vfs::kBF16) return fails(nullptr, VFS_ERR_ARG, "bad storage");
    vfs_server *v = new (std::nothrow) vfs_server();
    if (!v) return fails(nullptr, VFS_ERR_NOMEM, "host allocation failed");
    v->P = Params{0, 0, 0, 32, p->storage, p->beta, p->gamma, p->alpha0, p->alpha1, p->timestep_lambda, p->lambda,
                  p->fScale, p->limitRange};
    v->budget = p->hbm_budget_bytes;
    v->max_active = p->max_active;
    v->max_slots = p->max_streams;
    v->h_desc.resize((size_t)v->max_slots);
    hipError_t e = hipGetDevice(&v->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&v->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&v->d_desc, sizeof(vfs::PairDesc) * v->max_slots);
    if (e != hipSuccess) {
        fails(nullptr, VFS_ERR_DEVICE, "server setup: %s", hipGetErrorString(e));
        if (v->d_desc) (void)hipFree(v->d_desc);
        if (v->stream) (void)hipStreamDestroy(v->stream);
        if (v->copy_stream) (void)hipStreamDestroy(v->copy_stream);
        delete v;
        return VFS_ERR_DEVICE;
    }
    for (int i = v->max_slots - 1; i >= 0; i--) v->free_slots.push_back(i);
    *out = v;
    return VFS_OK;
}

int vfs_server_destroy(vfs_server *v)
{
    if (!v) return VFS_ERR_ARG;
    (void)hipSetDevice(v->device);
    (void)hipStreamSynchronize(v->stream);
    (void)hipStreamSynchronize(v->copy_stream);
    auto drop = [](std::vector<SrvRequest *> &l) {
        for (SrvRequest *r : l) {
            for (hipEvent_t e : {r->in_ready, r->ops_done, r->done})
                if (e) (void)hipEventDestroy(e);
            if (r->own_input) (void)hipHostFree(r->own_input);
            if (r->result) (void)hipHostFree(r->result);
            delete r;
        }
        l.clear();
    };
    for (SrvStream &st : v->streams) {
        std::vector<SrvRequest *> q(st.queue.begin(), st.queue.end());
        st.queue.clear();
        drop(q);
    }
    drop(v->active);
    drop(v->finishing);
    drop(v->done);
    for (hipEvent_t e : v->tick_events) (void)hipEventDestroy(e);
    for (hipEvent_t e : v->event_pool) (void)hipEventDestroy(e);
    for (auto &b : v->pinned_pool) (void)hipHostFree(b.second);
    for (SrvStream &st : v->streams) evict(v, st);
    (void)hipFree(v->d_desc);
    (void)hipStreamDestroy(v->stream);
    (void)hipStreamDestroy(v->copy_stream);
    delete v;
    return VFS_OK;
}

const char *vfs_server_last_error(const vfs_server *v) { return v ? v->err.c_str() : g_create_error.c_str(); }

int vfs_server_add_stream(vfs_server *v, const vfs_stream_params *p, int *stream_id)
{
    if (!v || !p || !stream_id) return VFS_ERR_ARG;
    if (p->width < 1 || p->height < 1 || p->nLevels < 1 || p->nLevels > kMaxLevels || !p->mask || !p->tv || !p->cv)
        return fails(v, VFS_ERR_ARG, "bad stream geometry or fields");
    if (p->warm_level < 0 || p->warm_level >= p->nLevels)
        return fails(v, VFS_ERR_ARG, "warm start level %d outside [0, %d)", p->warm_level, p->nLevels);
    SrvStream st;
    st.width = p->width;
    st.height = p->height;
    st.nLevels = p->nLevels;
    st.nWarpIters = p->nWarpIters;
    st.nSolverIters = p->nSolverIters;
    st.warm_level = p->warm_level;
    st.warm_warps = p->warm_warps;
    st.warm_iters = p->warm_iters;
    const size_t n = (size_t)p->width * p->height;
    st.mask.assign(p->mask, p->mask + n);
    st.tv.assign(p->tv, p->tv + 2 * n);
    st.cv.assign(p->cv, p->cv + 2 * n);
    st.bytes = stream_bytes(st, v->P.fScale);
    if (st.bytes > v->budget) return fails(v, VFS_ERR_NOMEM, "stream needs %zu bytes, budget %zu", st.bytes, v->budget);
    *stream_id = (int)v->streams.size();
    v->streams.push_back(std::move(st));
    return VFS_OK;
}

}  // extern "C"

namespace {

// A request of stream_id with its result buffer; `input` and `own_input` are the caller's.
int new_request(vfs_server *v, int stream_id, SrvRequest **out)
{
    if (stream_id < 0 || stream_id >= (int)v->streams.size()) return fails(v, VFS_ERR_ARG, "bad stream id");
    const SrvStream &st = v->streams[stream_id];
    const size_t n = (size_t)st.width * st.height;
    SrvRequest *r = new (std::nothrow) SrvRequest();
    if (!r) return fails(v, VFS_ERR_NOMEM, "host allocation failed");
    r->result = take_pinned(v, n + kPinnedExtra);
    if (!r->result) {
        delete r;
        return fails(v, VFS_ERR_NOMEM, "pinned result buffer of %zu bytes failed", sizeof(float) * (n + kPinnedExtra));
    }
    *out = r;
    return VFS_OK;
}

void queue_request(vfs_server *v, SrvRequest *r, int stream_id, float *u_out, long long tag)
{
    r->tag = tag;
    r->stream = stream_id;
    r->out = u_out;
    r->t_submit = now_ms();
    r->seq = ++v->seq;
    v->streams[stream_id].queue.push_back(r);
    v->queued++;
}

}  // namespace

extern "C" {

int vfs_server_submit(vfs_server *v, int stream_id, const float *i0, const float *i1, float *u_out, long long tag)
{
    if (!v || !i0 || !i1) return VFS_ERR_ARG;
    SrvRequest *r = nullptr;
    int rc = new_request(v, stream_id, &r);
    if (rc) return rc;
    const size_t n = (size_t)v->streams[stream_id].width * v->streams[stream_id].height;
    r->own_input = take_pinned(v, 2 * n);
    if (!r->own_input) {
        give_pinned(v, r->result, n + kPinnedExtra);
        delete r;
        return fails(v, VFS_ERR_NOMEM, "pinned staging of %zu bytes failed", sizeof(float) * 2 * n);
    }
    memcpy(r->own_input, i0, sizeof(float) * n);
    memcpy(r->own_input + n, i1, sizeof(float) * n);
    r->input = r->own_input;
    queue_request(v, r, stream_id, u_out, tag);
    return VFS_OK;
}

int vfs_server_submit_pinned(vfs_server *v, int stream_id, const float *frame, float *u_out, long long tag)
{
    if (!v || !frame) return VFS_ERR_ARG;
    SrvRequest *r = nullptr;
    int rc = new_request(v, stream_id, &r);
    if (rc) return rc;
    r->input = frame;
    queue_request(v, r, stream_id, u_out, tag);
    return VFS_OK;
}

int vfs_host_alloc(size_t bytes, void **out)
{
    if (!out || bytes == 0) return VFS_ERR_ARG;
    *out = nullptr;
    const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(nullptr, VFS_ERR_NOMEM, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return VFS_OK;
}

int vfs_host_free(void *p)
{
    if (!p) return VFS_ERR_ARG;
    return hipHostFree(p) == hipSuccess ? VFS_OK : VFS_ERR_DEVICE;
}

int vfs_server_step(vfs_server *v, int *active, int *queued)
{
    if (!v) return VFS_ERR_ARG;
    HIPCHKS(v, hipSetDevice(v->device));
    int rc = retire_finished(v);
    if (!rc) rc = admit(v);
    if (!rc && (int)v->tick_events.size() < 8) rc = launch_step(v);
    else if (!rc && !v->tick_events.empty()) {
        // bounded run-ahead: wait for the oldest tick in flight
        HIPCHKS(v, hipEventSynchronize(v->tick_events.front()));
    }
    if (active) *active = (int)(v->active.size() + v->finishing.size());
    if (queued) *queued = (int)v->queued;
    return rc;
}

int vfs_server_collect(vfs_server *v, vfs_server_result *out, int max, int *n)
{
    if (!v || !n || (max > 0 && !out)) return VFS_ERR_ARG;
    int k = 0;
    while (k < max && !v->done.empty()) {
        SrvRequest *r = v->done.front();
        v->done.erase(v->done.begin());
        out[k].tag = r->tag;
        out[k].stream = r->stream;
        out[k].warm = r->warm;
        out[k].latency_ms = r->t_done - r->t_submit;
        out[k].ops = (int)r->prog.ops.size();
        out[k].nonfinite = r->nonfinite;
        out[k].energy = r->energy;
        for (hipEvent_t e : {r->in_ready, r->ops_done, r->done}) give_event(v, e);
        const SrvStream &st = v->streams[r->stream];
        const size_t npx = (size_t)st.width * st.height;
        give_pinned(v, r->result, npx + kPinnedExtra);
        if (r->own_input) give_pinned(v, r->own_input, 2 * npx);
        delete r;
        k++;
    }
    *n = k;
    return VFS_OK;
}

int vfs_server_stats(const vfs_server *v, vfs_server_stats_t *st)
{
    if (!v || !st) return VFS_ERR_ARG;
    st->ticks = v->ticks;
    st->entries = v->launched_entries;
    st->hbm_used = v->used;
    int resident = 0;
    for (const SrvStream &s : v->streams) resident += s.solver != nullptr;
    st->resident_streams = resident;
    return VFS_OK;
}

}  // extern "C"
