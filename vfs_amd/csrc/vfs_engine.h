// Internal interface between the host solver (solver.hip) and the gfx950 kernels
// (kernels.hip) of the StereoTgv solve.
//
// A launch covers a *set of frame pairs*: every pair has a descriptor (PairDesc, in
// device memory) holding its level geometry, its pyramid planes and the base of its work
// arena.  Workgroups own rectangular tiles of one of kShapes shapes (chosen per launch); a
// prefix table of tile counts maps the flat workgroup index to (pair, tile), so pairs of
// different resolutions share one launch.  A single solve is the same code with one pair.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vfs {

constexpr int kMaxLevels = 16;
// Inner-iteration steps of a warp held in the schedule table of a pair descriptor; the
// multi-iteration ops run only within it.
constexpr int kMaxSched = 128;

// Tile shapes.  Shapes 0 .. kTickShapes - 1 are those of the tick kernel: a workgroup is
// 256 work-items laid out tw x (256 / tw); each work-item covers th / (256 / tw) rows (1
// or 2) of the tw x th tile.  Shape 0 is the default; the others are candidates of the
// per-level tile-shape tuner.  The multi-iteration ops (OP_ITER2, OP_ITER4, OP_ITER8) take the
// shapes of multi_iter_shape(), including the small tiles 8x8 and 4x4 for the coarsest,
// latency-bound levels, 32x32, 20x20, 36x25, 43x34, 34x40, 53x25 and 27x25 for the
// largest, and 14x13 and 7x7 between them (53x25, 34x40, 43x34 and 36x25 split an 848x800
// image into 512, 500, 480 and 768 tiles, two or three workgroups for each of 256 compute
// units, 53x25 with the smallest staged region of the two-round ones; 27x25, 14x13 and
// 7x7 split 424x400, 212x200 and 106x100 into 256, 256 and 240: one each).
// Shape 18 (53x50: 848x800 in 16 x 16 tiles) serves the whole-warp persistent op only.
constexpr int kShapes = 19, kTickShapes = 7;
struct ShapeDim {
    int tw, th;
};
constexpr ShapeDim kShape[kShapes] = {{64, 8}, {64, 4}, {128, 4}, {32, 16}, {32, 8}, {16, 32}, {16, 16},
                                      {8, 8}, {4, 4}, {32, 32}, {20, 20}, {36, 25}, {27, 25}, {14, 13}, {7, 7}, {43, 34}, {34, 40}, {53, 25},
                                      {53, 50}};
__host__ __device__ constexpr bool multi_iter_shape(int shape)
{
    return shape == 1 || shape == 3 || (shape >= 4 && shape != 5 && shape < 18);
}
// Positions a K-iteration workgroup stages: the tile grown by 2K pixels to the left / top
// and K to the right / bottom.
__host__ __device__ constexpr int iterk_region(int shape, int K)
{
    return (kShape[shape].tw + 3 * K) * (kShape[shape].th + 3 * K);
}
// A K-iteration launch takes one work-item per staged position up to 512 positions and
// 1024 above (two or three positions each); regions of 513 .. 1536 positions may instead
// take 512 work-items with two or three positions each (more workgroups per compute
// unit).  Launch shapes carry that choice as kHalfWorkgroup added to the tile shape index.
constexpr int kHalfWorkgroup = 32;
__host__ __device__ constexpr bool half_workgroup_ok(int shape, int K)
{
    return iterk_region(shape, K) > 512 && iterk_region(shape, K) <= 1536;
}
__host__ __device__ constexpr int shape_index(int launch_shape) { return launch_shape & (kHalfWorkgroup - 1); }
// Staged positions per work-item of a K-iteration launch in `shape` (K = 1: the whole-warp
// persistent op, whose staged region is the tile grown by 2 / 1 as well).
__host__ __device__ constexpr int staged_per_item(int shape, int K, bool half)
{
    const int nr = iterk_region(shape, K), nt = half ? 512 : nr >= 1024 ? 1024 : (nr + 63) / 64 * 64;
    return (nr + nt - 1) / nt;
}
static_assert(kShapes <= kHalfWorkgroup, "shape indices fit below the half-workgroup flag");
// Tile shapes of K-iteration launches: every multi-iteration shape for K = 2 and 4; for
// K = 8 those whose staged region holds at most 3072 positions (three per work-item).
__host__ __device__ constexpr bool iterk_ok(int shape, int K)
{
    return multi_iter_shape(shape) && (K <= 4 || iterk_region(shape, K) <= 3072);
}

// Whole-warp persistent op (OP_ITERP): every workgroup of the launch owns one tile for
// all of the op's iterations and keeps the tile's state on chip; after each iteration it
// publishes u_ and v_ of its two outermost rows and columns, from which its neighbours
// refresh the one- and two-pixel rings they stage around their own tiles (the tile grown
// by 2 pixels to the left / top and 1 to the right / bottom).  All workgroups must be
// resident at once.  Shapes: staged region of at most 3072 positions (three per
// work-item), tiles at least 3 pixels on each side.
__host__ __device__ constexpr int iterp_region(int shape) { return (kShape[shape].tw + 3) * (kShape[shape].th + 3); }
__host__ __device__ constexpr bool persist_shape(int shape)
{
    return (shape == 7 || shape == 8 || (shape >= 11 && shape <= 14) || shape == 17 || shape == 18) &&
           iterp_region(shape) <= 3072;
}
// Exchange words of one tile and parity: u_, v_.x, v_.y of columns tw-1, tw-2, rows
// th-1, th-2, column 0 and row 0, and the tile's progress word (StripP in tgv_device.h).
__host__ __device__ constexpr int persist_words(int shape) { return 9 * (kShape[shape].tw + kShape[shape].th) + 1; }
// The order in which its work-items take the staged positions (position q = tid + k NT):
// the tile's border first (columns 0, tw - 2, tw - 1 and rows 0, th - 2, th - 1 of the
// tile, the positions its strips hold), then the rest of the staged region (the tile grown
// by 2 / 1), each in raster order.  persist_rank maps a raster index of the staged region
// to its place in that order.
__host__ __device__ constexpr int persist_border(int shape) { return 3 * (kShape[shape].tw + kShape[shape].th) - 9; }
__host__ __device__ constexpr int persist_rank(int shape, int i)
{
    const int tw = kShape[shape].tw, th = kShape[shape].th, rw = tw + 3;
    const int dx = i % rw - 2, dy = i / rw - 2;
    // border positions in tile rows [0, y), and in tile row y before column x (x <= tw)
    auto rows_before = [=](int y) {
        const int full = (y > 0) + (y > th - 2) + (y > th - 1);
        return full * tw + (y - full) * 3;
    };
    auto in_row_before = [=](int y, int x) {
        return (y == 0 || y == th - 2 || y == th - 1) ? x : (x > 0) + (x > tw - 2) + (x > tw - 1);
    };
    const bool in_tile = dx >= 0 && dx < tw && dy >= 0 && dy < th;
    if (in_tile && (dy == 0 || dy == th - 2 || dy == th - 1 || dx == 0 || dx == tw - 2 || dx == tw - 1))
        return rows_before(dy) + in_row_before(dy, dx);
    int before = rows_before(dy < 0 ? 0 : dy > th ? th : dy);
    if (dy >= 0 && dy < th) before += in_row_before(dy, dx < 0 ? 0 : dx);
    return persist_border(shape) + i - before;
}

// Blocked persistent op (OP_ITERPK): the K-iteration op's workgroups made persistent.  A
// workgroup stages its tile grown by 2K / K once, runs blocks of K iterations as the
// K-iteration op does (an even count in all, the last block shortened to the rest), and
// between blocks refreshes its ring (everything it staged outside
// its tile) from the state its neighbours publish: u, u_, v_, p and q of the bands of
// their tiles that neighbouring rings cover.  Shapes: the tile at least 2K pixels on each
// side (a ring then reads its 8 neighbours only), at most two staged positions per
// work-item, and one of the tilings that serve the coarse levels.
__host__ __device__ constexpr bool persistk_shape(int shape, int K)
{
    return (K == 2 || K == 4 || K == 8) && shape >= 0 && shape < kShapes &&
           (shape == 6 || shape == 7 || shape == 8 || shape == 10 || shape == 12 || shape == 13) &&
           kShape[shape].tw >= 2 * K && kShape[shape].th >= 2 * K && iterk_region(shape, K) <= 2048;
}
// Its exchange words of one tile and parity: the tile's u, u_, v_.x, v_.y, p.x, p.y, q.x ..
// q.w (value-major, row-major within a value; only the bands are written), then the
// tile's progress word.
__host__ __device__ constexpr int persistk_words(int shape) { return 10 * kShape[shape].tw * kShape[shape].th + 1; }
// Tiles of one persistent launch at most (bounds the exchange buffer).
constexpr int kMaxPersistTiles = 1024;
// Launch bookkeeping of the persistent op, one per pair (device memory): the launch
// counter that makes exchange tags unique, the workgroups that have arrived and
// finished, and a sticky abort flag set when a workgroup waited too long (not every
// workgroup resident); the host clears it after re-running the solve without the op.
struct XSync {
    unsigned int epoch, arrive, done, abort;
    // the first give-up since the flag was cleared: the launch's tile count << 12 | what it
    // waited for (1 arrivals, 2 ring words, 3 a neighbour's progress word) << 8 | the
    // iteration (low 8 bits)
    unsigned int why, pad[3];
};

// Storage precision of the image pyramid, warped image, derivatives, tensor and duals.
// Arithmetic is FP32 in every mode.
enum Storage : int { kF32 = 0, kF16 = 1, kBF16 = 2 };

// Work planes of one pair: slot k starts at arena + k * plane (floats).  2- and 4-channel
// planes take 2 / 4 consecutive slots.  Planes marked (st) hold the storage type.  The
// state planes u, u_, v_ and the duals (p, q) form ping-pong groups of three buffers each:
// the current state, the other state (written by the next iteration) and a spare that
// the multi-iteration ops write; warpUV has two buffers.
enum Plane : int {
    PL_U = 0,                // u: 3 buffers
    PL_UB = 3,               // u_: 3 buffers
    PL_ULAST = 6,
    PL_I1W = 7,              // warped I1 (st)
    PL_IU = 8, PL_IZ = 9,    // derivatives (st)
    PL_A = 10, PL_B = 11, PL_C = 12,  // tensor (st)
    PL_V = 13,               // v_: 3 buffers (float2)
    PL_P = 19,               // p: 3 buffers (2 x st)
    PL_TP = 25,              // Tp (float2)
    PL_TV2 = 27,             // tv2 (float2)
    PL_UV0 = 29, PL_UV1 = 31,  // warpUV ping-pong (float2)
    PL_Q = 33,               // q: 3 buffers (4 x st)
    PL_I0D = 45,             // the I0 half of the derivative along the trajectory (float; per level)
    PL_COUNT = 46
};

// Ping-pong groups and their roles.  `roles` holds, for group g, the buffer index (0..2)
// of the current state in bits 4g..4g+1 and of the other state in bits 4g+2..4g+3; the
// remaining index is the spare.  Bit 16 selects the current warpUV buffer.
enum Group : int { G_U = 0, G_UB = 1, G_V = 2, G_PQ = 3 };
constexpr int RB_UV = 1 << 16;
constexpr int kRolesInit = 0x4444;  // every group: current 0, other 1, spare 2

__host__ __device__ inline int rcur(int roles, int g) { return (roles >> (4 * g)) & 3; }
__host__ __device__ inline int roth(int roles, int g) { return (roles >> (4 * g + 2)) & 3; }
__host__ __device__ inline int rspare(int roles, int g) { return 3 - rcur(roles, g) - roth(roles, g); }
__host__ __device__ inline int rset(int roles, int g, int cur, int oth)
{
    return (roles & ~(15 << (4 * g))) | ((cur | (oth << 2)) << (4 * g));
}
// One iteration: the other state becomes current.
__host__ __device__ inline int rswap(int roles, int g) { return rset(roles, g, roth(roles, g), rcur(roles, g)); }
// Two iterations in one op: the spare (written last) becomes current, the other state
// stays, the old current becomes the spare.
__host__ __device__ inline int rrotate(int roles, int g) { return rset(roles, g, rspare(roles, g), roth(roles, g)); }
// The four groups an inner iteration advances.
__host__ __device__ inline int rswap_iter(int roles)
{
    return rswap(rswap(rswap(rswap(roles, G_U), G_UB), G_V), G_PQ);
}
__host__ __device__ inline int rrotate_iter(int roles)
{
    return rrotate(rrotate(rrotate(rrotate(roles, G_U), G_UB), G_V), G_PQ);
}
// Slot of buffer k of the u, u_, v_, p and q planes.
__host__ __device__ inline int slot_u(int k) { return PL_U + k; }
__host__ __device__ inline int slot_ub(int k) { return PL_UB + k; }
__host__ __device__ inline int slot_v(int k) { return PL_V + 2 * k; }
__host__ __device__ inline int slot_p(int k) { return PL_P + 2 * k; }
__host__ __device__ inline int slot_q(int k) { return PL_Q + 4 * k; }
__host__ __device__ inline int slot_uv(int roles, bool current) { return ((roles & RB_UV) != 0) == current ? PL_UV1 : PL_UV0; }

// Energy / residual monitor (OP_ENERGY, OP_EFINAL): FP64 sums over the processed pixels
// of one pair at the end of a warp's inner iterations.  energy = data + alpha1 |T (grad u -
// v_)| + alpha0 |grad v_|, data = lambda |Iu (u - u_last) + Iz|, residual = |u - u_prev|
// (the last iteration's change); pixels whose terms are not finite are counted in
// `nonfinite` and left out of the sums.  A record is written per (level, warp) slot.
struct EnergyPart {
    double energy, data, residual;
    unsigned int nonfinite, pixels;
};
static_assert(sizeof(EnergyPart) == 32, "EnergyPart layout");
// Pixels per OP_ENERGY workgroup (256 work-items, 4 pixels each, in row-major order of the
// level): the partial sums, and so the result, do not depend on the launch's tile shape.
constexpr int kEnergyChunk = 1024;
// Warp slots per level in a pair's record array (later warps share the last slot).
constexpr int kMonWarps = 128;
__host__ __device__ inline int mon_slot(int level, int warp) { return level * kMonWarps + (warp < kMonWarps ? warp : kMonWarps - 1); }
// Partial sums of an OP_ENERGY launch for a level of w x h pixels.
__host__ __device__ inline int energy_chunks(int w, int h) { return (int)(((long long)w * h + kEnergyChunk - 1) / kEnergyChunk); }

struct PairDesc {
    float *arena;
    long long plane;       // floats per plane slot (>= s0 * h0, multiple of 64)
    const float *in_i0;    // level-0 input images (float)
    const float *in_i1;
    const float2 *cv;      // calibration offsets (level 0)
    int w[kMaxLevels], h[kMaxLevels], s[kMaxLevels];
    void *I0[kMaxLevels];  // image pyramids (st); in FP32 storage I0[0] == in_i0
    void *I1[kMaxLevels];
    const float *mask[kMaxLevels];
    const float2 *tv[kMaxLevels];
    // step sizes (sigma, tau, mu) of inner iteration k of every warp
    // (stereotgv.cu:449-477) and k1 = alpha1 sigma / 3 (float product, correctly rounded
    // division); the multi-iteration ops read them
    float4 sched[kMaxSched];
    EnergyPart *mon_rec;   // kMaxLevels x kMonWarps records (mon_slot)
    EnergyPart *mon_part;  // energy_chunks(w0, h0) partial sums of the running OP_ENERGY
    // OP_ITERP / OP_ITERPK: exchange words ([tile][parity][persist_words or
    // persistk_words], value | tag << 32) and the launch bookkeeping
    unsigned long long *xbuf;
    XSync *xsync;
};

// Scalars of one inner iteration (schedule of stereotgv.cu:449-477).
struct IterParams {
    float k1;      // alpha1 * sigma / eta_p
    float k0;      // alpha0 * sigma / eta_q
    float tau;
    float mu;
    float lambda;
    float alpha0;
    float alpha1;
    int first;       // p and q were just zeroed (stereotgv.cu:431-432): read them as 0
    int write_tp;    // also store Tp (read later only where a neighbour is outside the mask)
    int preload_tp;  // fetch the Tp-plane neighbours up front (latency-bound levels)
    int deriv_level;  // >= 0: the warp's derivatives Iu, Iz at that level are computed
                      // here (EF_DERIV) and stored, instead of loaded
};

// Ops of a solve program.  A program is the op sequence of one solve (built on the
// host); k_tick executes one op for each of up to kMaxEntries (pair, op) entries.
enum OpKind : int {
    OP_ZERO = 0,   // zero the work arena (fresh state)
    OP_CALIB = 1,  // calibration warp of I1 into the level-0 pyramid slot
    OP_DOWN = 2,   // I0/I1 pyramid level `level` from level-1
    OP_CINIT = 3,  // coarsest level: warpUV = u * tv
    OP_WINIT = 4,  // warm start: u, u_, warpUV of level-1 downscaled (x 1/fScale) to `level`
    OP_TENSOR = 5,
    OP_PREP = 6,   // tv2, I1w, u_last, v_ = 0
    OP_DERIV = 7,
    OP_ITER = 8,   // one fused TGV iteration
    OP_END = 9,    // [median +] limit + update at the end of a warp
    OP_UP = 10,    // upscale u, u_, warpUV from `level` to level-1
    OP_ITER2 = 11,  // two fused TGV iterations (recomputed halo)
    OP_ITER4 = 12,  // four fused TGV iterations (recomputed halo)
    OP_ENDPREP = 13, // END of a warp and PREP of the next one at the same level
    OP_ITER8 = 14,  // eight fused TGV iterations (recomputed halo)
    OP_SYNC = 15,   // spare buffers of u, u_, v_ take the current ones beyond the level's extent
    OP_ENERGY = 16, // energy / residual partial sums of the level (kEnergyChunk pixels per workgroup)
    OP_EFINAL = 17, // one workgroup per pair: sums the partials in order into the (level, warp) record
    OP_ITERP = 18,  // an even number of fused TGV iterations (a whole warp), persistent workgroups
    OP_ITERPK = 19  // blocks of K fused TGV iterations (halo recomputed), persistent workgroups
};

// Iterations of an inner-iteration op (OP_ITERP: see iterp_count).
inline int op_iterations(int kind)
{
    return kind == OP_ITER8 ? 8 : kind == OP_ITER4 ? 4 : kind == OP_ITER2 ? 2 : kind == OP_ITER ? 1 : 0;
}
// OP_ITERP / OP_ITERPK entries carry (iterations << 16) | (ITERPK: K << 8) | first
// iteration's index in `it0`.
__host__ __device__ inline int iterp_count(int it0) { return it0 >> 16; }
__host__ __device__ inline int iterp_first(int it0) { return it0 & 0xff; }
__host__ __device__ inline int iterp_block(int it0) { return (it0 >> 8) & 0xff; }
inline bool op_multi(int kind) { return kind == OP_ITER2 || kind == OP_ITER4 || kind == OP_ITER8; }

// Entries per launch: one per lane of a wave64 (the kernel's entry lookup), and 64
// entries of 56 bytes keep the kernel arguments under 4 KiB.
constexpr int kMaxEntries = 64;

// Flags of a TickEntry.
enum EntryFlag : int {
    EF_FIRST = 1,          // inner-iteration ops: p and q were just zeroed
    EF_WRITE_TP = 2,       // ITER: the iteration also stores Tp
    EF_PRELOAD_TP = 4,     // ITER: fetch the Tp-plane neighbours with the other operands
    EF_WRITE_TP_LAST = 8,  // ITER2 / ITER4 / ITER8: the op's last iteration stores Tp
    EF_EFINAL = 16,        // END / ENDPREP: one extra workgroup adds the level's energy partials
                           // into record (level, it0), the work of OP_EFINAL
    EF_SYNC = 32,          // DERIV: extra workgroups do OP_SYNC's copy beyond the level's extent
    EF_PREP = 64,          // TENSOR: the same workgroups also run OP_PREP at each pixel
    EF_DERIV = 128         // ITER2 / ITER4 / ITER8 / ITERP / ITERPK, a warp's first op: the
                           // derivatives (OP_DERIV's work) computed at every staged position
                           // and stored for the tile, in place of an OP_DERIV op (the host
                           // folds it where the kernel stages at most two positions per
                           // work-item, three in OP_ITERP: in the K-iteration kernels with
                           // three the derivatives' dependent loads would be taken before the
                           // staging of every launch that carries the flag)
};

// One (pair, op) of a launch.  The geometry and arena of the op's tile level travel in
// the kernel arguments, so the ITER op needs no descriptor load before its data loads.
struct TickEntry {
    float *arena;              // the pair's work arena
    const float *mask;         // mask plane of the tile level
    int plane;                 // floats per plane slot
    int tile_start;            // first workgroup of this entry (prefix sum)
    unsigned short pair;       // index into the descriptor array
    unsigned short w, h, s;    // tile level dims and row stride (OP_ZERO: level-0 stride x height)
    unsigned char kind, level, masked, flags;
    unsigned int roles;        // buffer roles before the op (see Group)
    float sigma, tau, mu;      // ITER: step sizes of the iteration
    int it0;                   // ITER2 / ITER4 / ITER8: index of the first iteration in the warp;
                               // ITERP / ITERPK: (iterations << 16) | (ITERPK: K << 8) |
                               // that index;
                               // EFINAL: the warp of the record
};
static_assert(sizeof(TickEntry) == 56, "TickEntry layout");

struct TickArgs {
    int n;      // entries
    int total;  // workgroups
    float lambda, alpha0, alpha1, beta, gamma, fscale, scale, limit;
    int xblock;  // calcTensor launch block width (column limit of the tensor)
    TickEntry e[kMaxEntries];
};

// Level whose pixels an op's tiles cover (OP_UP writes the finer level).
inline int op_tile_level(int kind, int level) { return kind == OP_UP ? level - 1 : kind == OP_SYNC ? 0 : level; }
// Ops whose tiles cover the level-0 planes' whole flat extent, row padding included.
inline bool op_flat(int kind) { return kind == OP_ZERO || kind == OP_SYNC; }

// Flat positions of the level-0 planes beyond level `level`'s extent (stride x height), in
// chunks of kSyncChunk per workgroup (EF_SYNC).
constexpr int kSyncChunk = 1024;
inline int sync_chunks(int level, const int *h, const int *s)
{
    const long long n = (long long)s[0] * h[0] - (long long)s[level] * h[level];
    return n > 0 ? (int)((n + kSyncChunk - 1) / kSyncChunk) : 0;
}

// Workgroups an op of `kind` at `level` with `flags` needs for a pair with dims w/h/s in
// tile `shape`.
inline int op_tiles(int kind, int level, const int *w, const int *h, const int *s, int shape, int flags = 0)
{
    if (kind == OP_ENERGY) return energy_chunks(w[level], h[level]);
    if (kind == OP_EFINAL) return 1;
    const int tl = op_tile_level(kind, level);
    const int ww = op_flat(kind) ? s[0] : w[tl];
    const ShapeDim d = kShape[shape_index(shape)];
    int n = ((ww + d.tw - 1) / d.tw) * ((h[tl] + d.th - 1) / d.th);
    if ((kind == OP_END || kind == OP_ENDPREP) && (flags & EF_EFINAL)) n += 1;
    if (kind == OP_DERIV && (flags & EF_SYNC)) n += sync_chunks(level, h, s);
    return n;
}

// Launches one op for the entries of `args` (all of one kind) in tile shape `shape`;
// hipErrorInvalidValue (nothing launched) when no kernel exists for that kind and shape.
// OP_ITERP / OP_ITERPK launches hold one entry.
hipError_t launch_tick(const PairDesc *desc, const TickArgs &args, int storage, int shape, hipStream_t st);
// Workgroups of the persistent op (tile shape `shape`; K = 1: OP_ITERP, else OP_ITERPK
// with blocks of K) that the device holds at once: its compute units times the workgroups
// one of them holds; 0 without such a kernel.
int persist_capacity(int storage, int shape, int K = 1);

// ---- single-plane kernels (load time, outputs) ----
struct Dims {
    int w, h, s;
};
void launch_fill_zero(void *ptr, size_t bytes, hipStream_t st);
void launch_cv8u(const uint8_t *in, int in_pitch, float *out, Dims d, hipStream_t st);
void launch_cv8uc3(const uint8_t *in, int in_pitch, float *out, Dims d, hipStream_t st);
void launch_downscale_f2(const float2 *src, Dims fine, float2 *out, Dims coarse, hipStream_t st);
void launch_downscale_nn(const float *src, Dims fine, float *out, Dims coarse, hipStream_t st);
void launch_convert_kb(const float2 *uv, double focalx, double focaly, double cx, double cy, double d1,
                       double d2, double d3, double d4, double t1, double t2, double t3, float *X,
                       float *depth, Dims d, hipStream_t st);
void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
// FlowToHSV colour coding of a flow plane into 3 floats (B, G, R) per pixel.
void launch_flow_hsv(const float2 *uv, Dims d, float *bgr, float scale, hipStream_t st);
// Copies a storage-typed plane into a float plane (inspection).
void launch_widen(const void *src, int storage, int nch, float *dst, Dims d, hipStream_t st);

}  // namespace vfs
