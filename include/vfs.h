/*
 * libvfs: C ABI of the gfx950 StereoTgv TGV-L1 fisheye-stereo solver.
 *
 * Every entry point replaces one member of the reference class `StereoTgv`
 * (menandro/vfs stereotgv/stereotgv.h:14-252); the comment on each names the
 * reference interface (file:line) it stands in for.  Conventions:
 *   - an opaque handle per solver instance; no process-global mutable state, so
 *     one handle per thread is safe;
 *   - caller-owned host buffers, row pitch given explicitly in ELEMENTS of the
 *     buffer's scalar type (floats for float planes, bytes for u8 planes);
 *     multi-channel planes are interleaved (2 floats per pixel for vector fields
 *     and flow, 3 floats per pixel for X);
 *   - every call returns VFS_OK (0) or a negative status and never exits the
 *     process; vfs_last_error() describes the last failure of a handle.
 * No HIP, CUDA or OpenCV type appears in this header.
 */
#ifndef VFS_H
#define VFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFS_OK 0
#define VFS_ERR_ARG (-1)     /* bad argument: null pointer, size, pitch, level */
#define VFS_ERR_STATE (-2)   /* call out of order (e.g. solve before initialize) */
#define VFS_ERR_DEVICE (-3)  /* HIP runtime / kernel failure */
#define VFS_ERR_NOMEM (-4)   /* device or host allocation failed */
#define VFS_ERR_NUMERIC (-5) /* the solve produced non-finite values (the outputs are still
                                copied); see vfs_get_energy */

typedef struct vfs_solver vfs_solver;

/* StereoTgv::StereoTgv() (stereotgv.cu:32-36): block 32x1, stride alignment 32. */
int vfs_create(vfs_solver **out);

/* StereoTgv::StereoTgv(int blockWidth, int blockHeight, int strideAlignment)
 * (stereotgv.cu:38-42).  blockWidth sets the column coverage of the tensor launch
 * (calcTensor.cu:81,227 launch ceil(h/blockWidth) blocks of blockWidth columns);
 * strideAlignment sets the row pitch of every plane (stereotgv.cu:571-580).
 * blockHeight is accepted and ignored (launch shapes are the library's own). */
int vfs_create_ex(vfs_solver **out, int blockWidth, int blockHeight, int strideAlignment);

/* ~StereoTgv() (stereotgv.h:18).  Unlike the reference, releases every buffer. */
int vfs_destroy(vfs_solver *s);

/* StereoTgv::initialize (stereotgv.h:107-109, stereotgv.cu:44-168).  Level sizes
 * follow the reference: h_{l+1} = (int)((float)h_l / fScale), same for w. */
int vfs_initialize(vfs_solver *s, int width, int height, float beta, float gamma, float alpha0,
                   float alpha1, float timestep_lambda, float lambda, int nLevels, float fScale,
                   int nWarpIters, int nSolverIters);

/* Public field StereoTgv::limitRange (stereotgv.h:45, default 1.0f). */
int vfs_set_limit_range(vfs_solver *s, float limit);
int vfs_get_limit_range(const vfs_solver *s, float *limit);

/* StereoTgv::copyMaskToDevice(cv::Mat CV_32F) (stereotgv.h:115, stereotgv.cu:241-252):
 * width x height floats, nonzero = solve here; builds the nearest-neighbour pyramid. */
int vfs_copy_mask(vfs_solver *s, const float *mask, int pitch);

/* StereoTgv::loadVectorFields(cv::Mat tv CV_32FC2, cv::Mat cv CV_32FC2)
 * (stereotgv.h:110, stereotgv.cu:171-199): trajectory field (downscaled per level)
 * and calibration offsets (level 0).  Pitches in floats (>= 2 * width). */
int vfs_load_vector_fields(vfs_solver *s, const float *tv, int tv_pitch, const float *cv, int cv_pitch);

/* StereoTgv::copyImagesToDevice(unsigned char*, unsigned char*) (stereotgv.h:112,
 * stereotgv.cu:201-213) and the CV_8U branch of copyImagesToDevice(cv::Mat, cv::Mat)
 * (stereotgv.cu:220-226): gray = u8 / 256. */
int vfs_copy_images_u8(vfs_solver *s, const uint8_t *i0, const uint8_t *i1, int pitch);

/* CV_32F branch of StereoTgv::copyImagesToDevice(cv::Mat, cv::Mat) (stereotgv.cu:227-230). */
int vfs_copy_images_f32(vfs_solver *s, const float *i0, const float *i1, int pitch);

/* CV_8UC3 branch of StereoTgv::copyImagesToDevice(cv::Mat, cv::Mat) (stereotgv.cu:231-237):
 * gray = (c0 + c1 + c2) / 3 / 256.  pitch in bytes. */
int vfs_copy_images_u8c3(vfs_solver *s, const uint8_t *i0, const uint8_t *i1, int pitch);

/* StereoTgv::solveStereoForwardMasked() (stereotgv.h:117, stereotgv.cu:389-515).
 * Asynchronous on the solver's stream; the copy-out calls wait for it. */
int vfs_solve_masked(vfs_solver *s);

/* StereoTgv::solveStereoForward() (stereotgv.h:116, stereotgv.cu:256-387): the
 * unmasked form (circle gate r < width/2.2, no median). */
int vfs_solve(vfs_solver *s);

/* Warm-started re-solve (no counterpart in the reference, whose solves always restart
 * from zero, stereotgv.cu:395-397): the previous solve's u, u_ and warpUV are downscaled
 * to `start_level` (scaled by 1/fScale per level, the scaled Downscale of
 * downscale.cu:148-237) and the masked (masked != 0) or unmasked solve runs from there
 * to level 0 with `warps` warps of `iters` iterations per level.  Requires a previous
 * solve of this handle; new images are normally copied in first. */
int vfs_solve_warm(vfs_solver *s, int masked, int start_level, int warps, int iters);

/* Copy-outs of results (vfs_copy_disparity_1d, vfs_copy_flow, vfs_copy_stereo,
 * vfs_copy_disparity_vis, vfs_copy_warped_image) and vfs_synchronize return
 * VFS_ERR_NUMERIC, after copying, when an energy record of the last solve counted
 * non-finite values (option "energy" >= 1; the reference returns 0 and hands NaN on). */

/* StereoTgv::copy1DDisparityToHost (stereotgv.h:121, stereotgv.cu:533-540): u, width x height. */
int vfs_copy_disparity_1d(vfs_solver *s, float *out, int pitch);

/* StereoTgv::copyDisparityToHost (stereotgv.h:122, stereotgv.cu:542-549): the 2-D flow
 * warpUV, 2 floats per pixel. */
int vfs_copy_flow(vfs_solver *s, float *out, int pitch);

/* StereoTgv::copyStereoToHost (stereotgv.h:118-120, stereotgv.cu:517-531): ConvertKB
 * (conversion.cu:3-69, evaluated in FP64 on the GPU) of warpUV into depth (1 float per
 * pixel) and X (3 floats per pixel).  Either output may be NULL. */
int vfs_copy_stereo(vfs_solver *s, float *depth, int depth_pitch, float *X, int X_pitch, float focalx,
                    float focaly, float cx, float cy, float d1, float d2, float d3, float d4, float t1,
                    float t2, float t3);

/* StereoTgv::copyDisparityVisToHost(cv::Mat&, float flowScale) (stereotgv.h:123,
 * stereotgv.cu:551-559): FlowToHSV colour coding of the flow warpUV / flowScale
 * (flowToHsv.cu:3-121), 3 floats per pixel in (B, G, R) order, values in [0, 1].
 * Pitch in floats (>= 3 * width). */
int vfs_copy_disparity_vis(vfs_solver *s, float *out, int pitch, float flowScale);

/* StereoTgv::copyWarpedImageToHost (stereotgv.h:124, stereotgv.cu:561-568): the last
 * warped I1 (valid at level 0 after a solve). */
int vfs_copy_warped_image(vfs_solver *s, float *out, int pitch);

/* No counterpart in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same copy as vfs_copy_warped_image, which in FP16 / BF16 storage allocates a temporary
 * float plane per call, with that plane staged in a static device array of 16 MiB (4 Mi
 * floats) that the library holds on each device until the process ends; it calls no
 * allocation routine.  A warped image of stride x height above 4 Mi floats is rejected
 * with VFS_ERR_ARG.  Calls that overlap on one device share the array and race. */
int vfs_copy_warped_image_noalloc(vfs_solver *s, float *out, int pitch);

/* No counterpart in the reference (its checkCudaErrors exits the process): text of the
 * last failure on this handle, "" if none.  Valid until the next call on the handle. */
const char *vfs_last_error(const vfs_solver *s);

/* Pyramid geometry (the reference's public pW/pH/pS vectors, stereotgv.h:68-70). */
int vfs_get_level_dims(const vfs_solver *s, int level, int *w, int *h, int *stride);

/* Inspection of an internal plane at `level` dims into a compact host array of
 * w*h*nch floats (the reference exposes these as public device pointers,
 * stereotgv.h:47-100).  Names: "u", "u_", "u_last", "a", "b", "c", "Iu", "Iz",
 * "i1warp", "I0", "I1", "mask", "warpUV", "tv2", "tv", "p", "q", "v_", "Tp"; and
 * "u.other", "u_.other", "v_.other": the ping-pong partner of u, u_, v_ (the
 * reference's second buffer of each pair, whose stale values its mask-edge reads see).
 * Returns the channel count (> 0) or a negative status. */
int vfs_read_plane(vfs_solver *s, const char *name, int level, float *out);

/* Blocks until all work queued on the solver's stream has finished. */
int vfs_synchronize(vfs_solver *s);

/* Energy monitor of a solve (no counterpart in the reference).  With option "energy" 1
 * (default) every level's last warp, with 2 every warp, ends with an FP64 reduction on the
 * GPU over the processed pixels (mask != 0; unmasked: inside the circle), from u, v_ and
 * the derivatives after the warp's last inner iteration:
 *   energy   = sum lambda |Iu (u - u_last) + Iz| + alpha1 |T (grad u - v_)| + alpha0 |grad v_|
 *   data     = the sum of the first term
 *   residual = sum |u - u_prev| (the last inner iteration's change of u)
 *   nonfinite = pixels whose terms are not finite (left out of the sums), pixels = processed
 * T = [[a, c], [c, b]] is the diffusion tensor; the gradients are the dual step's forward
 * differences with the backward fallback at mask / image edges (updateDualVariables.cu).
 * The sums run in a fixed order that does not depend on launch choices or batching.
 * vfs_get_energy waits for the last solve and returns the record of (level, warp), or
 * VFS_ERR_STATE if that solve wrote none (warps >= 128 share slot 127). */
typedef struct {
    double energy, data, residual;
    int nonfinite, pixels;
} vfs_energy_t;
int vfs_get_energy(vfs_solver *s, int level, int warp, vfs_energy_t *out);

/* Execution options:
 *   "graph"   1 = replay the solve as a captured HIP graph (default), 0 = launch
 *             kernels one by one;
 *   "storage" 0 = FP32 (default), 1 = FP16, 2 = BF16 storage of the image pyramid,
 *             warped image, derivatives, tensor and duals p, q; arithmetic stays FP32;
 *   "autotune" 1 = pick each level's tile shape, Tp prefetch and iterations per launch
 *             by timing the inner iteration in every choice at the first solve of a
 *             geometry (default; the choices are cached per process), 0 = fixed 64x8
 *             tiles, one iteration per launch;
 *   "shape"   -1 = as "autotune" says (default), 0..6 = force one tile shape for every
 *             level: 64x8, 64x4, 128x4, 32x16, 32x8, 16x32, 16x16;
 *   "per_launch" -1 = as "autotune" says (default), 1, 2, 4 or 8 = that many inner
 *             iterations per launch wherever allowed (the workgroup recomputes a halo
 *             of 3 pixels per iteration; levels with a mask pixel in column 0 run one;
 *             eight only in the tiles 64x4, 32x16, 32x8, 16x16, 20x20, 36x25, 27x25, 14x13,
 *             8x8, 7x7 and 4x4, else
 *             four);
 *   "pair"    the same as "per_launch" -1 / 1 / 2 written as -1 / 0 / 1;
 *   "energy"  0 = no energy records, 1 = after the last warp of every level (default),
 *             2 = after every warp (vfs_get_energy);
 *   "persist" -1 = as "autotune" says (default), 0 = never, 1 = wherever possible: a cold
 *             solve runs each warp's inner iterations in one launch of persistent
 *             workgroups, one per tile, that keep their tiles on chip and exchange
 *             borders every iteration (levels without a mask pixel in column 0, as many
 *             tiles as the device holds workgroups at once).  If such a launch finds not
 *             every workgroup resident (other kernels on the device), it gives up, and the
 *             solve runs again without persistent launches at the next call that waits
 *             for it; vfs_get_option "persist_reruns" counts those runs, and
 *             "persist_why" tells what the last launch that gave up waited for (its tile
 *             count << 12 | 1 arrivals, 2 ring values, 3 a neighbour's progress << 8 |
 *             iteration);
 *   "persist_give_up" 1 = the next solve's persistent launches find their abort flag set
 *             and leave at once, so that the solve runs again without them (tests of
 *             that path).
 * Results do not depend on "graph", "autotune", "shape", "per_launch", "pair" or
 * "persist".
 * No counterpart in the reference (FP32 only, stereotgv.h:47-100; its block shape is a
 * constructor argument, stereotgv.cu:32-42). */
int vfs_set_option(vfs_solver *s, const char *key, int value);
int vfs_get_option(const vfs_solver *s, const char *key, int *value);

/* The launch choices used at `level` by the masked (masked != 0) or unmasked solve:
 * tile shape index of the inner iterations (0..6 as option "shape"; with several
 * iterations per launch also 7..17 = 8x8, 4x4, 32x32, 20x20, 36x25, 27x25, 14x13, 7x7,
 * 43x34, 34x40, 53x25, and 1, 3, 4, 6 of the former), whether the inner iteration prefetches its Tp-plane
 * neighbours, and the inner iterations per launch, 1, 2, 4 or 8, or VFS_PERSISTENT when
 * cold solves run a warp's iterations in one persistent launch that exchanges tile
 * borders every iteration (tile shapes 7, 8, 11, 12, 13, 14, 17 or 18 = 53x50), or
 * VFS_PERSISTENT | K (K = 2, 4, 8) when that launch runs blocks of K iterations with
 * recomputed halos and exchanges the halo rings between blocks (tile shapes 6, 7, 8, 10,
 * 12, 13 with sides of at least 2K and a staged region of at most 2048 positions)
 * (`per_launch` may be NULL).  Runs the tuner first when it has not run for this
 * geometry.
 * No counterpart in the reference. */
int vfs_get_level_tuning(vfs_solver *s, int masked, int level, int *shape, int *preload, int *per_launch);
/* Sets the launch choices of one level (reproducible profiling, expert use); they hold
 * until the next initialize or option change.  `per_launch` (1, 2, 4 or 8) may carry
 * VFS_HALF_WORKGROUP: multi-iteration workgroups of 512 work-items, two or three
 * positions each, where the tile's staged region holds 513 .. 1536 positions; or it is
 * VFS_PERSISTENT, optionally | 2, 4 or 8: cold solves run each warp's iterations in one
 * persistent launch in tile shape `shape`, exchanging borders every iteration or halo
 * rings every K iterations (other solves keep the level's other choices).
 * No counterpart in the reference. */
#define VFS_HALF_WORKGROUP 0x100
#define VFS_PERSISTENT 0x200
int vfs_set_level_tuning(vfs_solver *s, int masked, int level, int shape, int preload, int per_launch);

/* Timing with HIP events on the solver's stream: runs `n` solves (masked != 0:
 * solveStereoForwardMasked) back to back on resident inputs and returns their total
 * wall time in ms. */
int vfs_time_solves(vfs_solver *s, int masked, int n, float *ms_total);

/* Timing of the inner TGV iteration at `level` (the dominant work): `reps` iterations
 * (rounded up to whole launches when the level runs several iterations per launch)
 * launched back to back on the current solver state, HIP events around them.  Returns
 * the time of `reps` iterations in ms and, in the last argument, the iterations per
 * launch (1, 2, 4 or 8).
 * The launches run on a copy of the solver's work buffers: the solver's state is left as
 * the last solve left it (a following warm start continues from it). */
int vfs_time_inner_iterations(vfs_solver *s, int level, int reps, float *ms_total, int *iters_per_launch);

/* Timing of a plain device-to-device copy kernel moving `bytes` (read + write), `reps`
 * times: the achievable-bandwidth reference for the solver's kernels. */
int vfs_time_copy(vfs_solver *s, size_t bytes, int reps, float *ms_total);

/* ---- batches: no counterpart in the reference, which solves one pair per call ----
 * A batch groups initialized solvers (each with its own mask, fields, images and
 * buffers; resolutions may differ) that share nLevels, nWarpIters, nSolverIters, the
 * solver parameters, limitRange and the storage option.  vfs_batch_solve runs
 * solveStereoForwardMasked (masked != 0) or solveStereoForward for all of them with
 * every kernel launch covering all pairs; afterwards each member's copy-out calls
 * return its own results.  The members must not be solved concurrently elsewhere. */
typedef struct vfs_batch vfs_batch;
int vfs_batch_create(vfs_batch **out, vfs_solver *const *solvers, int n);
int vfs_batch_destroy(vfs_batch *b);
int vfs_batch_solve(vfs_batch *b, int masked);
/* vfs_solve_warm for every member (all members solved last by the same program). */
int vfs_batch_solve_warm(vfs_batch *b, int masked, int start_level, int warps, int iters);
int vfs_batch_synchronize(vfs_batch *b);
/* Launch choices of the batch's launches at `level` (as vfs_get_level_tuning); the batch
 * times them on scratch copies of its largest members at its first solve. */
int vfs_batch_get_level_tuning(vfs_batch *b, int masked, int level, int *shape, int *preload, int *per_launch);
/* Fixes the batch's launch choices of one level (as vfs_set_level_tuning; reproducible
 * profiling): they replace the tuner's from the next batch solve on, with one iteration
 * per launch where a member's mask forbids more. */
int vfs_batch_set_level_tuning(vfs_batch *b, int masked, int level, int shape, int preload, int per_launch);
/* `n` batch solves back to back, HIP events around them; total ms. */
int vfs_batch_time_solves(vfs_batch *b, int masked, int n, float *ms_total);
const char *vfs_batch_last_error(const vfs_batch *b);

/* ---- serving: no counterpart in the reference ----
 * A continuous-batching server for many camera streams on one device.  Each stream keeps
 * its solver state resident (pyramids, duals, previous disparity) while the HBM budget
 * allows; its first frame (and the first after an eviction) is a full solve, every later
 * frame warm-starts from the stream's previous result (vfs_solve_warm semantics).  A
 * submitted frame is admitted as soon as its stream is idle, an active slot is free and
 * the stream is resident; each vfs_server_step launches one kernel that advances by one op
 * of its own schedule every admitted frame whose next op runs on that kernel, so frames
 * arriving while others are on the GPU join the running work.  A frame's result does not depend on what else is
 * running.  Solver parameters are per server; geometry and schedule are per stream. */
typedef struct vfs_server vfs_server;

typedef struct {
    size_t hbm_budget_bytes;  /* device memory the resident streams may use */
    int max_streams;          /* resident streams at most */
    int max_active;           /* frames in flight at most (1..64) */
    int storage;              /* 0 FP32, 1 FP16, 2 BF16 (vfs_set_option "storage") */
    float beta, gamma, alpha0, alpha1, timestep_lambda, lambda, fScale, limitRange;
} vfs_server_params;

typedef struct {
    int width, height;
    int nLevels, nWarpIters, nSolverIters;  /* full (cold) solve schedule */
    int warm_level, warm_warps, warm_iters; /* warm start schedule */
    const float *mask;                      /* width*height, copied */
    const float *tv, *cv;                   /* width*height*2 each, copied */
} vfs_stream_params;

typedef struct {
    long long tag;
    int stream;
    int warm;           /* 1 if the frame warm-started */
    int ops;            /* ops of its schedule */
    double latency_ms;  /* submit -> result in the caller's buffer */
    int nonfinite;      /* non-finite values counted by the frame's last energy record (level 0) */
    double energy;      /* that record's energy (vfs_energy_t) */
} vfs_server_result;

typedef struct {
    long long ticks;    /* kernel launches */
    long long entries;  /* (frame, op) entries over all launches */
    size_t hbm_used;
    int resident_streams;
} vfs_server_stats_t;

int vfs_server_create(vfs_server **out, const vfs_server_params *params);
int vfs_server_destroy(vfs_server *v);
int vfs_server_add_stream(vfs_server *v, const vfs_stream_params *params, int *stream_id);
/* Queues a frame (float images width*height, copied).  If u_out is not NULL it receives
 * the 1-D disparity (width*height floats) when the frame completes; it must stay valid
 * until the frame is collected. */
int vfs_server_submit(vfs_server *v, int stream_id, const float *i0, const float *i1, float *u_out,
                      long long tag);
/* Queues a frame without copying it: `frame` holds i0 then i1 (2*width*height floats) in
 * pinned host memory (vfs_host_alloc) and must stay unchanged until the frame is
 * collected; the server only reads it (one buffer may back several queued frames).  The
 * images go to the device on the server's copy stream, overlapping the launches. */
int vfs_server_submit_pinned(vfs_server *v, int stream_id, const float *frame, float *u_out, long long tag);
/* Page-locked host memory for vfs_server_submit_pinned (and any other host buffer). */
int vfs_host_alloc(size_t bytes, void **out);
int vfs_host_free(void *p);
/* One scheduling step: retire finished frames, admit queued ones, launch one kernel (waits
 * only when 8 steps are already in flight).  Reports frames in flight and queued. */
int vfs_server_step(vfs_server *v, int *active, int *queued);
/* Moves up to `max` finished frames into `out`; *n receives the count. */
int vfs_server_collect(vfs_server *v, vfs_server_result *out, int max, int *n);
int vfs_server_stats(const vfs_server *v, vfs_server_stats_t *stats);
const char *vfs_server_last_error(const vfs_server *v);

#ifdef __cplusplus
}
#endif

#endif /* VFS_H */
