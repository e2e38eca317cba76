/*
 * dtk.h -- C ABI of the MI355X-native generalised geodesic distance
 * transform (GGDT) library (libdtk.so, HIP kernels for gfx950).
 *
 * This is the drop-in boundary for the reference's GGDT path
 * (baptisteesteban/distance_transform_benchmarks).  The reference exposes a
 * C++ API only; the C++ wrappers in include/dt/ (namespace dt) keep its
 * signatures and call these entry points.  Each entry point below names the
 * reference interface it replaces.
 *
 * Conventions
 *   - Images are pitched 2-D buffers in device memory (pitch in BYTES), the
 *     layout of dt::image2d_view (dt/include/dt/image2d_view.hpp:194-199):
 *     address(x, y) = data + y * pitch + x * elem_size.
 *   - Pixel types: DTK_U8 (grey), DTK_RGB8 (3 bytes r,g,b, the layout of
 *     dt::rgb8, dt/include/dt/rgb.hpp:7-12), DTK_F32.
 *   - Masks: DTK_U8 (mask != 0 marks a seed; D0 = v * (mask == 0),
 *     dt/src/utils.cu:12) or DTK_F32 soft masks s in [0, 1] (D0 = v * (1 - s),
 *     an extension of this library).
 *   - Errors are returned as dtk_status codes, never thrown across the ABI;
 *     dtk_last_error() returns a message for the calling thread.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are stream-ordered; the raster and blocks schedules read a
 *     convergence flag back to the host once per round, the task schedule
 *     does not synchronise.
 */
#ifndef DTK_H
#define DTK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
  DTK_OK                   = 0,
  DTK_ERR_INVALID_ARGUMENT = 1,
  DTK_ERR_OUT_OF_MEMORY    = 2,
  DTK_ERR_HIP              = 3,
  DTK_ERR_UNSUPPORTED      = 4,
  DTK_ERR_NOT_CONVERGED    = 5
} dtk_status;

typedef enum
{
  DTK_U8   = 0,
  DTK_RGB8 = 1,
  DTK_F32  = 2
} dtk_dtype;

typedef enum
{
  DTK_C4 = 4,
  DTK_C8 = 8
} dtk_connectivity;

typedef enum
{
  DTK_SCHED_RASTER = 0, /* full-image row/column passes (reference "raster")   */
  DTK_SCHED_BLOCKS = 1, /* host-driven rounds over dirty tiles ("blocks")       */
  DTK_SCHED_TASK   = 2  /* persistent kernel over a device work queue ("task")  */
} dtk_schedule;

/* A pitched 2-D image in device (or, for the transfer calls, host) memory. */
typedef struct
{
  void*     data;
  int       width;
  int       height;
  long long pitch; /* bytes between rows */
  int       dtype; /* dtk_dtype */
} dtk_image;

/* Parameters of one call.  Initialise with dtk_params_default(). */
typedef struct
{
  float lambda;         /* 0 <= lambda <= 1; l_grad = lambda, l_eucl = 1 - lambda */
  float v;              /* non-seed value (uint8 mask) / nu (soft mask); finite, >= 0; default 1e10 */
  int   connectivity;   /* DTK_C8 (reference) or DTK_C4 */
  int   schedule;       /* dtk_schedule */
  int   max_iterations; /* 0: run to convergence; N > 0: exactly N raster iterations (raster only) */
  int   tile_rows;      /* rows of one wave's region: 32 or 16 (0 = default: 16 for task, 32 for blocks) */
  int   max_sweeps;     /* sweep cap per tile visit (0 = until the tile is stable) */
  int   waves_per_cu;   /* persistent waves per CU (0 = occupancy limit) */
  int   cost_plane;      /* tile schedules, lambda > 0: 0 = compute edge costs from the image on every
                            visit (default), 1 = compute them once per call into a workspace plane
                            (16 B/px C8, 8 B/px C4) that every visit reads */
  int   tile_waves;      /* task schedule: waves stacked per tile, 1, 2 or 4; a tile owns 62 x tile_waves *
                            (tile_rows - 2) pixels.  0 = library default (4 with 16-row, 2 with 32-row waves) */
  int   queue_order;     /* task schedule: order in which dirty tiles are solved (the result never depends
                            on it): DTK_ORDER_DEFAULT, DTK_ORDER_FIFO, DTK_ORDER_KEY or DTK_ORDER_SEED */
  int*  round_per_blocks; /* optional device int[dtk_tile_count_for()]: visits per tile (blocks and task
                            schedules, single-image calls), added to the existing counts */
} dtk_params;

/* Queue orders of the task schedule (dtk_params.queue_order).
 *   FIFO  one ring: tiles are solved in the order they became dirty;
 *   KEY   64 priority buckets by the smallest improved value a wake carries
 *         (two buckets per octave of the distance): tiles near the current
 *         front of final values go first, which cuts re-visits;
 *   SEED  64 priority buckets by a static tile priority: the city-block
 *         tile distance to the nearest seed tile, quantised by its cdf, the
 *         priority of the reference's task queue (dt/src/block_priority.cu:74-155,
 *         dt/src/task_queue.cuh:40-63) on this library's tiling.
 * KEY and SEED defer a woken tile whose bucket lies beyond the current
 * front into a later ring; a worker that finds the main ring empty moves
 * the lowest later buckets into it.  DEFAULT is KEY for batched calls and
 * plans (n > 1), FIFO for one image. */
typedef enum
{
  DTK_ORDER_DEFAULT = 0,
  DTK_ORDER_FIFO    = 1,
  DTK_ORDER_KEY     = 2,
  DTK_ORDER_SEED    = 3
} dtk_queue_order;

/* Filled by calls that report statistics (may be NULL). */
typedef struct
{
  int      rounds;     /* host rounds (raster iterations / block rounds); 1 for task */
  uint32_t visits;     /* tile visits (blocks / task) */
  uint32_t sweeps;     /* tile sweeps (blocks / task) */
  int      tiles;      /* tiles of the call */
  int      tile_rows;  /* region rows used */
  int      workers;    /* persistent workers launched (task) */
  uint32_t cycles_wait;  /* task: summed worker shader-clock cycles / 1024 waiting for a tile */
  uint32_t cycles_visit; /*       ... solving tiles */
  uint32_t cycles_pub;   /*       ... publishing results (fence, wake-ups, re-queue) */
  uint32_t cycles_setup; /*   inside visits: tile loads and edge costs */
  uint32_t cycles_sweep; /*   inside visits: sweeps and layout changes */
  uint32_t cycles_store; /*   inside visits: border check and write-back */
  uint32_t cycles_release; /* inside publishing: the release fence after a changing visit */
  uint32_t cycles_acquire; /* inside waiting: claiming the popped tile and the acquire fence */
  uint32_t drains;       /* task, KEY / SEED orders: drains of the later rings that moved tiles */
  uint32_t moved;        /*   tiles those drains moved into the main ring */
  uint32_t deferred;     /*   wakes deferred into a later ring (seeds excluded) */
  uint32_t cycles_drain; /*   summed cycles / 1024 of the drains (inside waiting) */
  uint32_t kept;         /* task: visits that followed a visit of the same tile at once, by the same
                            worker, because the tile was woken during that visit (counted in visits) */
} dtk_stats;

void dtk_params_default(dtk_params* p);

/* Message of the last failed call on this thread ("" if none). */
const char* dtk_last_error(void);

/* Library version string. */
const char* dtk_version(void);

/* GGDT of one image, in place into `dist` (DTK_F32, same size).
 * Replaces dt::generalised_distance_transform (schedule RASTER,
 * dt/include/dt/generalised_distance_transform.hpp:7-8),
 * dt::generalised_distance_transform_blocks (BLOCKS, :13-15) and
 * dt::generalised_distance_transform_task (TASK, :21-23).  The converged
 * result is identical for every schedule.  The RASTER schedule keeps every
 * strip of a pass resident at once; on a 256-CU MI355X that allows at
 * least about 73 000 rows and 98 000 columns (C8 with lambda > 0, the
 * tightest case), and a larger image gives DTK_ERR_UNSUPPORTED. */
dtk_status dtk_ggdt(const dtk_image* img, const dtk_image* mask, dtk_image* dist, const dtk_params* p, void* stream,
                    dtk_stats* stats);

/* GGDT of n independent images of any sizes in ONE persistent launch that
 * shares a single tile queue (schedule TASK).  No reference counterpart
 * (the reference computes one image per call); each output equals the
 * single-image call. */
dtk_status dtk_ggdt_batched(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                            const dtk_params* p, void* stream, dtk_stats* stats);

/* Batched plan captured as a HIP graph: buffers are bound at creation, each
 * dtk_batch_plan_launch replays init + seed + persistent solve.  No
 * reference counterpart. */
typedef struct dtk_batch_plan dtk_batch_plan;
dtk_status dtk_batch_plan_create(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                                 const dtk_params* p, void* stream, dtk_batch_plan** plan);
dtk_status dtk_batch_plan_launch(dtk_batch_plan* plan, void* stream);
dtk_status dtk_batch_plan_destroy(dtk_batch_plan* plan);

/* Autotuner of batched plans.  First times the queue orders (KEY, FIFO,
 * SEED) at the default tile shape, then, under the fastest order, candidate
 * tile shapes (16-row waves stacked 4, 2 or 1 high, 32-row waves stacked 2
 * or 1 high), sweep caps per visit (max_sweeps 0 / 12) and persistent waves
 * per CU (the occupancy limit, 3/4 and 1/2 of it) on the given batch: each
 * candidate is a captured plan, launched once to warm up and then `reps`
 * times between HIP events on `stream`.  Writes *p with the fastest
 * queue_order / tile_rows / tile_waves / max_sweeps / waves_per_cu into
 * *best (default_ms: the library defaults).  Every candidate
 * computes the same bit-exact maps (the fixed point does not depend on the
 * schedule), so tuning never changes results; dists are overwritten.
 * Results are memoised per process and device by a signature of the batch
 * (count, pixel total, largest side, dtypes, connectivity, lambda, v).
 * No reference counterpart. */
typedef struct
{
  int   candidates; /* plans timed (0 when served from the memo) */
  int   from_cache; /* 1: *best came from the memo */
  float best_ms;    /* mean launch time of the chosen candidate */
  float default_ms; /* mean launch time of the default parameters */
} dtk_tune_report;
dtk_status dtk_batch_autotune(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                              const dtk_params* p, void* stream, int reps, dtk_params* best, dtk_tune_report* report);

/* Argmin predecessor codes of a converged map (uint8 image, same size as
 * dist): 0..7 = neighbour offset index in the order of dt::c8
 * (dt/include/dt/c8.hpp:19-28): (-1,0) (-1,-1) (0,-1) (1,-1) (1,0) (1,1)
 * (0,1) (-1,1); 8 = root (D == D0).  No reference counterpart. */
dtk_status dtk_ggdt_pred(const dtk_image* img, const dtk_image* mask, const dtk_image* dist, dtk_image* pred,
                         const dtk_params* p, void* stream);

/* Predecessor codes of n converged maps in one launch: arrays of n
 * descriptors, as for dtk_ggdt_pred; every image of the batch has the same
 * dtype, and so has every mask.  No reference counterpart. */
dtk_status dtk_ggdt_pred_batched(int n, const dtk_image* imgs, const dtk_image* masks, const dtk_image* dists,
                                 dtk_image* preds, const dtk_params* p, void* stream);

/* Adjoint of L = sum(W * D) through the predecessor forest, for DTK_F32
 * images and soft DTK_F32 masks: writes dL/dimage and dL/dmask (DTK_F32).
 * `grad_dist` holds dL/dD (= W).  No reference counterpart. */
dtk_status dtk_ggdt_backward(const dtk_image* img, const dtk_image* mask, const dtk_image* dist,
                             const dtk_image* pred, const dtk_image* grad_dist, dtk_image* grad_img,
                             dtk_image* grad_mask, const dtk_params* p, void* stream);

/* The adjoint of n images in one launch of each kernel: arrays of n
 * descriptors, as for dtk_ggdt_backward (grad_masks may be NULL; every mask
 * of the batch has the same dtype).  The subtree sums
 * climb predecessor chains (one dependent load per step, an atomic exchange
 * where chains merge), so an image costs as long as its longest chain; in
 * one launch the chains of all images climb together.  No reference
 * counterpart. */
dtk_status dtk_ggdt_backward_batched(int n, const dtk_image* imgs, const dtk_image* masks, const dtk_image* dists,
                                     const dtk_image* preds, const dtk_image* grad_dists, dtk_image* grad_imgs,
                                     dtk_image* grad_masks, const dtk_params* p, void* stream);

/* Tiles on a width x height image with single-wave regions of tile_rows
 * rows (0 = 32), i.e. 62 x (tile_rows - 2) owned pixels: enough entries for
 * round_per_blocks of dt::generalised_distance_transform_blocks
 * (dt/src/generalised_distance_transform_blocks.cu:18-19) under any default
 * tiling, whose tiles are taller.  dtk_tile_count_for gives a call's exact
 * count. */
int dtk_tile_count(int width, int height, int tile_rows);

/* Tiles a call with parameters *p uses on a width x height image, for any
 * schedule and tile shape: the size of round_per_blocks (visit counts per
 * tile of the blocks and task schedules).  0 for invalid arguments. */
int dtk_tile_count_for(int width, int height, const dtk_params* p);

/* Pitched device allocation (replaces cudaMallocPitch in
 * dt/include/dt/details/image2d_data_gpu.hpp:28). */
dtk_status dtk_malloc_pitch(void** ptr, long long* pitch, long long width_bytes, int height);
dtk_status dtk_free(void* ptr);

/* Pitched copies (replace dt::host_to_device / dt::device_to_host,
 * dt/src/transfert.cpp:10-28).  Synchronous. */
dtk_status dtk_host_to_device(const dtk_image* src, dtk_image* dst);
dtk_status dtk_device_to_host(const dtk_image* src, dtk_image* dst);

/* Raw pitched 2-D copy: kind 1 = host to device, 2 = device to host,
 * 3 = device to device.  Synchronous. */
dtk_status dtk_copy2d(void* dst, long long dst_pitch, const void* src, long long src_pitch, long long width_bytes,
                      int height, int kind);

/* Kernel timing for benchmarks (no reference counterpart; the reference
 * times whole calls with CUDA events, benchmarks/distance_transform.cpp:78-89).
 * Between begin and end, every task-schedule call records HIP events on its
 * own stream around its setup kernels (reset, init, seed) and around the
 * persistent solve; end waits for them and returns the summed durations and
 * the number of calls.  One profiler per process. */
dtk_status dtk_profile_begin(void);
dtk_status dtk_profile_end(double* setup_ms, double* solve_ms, int* calls);

/* GGDT of one image like dtk_ggdt with the task schedule, allocating
 * nothing: no hipMalloc, hipFree, pinned host memory, event or stream.  No
 * reference counterpart; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * The call's queue, tile statistics and control words live in a scratch
 * array of 16 MiB that every device holds from library load until the
 * process ends; a call that needs more (about 8192^2 pixels in FIFO order,
 * 4096^2 in KEY or SEED order) is rejected with DTK_ERR_UNSUPPORTED.  The
 * kernels clear what must start at zero on every call.  Every call on a
 * device shares that device's scratch, so calls that overlap on one device
 * (several streams, several threads) race: order them on one stream.  The
 * cost plane is not available (DTK_ERR_UNSUPPORTED).  Device-side errors
 * are returned at once when `stats` is non-NULL (the call then waits for
 * the stream), else kept until dtk_sync_errors. */
dtk_status dtk_ggdt_noalloc(const dtk_image* img, const dtk_image* mask, dtk_image* dist, const dtk_params* p,
                            void* stream, dtk_stats* stats);

/* Asynchronous errors.  The task schedule (dtk_ggdt, dtk_ggdt_batched,
 * batch plans) returns before its kernels finish, so a device-side error
 * (a soft mask value outside [0, 1] or NaN, a queue watchdog) cannot be
 * its return value.  The kernels leave it in a per-device mailbox that
 * every later call on the device reads first (a host read, no wait when
 * there is no error): that call then synchronises the device and returns
 * the error, once, with "(an earlier asynchronous call on this device)" in
 * its message.  dtk_sync_errors waits for `stream` and returns any such
 * error now.  Calls with a non-NULL stats pointer synchronise and return
 * their own error directly.  (The reference reports CUDA errors after a
 * synchronous call, dt/src/generalised_distance_transform.cu:136-137.) */
dtk_status dtk_sync_errors(void* stream);

/* Device self-check of the assumptions the kernels rely on: DPP lane
 * direction of wave shifts and correctly rounded sqrtf over every RGB8
 * squared distance.  out[0] = DPP ok, out[1] = number of sqrtf mismatches
 * (host reference), out[2] = device compute units. */
dtk_status dtk_selftest(int* out3);

/* Device properties used for launch shapes. */
int dtk_device_cu_count(void);

/* Display of a distance map on the device: the step after the transform in
 * the reference CLI (bin/distance_transform.cpp:66-70).  All images are
 * device images; the calls are asynchronous on stream except
 * dtk_minmax_f32, which returns the values to the host.
 *   dtk_minmax_f32    min / max of a f32 image as dt::minmax computes them
 *                     (dt/include/dt/normalize.hpp:28-43: NaNs skipped, max
 *                     starts at the smallest positive float);
 *   dtk_normalize_u8  f32 -> u8, (uint8)(double(x - min) / double(max - min)
 *                     * 255) truncated (dt/include/dt/normalize.hpp:45-62;
 *                     0/0 gives 0);
 *   dtk_inferno       u8 -> rgb8 through the inferno table
 *                     (dt/src/inferno.cpp:263-274);
 *   dtk_colorize      f32 -> rgb8, normalise and colour map in one pass. */
dtk_status dtk_minmax_f32(const dtk_image* img, float* min_max, void* stream);
dtk_status dtk_normalize_u8(const dtk_image* in, dtk_image* out, void* stream);
dtk_status dtk_inferno(const dtk_image* in, dtk_image* out, void* stream);
dtk_status dtk_colorize(const dtk_image* in, dtk_image* out, void* stream);

/* Tile size of a call with parameters *p (tile schedules): tiles are
 * tile_w x tile_h owned pixels, row-major, ceil(W / tile_w) a row; the
 * layout of round_per_blocks.  Returns 0 for invalid parameters. */
int dtk_tile_size_for(const dtk_params* p, int* tile_w, int* tile_h);

/* Per-tile values spread over their tiles' pixels: out(x, y) =
 * values[(y / tile_h) * tiles_x + x / tile_w] as float, for `values` a
 * device array of tiles_x * tiles_y int32 (value_bytes 4, e.g. visit counts)
 * or uint8 (value_bytes 1, e.g. tile priorities) and `out` a DTK_F32 device
 * image that the tile grid covers.  With dtk_colorize this gives the
 * reference's heat maps of per-block rounds and priorities: its
 * block_attribute_to_image kernel (bin/rounds_per_blocks.cu:13-23,
 * bin/priority_block.cu:12-22), for any tile size.  Asynchronous. */
dtk_status dtk_tile_image(const void* values, int value_bytes, int tiles_x, int tiles_y, int tile_w, int tile_h,
                          dtk_image* out, void* stream);

/* Tile priorities of the reference's task queue:
 * dt::compute_block_priorities (dt/include/dt/block_priority.hpp:7-8,
 * dt/src/block_priority.cu:74-155).  The image is cut into 32 x 32 tiles,
 * gw = ceil(W / 32) by gh = ceil(H / 32), N = gw * gh, row-major.
 *   mask        u8 (nonzero = seed) or f32 (nonzero = seed), on the device
 *   priorities  device array of N bytes: cdf[d] * 64 / (N + 1), where d is
 *               the tile's city-block distance (in tiles) to the nearest
 *               tile holding a seed and cdf[d] the number of tiles at
 *               distance <= d
 *   offsets     device array of 64 words: offsets[k] = number of tiles with
 *               priority < k (the reference's `cdf` argument)
 *   level0      host word: the number of tiles holding a seed
 * Synchronous on `stream` (it reads level0 back).  With no seed at all,
 * every priority is N * 64 / (N + 1) and level0 is 0. */
dtk_status dtk_block_priorities(const dtk_image* mask, uint8_t* priorities, uint32_t* offsets, uint32_t* level0,
                                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DTK_H */
