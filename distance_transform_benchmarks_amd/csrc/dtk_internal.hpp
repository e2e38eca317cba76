// Structures shared by the HIP kernels and the host side of libdtk.
//
// Tiling.  A tile owns a 62 x (RH-2) block of pixels and is solved by one
// wave64 over a 64 x RH "region": the owned block plus a one-pixel halo on
// every side.  Lane i of the wave holds region column i; with S row
// segments RH = 64 / S (S = 2: 64 x 32 regions, 62 x 30 owned pixels).
#pragma once

#include <stdint.h>

namespace dtk
{
  enum ImgKind : int
  {
    IMG_U8   = 0,
    IMG_RGB8 = 1,
    IMG_F32  = 2,
  };

  enum MaskKind : int
  {
    MASK_U8  = 0,
    MASK_F32 = 1,
  };

  constexpr int kTileW = 62; // owned columns per tile (64 lanes minus the halo)

  // Default region rows per wave of the task schedule (stacked four high).
  constexpr int kDefaultTaskRows = 16;

  // One image of a call (or of a batch).  Pitches are in bytes.
  struct ImgDesc
  {
    const unsigned char* img;
    long long            img_pitch;
    const unsigned char* mask;
    long long            mask_pitch;
    float*               D;
    long long            D_pitch;
    unsigned char*       pred;
    long long            pred_pitch;
    // Cached edge costs (tile schedules, lambda > 0): per pixel the costs of
    // its W, N, NW, NE edges (C8, 16 B) or W, N (C4, 8 B); NULL = recompute
    // them from the image on every tile visit.
    unsigned char*       cost;
    long long            cost_pitch;
    int                  W, H;
    int                  tiles_x, tiles_y;
    int                  tile0; // first global tile id of this image
    int                  pad_;
  };

  // One image of an adjoint launch (dtk_ggdt_backward[_batched]): its
  // inputs (d: image, mask, D, predecessor codes), dL/dD, the outputs and
  // its scratch (subtree sums G, the coarse words of the pixels whose
  // predecessor lies in another tile, and the tile-border pixels' path
  // ends), all device.
  struct BackJob
  {
    ImgDesc              d;
    const unsigned char* gd;
    long long            gd_pitch;
    unsigned char*       gi;
    long long            gi_pitch;
    unsigned char*       gm; // NULL: no dL/dmask
    long long            gm_pitch;
    double*              G;  // subtree sums
    unsigned long long*  cw; // per pixel; used at exits: their coarse words
    uint32_t*            xp; // per pixel; used at tile borders: where the in-tile path ends
  };

  // A call: one image (d0) or a batch described by a device table.
  struct Batch
  {
    ImgDesc         d0;
    const ImgDesc*  table;    // nimg entries (device), used when nimg > 1
    const uint32_t* tile_img; // global tile id -> image index (device), nimg > 1
    int             nimg;
    int             ntiles;   // total tiles of the call
  };

  // Arithmetic of the relaxation.  cur = (D(q) + e) + g with e = l_eucl*ld
  // and g = l_grad*|I(p)-I(q)| (dt/src/generalised_distance_transform.cu:45).
  // When lambda is 0 or 1 one of the two addends is an exact zero and the
  // kernels use a single add (fl(x + 0) == x for x >= +0): the cost arrays
  // then hold e (lambda = 0) or g (lambda = 1).
  struct KParams
  {
    int   img_kind;
    int   mask_kind;
    float l_grad;
    float e_axis; // l_eucl * 1
    float e_diag; // l_eucl * sqrtf(2)
    float v;      // non-seed value (uint8 mask) or nu (float mask)
    int   zero_lambda;
    int   max_sweeps; // sweep cap per tile visit (0 = until stable)
    int   rh;         // region rows (64 / S)
    int   c8;         // 8-connectivity (layout of the cached cost records)
    int   nw;         // stacked waves per tile (task schedule, rh == 32): a tile
                      // owns nw * (rh - 2) rows
  };

  // Device work queue of dirty tiles (persistent kernel; the protocol is in
  // dtk_queue.hpp).
  //   slots[CAP]     : main ring, (seq << 32) | tile; seq = 2*ticket while
  //                    free for that ticket, 2*ticket + 1 once filled.
  //   bslots[B][CAP] : one such later ring per priority bucket (bucket orders).
  //   state[t]       : TileState of tile t.
  //   ctrl           : counters, each hot word on its own 128-byte line.
  enum CtrlWord : int
  {
    CTRL_HEAD    = 0,  // FIFO: tickets claimed by poppers
    CTRL_TP      = 32, // FIFO, 64-bit: low 32 bits = tail (tickets issued to pushers),
                       //              high 32 bits = pending (tiles queued or running)
    CTRL_DONE    = 64, // set once pending reaches 0 (idle workers poll it)
    CTRL_ERROR   = 96,
    CTRL_VISITS  = 112,
    CTRL_SWEEPS  = 113,
    CTRL_CHANGED = 114,
    CTRL_KEPT    = 115, // task: visits that followed a visit of the same tile at once (end_visit's keep)
    CTRL_T_WAIT  = 116, // summed s_memtime cycles / 1024 of the task workers:
    CTRL_T_VISIT = 117, //   waiting for a tile, solving tiles,
    CTRL_T_PUB   = 118, //   publishing (release, wake-ups, re-queue)
    CTRL_T_SETUP = 119, //   inside visits: loads + edge costs,
    CTRL_T_SWEEP = 120, //   sweeps,
    CTRL_T_STORE = 121, //   border check + write-back,
    CTRL_T_REL   = 122, //   inside publishing: the release fence,
    CTRL_T_ACQ   = 123, //   inside waiting: claiming a popped tile + acquire fence
    CTRL_DRAINS  = 124, // bucket orders: drains that moved tiles,
    CTRL_MOVED   = 125, //   tiles they moved,
    CTRL_DEFER   = 126, //   tiles deferred into later rings,
    CTRL_T_DRAIN = 127, //   summed cycles / 1024 spent draining (inside waiting)
    CTRL_FRONT   = 128, // bucket orders: the front bucket (tiles beyond front + window are deferred)
    CTRL_DRAIN   = 160, // bucket orders: 1 while a worker drains the later rings
    CTRL_LHEAD   = 192, // bucket orders: per later ring, tickets drained (64 words)
    CTRL_LTAIL   = 256, // bucket orders: per later ring, tickets issued to pushers (64 words)
    CTRL_WORDS   = 320,
  };

  // Priority buckets of the bucket orders (one later ring each; one wave
  // lane per bucket when a drainer scans them).
  constexpr uint32_t kBuckets = 64;

  // How the queue orders dirty tiles (the result never depends on it).
  enum QueueOrder : int
  {
    ORDER_FIFO = 0, // first queued first solved
    ORDER_KEY  = 1, // deferral by the smallest improved value a wake carries
    ORDER_SEED = 2, // deferral by a static priority: tile distance to the seeds
  };

  // Tile states (one word per tile, changed only by read-modify-writes):
  //   0                  idle
  //   TILE_DIRTY         queued (in the ring, not yet popped)
  //   TILE_RUNNING       being solved by one wave
  //   RUNNING|DIRTY      being solved, woken again meanwhile (re-queued at the end,
  //                      or visited again at once by the same worker)
  // A waker does fetch_or(TILE_DIRTY); the one that sees 0 inserts the tile.
  enum TileState : uint32_t
  {
    TILE_RUNNING = 2u,
    TILE_DIRTY   = 4u,
  };

  struct Queue
  {
    unsigned long long* slots;  // main ring (kSlotStride words apart), or NULL
    unsigned long long* bslots; // later rings (kBuckets x CAP words), or NULL
    uint32_t*           state;
    uint32_t*           ctrl;
    float*              tmin; // per tile min / max of the initial map (seed selection)
    float*              tmax;
    uint32_t*           tkey;  // ORDER_KEY: per tile the smallest key of the wakes since its
                               // last pop (float bits); ORDER_SEED: distance scratch
    uint8_t*            tprio; // ORDER_SEED: per tile static priority (bucket)
    int*                tile_visits; // optional: visits per tile (instrumentation)
    uint32_t*           mailbox;     // pinned host word: error bits of the call, or NULL
    uint32_t            cap_mask;    // CAP - 1 (CAP a power of two)
    uint32_t            cap_log2;
    int                 ntiles;
    int                 order;  // QueueOrder
    int                 window; // bucket orders: buckets above the front still queued in the main ring
    int                 pad_;
  };

  enum ErrorBits : uint32_t
  {
    ERR_MASK_RANGE = 1u, // soft mask value outside [0, 1] or NaN
    ERR_WATCHDOG   = 2u, // a device-side wait exceeded kSpinLimit polls
    ERR_QUEUE      = 4u, // a tile id or bucket read back from the queue was out of range
  };

  // Upper bound on the polls of any device-side wait (queue slots, raster
  // progress flags).  Legitimate waits end long before; an expired wait
  // flags ERR_WATCHDOG and the kernel finishes instead of hanging.
  constexpr uint32_t kSpinLimit = 1u << 26;

  // Polls of a wait for a queue slot (a ring slot a pusher fills right after
  // reserving its ticket, or frees right after reading it): such a wait
  // ends within microseconds, and a poll is a memory round trip of about a
  // microsecond, so this limit stops a faulty wait within a few seconds.
  constexpr uint32_t kSlotWaitLimit = 1u << 20;

  // Ring slots are kSlotStride 8-byte words apart, so the slot an idle wave
  // polls shares its cache line with no other slot.
#ifndef DTK_SLOT_STRIDE
#define DTK_SLOT_STRIDE 16
#endif
  constexpr int kSlotStride = DTK_SLOT_STRIDE;

// Tile state words are kStateStride 4-byte words apart.
#ifndef DTK_STATE_STRIDE
#define DTK_STATE_STRIDE 1
#endif
  constexpr int kStateStride = DTK_STATE_STRIDE;

// Visits in a row that one worker may add to a tile it keeps (a tile woken
// during its own visit is visited again at once by the same worker, with its
// edge costs still in registers).  The bound keeps the ticket the worker
// claimed meanwhile from waiting on a long run of such visits.  4096^2 RGB8,
// ms at lambda 0 / 0.5 / 1: 0.57 / 1.34 / 1.41 with 4, 0.59 / 1.36 / 1.45
// with 16, 0.59 / 1.38 / 1.42 with 64 (1.55 / 1.61 at lambda > 0 without
// keeping).
#ifndef DTK_MAX_KEEPS
#define DTK_MAX_KEEPS 4
#endif
  constexpr uint32_t kMaxKeeps = DTK_MAX_KEEPS;

#ifndef DTK_POLL_NAPS
#define DTK_POLL_NAPS 2
#endif
  constexpr uint32_t kPollNaps = DTK_POLL_NAPS;

  // Longest wait of an idle persistent worker for a queued tile, in ticks of
  // the 100 MHz s_memrealtime clock (30 s).
  constexpr uint64_t kIdleLimitTicks = 30ull * 100000000ull;
} // namespace dtk
