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
  // its scratch (subtree sums G; per pixel a word holding the count of
  // children still to arrive (high 32 bits) and the float sum of their
  // subtree sums so far (low 32 bits)), all device.
  struct BackJob
  {
    ImgDesc              d;
    const unsigned char* gd;
    long long            gd_pitch;
    unsigned char*       gi;
    long long            gi_pitch;
    unsigned char*       gm; // NULL: no dL/dmask
    long long            gm_pitch;
    float*               G;
    unsigned long long*  cw;
    unsigned char*       meta; // per pixel: child count << 4 | predecessor code
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

  // Device work queue of dirty tiles (persistent kernel).
  //   slots[CAP]: (seq << 32) | tile; seq = 2*ticket while free for that
  //               ticket, 2*ticket + 1 once filled.
  //   state[t]  : TileState of tile t.  A tile is in the ring at most once
  //               and solved by at most one wave at a time; a wake-up while
  //               it runs sets TILE_DIRTY and the running wave re-queues it
  //               when it finishes.
  //   ctrl      : counters, each on its own 64-byte line.
  enum CtrlWord : int
  {
    CTRL_HEAD    = 0,
    CTRL_TP      = 16, // 64-bit: low 32 bits = tail (tickets issued to pushers),
                       //         high 32 bits = pending (tiles queued or running)
    CTRL_DONE    = 32, // set once pending reaches 0 (idle workers poll it, not CTRL_TP)
    CTRL_ERROR   = 48,
    CTRL_VISITS  = 64,
    CTRL_SWEEPS  = 80,
    CTRL_CHANGED = 96,
    CTRL_T_WAIT  = 100, // summed s_memtime cycles / 1024 of the task workers:
    CTRL_T_VISIT = 101, //   waiting for a tile, solving tiles,
    CTRL_T_PUB   = 102, //   publishing (release, wake-ups, re-queue)
    CTRL_T_SETUP = 103, //   inside visits: loads + edge costs,
    CTRL_T_SWEEP = 104, //   sweeps,
    CTRL_T_STORE = 105, //   border check + write-back,
    CTRL_T_REL   = 106, //   inside publishing: the release fence,
    CTRL_T_ACQ   = 107, //   inside waiting: claiming a popped tile + acquire fence
    CTRL_WORDS   = 128,
  };

  // Tile states (one word per tile, changed only by read-modify-writes so
  // every waker's D stores are in the release sequence the next visit
  // acquires from):
  //   0                  idle
  //   TILE_DIRTY         queued (in the ring, not yet popped)
  //   TILE_RUNNING       being solved by one wave
  //   RUNNING|DIRTY      being solved, woken again meanwhile (re-queued at the end)
  // A waker does fetch_or(TILE_DIRTY); the one that sees 0 inserts the tile.
  enum TileState : uint32_t
  {
    TILE_RUNNING = 2u,
    TILE_DIRTY   = 4u,
  };

  struct Queue
  {
    unsigned long long* slots;
    uint32_t*           state;
    uint32_t*           ctrl;
    float*              tmin; // per tile min / max of the initial map (seed selection)
    float*              tmax;
    int*                tile_visits; // optional: visits per tile (instrumentation)
    uint32_t            cap_mask;
    int                 pad_;
  };

  enum ErrorBits : uint32_t
  {
    ERR_MASK_RANGE = 1u, // soft mask value outside [0, 1] or NaN
    ERR_WATCHDOG   = 2u, // a device-side wait exceeded kSpinLimit polls
  };

  // Upper bound on the polls of any device-side wait (queue slots, raster
  // progress flags).  Legitimate waits end long before; an expired wait
  // flags ERR_WATCHDOG and the kernel finishes instead of hanging.
  constexpr uint32_t kSpinLimit = 1u << 26;

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

#ifndef DTK_POLL_NAPS
#define DTK_POLL_NAPS 2
#endif
  constexpr uint32_t kPollNaps = DTK_POLL_NAPS;

  // Longest wait of an idle persistent worker for a queued tile, in ticks of
  // the 100 MHz s_memrealtime clock (30 s).
  constexpr uint64_t kIdleLimitTicks = 30ull * 100000000ull;
} // namespace dtk
