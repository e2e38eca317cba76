// Host side of libdtk: the C ABI declared in include/dtk.h.
//
// Error model: every entry point returns a dtk_status; a message for the
// calling thread is kept for dtk_last_error().  Nothing throws across the
// ABI (the C++ wrappers in include/dt/ turn codes into the reference's
// exceptions, dt/src/generalised_distance_transform.cu:136-137).
//
// Memory: each device has one cached workspace (grown on demand with
// hipMalloc, freed only when it must grow) for the queue, tile statistics,
// batch tables and adjoint scratch.  Calls on one device share it, so calls
// that overlap on one device must be ordered on one stream (the reference
// is not re-entrant either: it rewrites a __constant__ symbol per call,
// dt/src/generalised_distance_transform.cu:112).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dtk.h"
#include "dtk_launch.hpp"

using namespace dtk;

namespace
{
  thread_local std::string g_err;

  dtk_status fail(dtk_status s, const std::string& msg)
  {
    g_err = msg;
    return s;
  }

#define DTK_HIP(expr, what)                                                                                           \
  do                                                                                                                  \
  {                                                                                                                   \
    hipError_t e_ = (expr);                                                                                           \
    if (e_ != hipSuccess)                                                                                             \
      return fail(e_ == hipErrorOutOfMemory ? DTK_ERR_OUT_OF_MEMORY : DTK_ERR_HIP,                                    \
                  std::string(what) + ": " + hipGetErrorString(e_));                                                  \
  } while (0)

  constexpr int kMaxDevices = 64;

  // Pinned host buffers for host-to-device uploads of call tables, used in
  // turn; each is reused only once the copy that read it has completed (its
  // event), so an upload never makes the caller wait for the device.
  constexpr int kStages = 8;
  struct Stage
  {
    void*      h   = nullptr;
    size_t     cap = 0;
    hipEvent_t ev  = nullptr;
    bool       pending = false;
  };

  struct Workspace
  {
    std::mutex  mu;
    void*       dbuf = nullptr;
    size_t      dcap = 0;
    void*       hbuf = nullptr; // pinned host buffer for synchronous read-backs
    size_t      hcap = 0;
    int         cu   = 0;
    // Stream order of the calls that use dbuf: the event of the last such
    // call and its stream.  A call on another stream first waits for it.
    hipEvent_t  done = nullptr;
    hipStream_t last = nullptr;
    bool        used = false;
    // Error mailbox in pinned host memory: the persistent kernels of an
    // asynchronous call store their error bits here; the next call reads it.
    uint32_t*   mailbox = nullptr;
    Stage       stage[kStages];
    int         next_stage = 0;
  };

  Workspace g_ws[kMaxDevices];

  // Kernel timing for benchmarks: while active, every task-schedule call
  // records three HIP events on its stream (before the reset/init/seed
  // kernels, before the persistent solve, after it).  No synchronisation
  // happens until dtk_profile_end.  One profiler per process.
  struct Profiler
  {
    std::mutex             mu;
    bool                   on = false;
    std::vector<hipEvent_t> ev;
    size_t                 used = 0;
  };
  Profiler g_prof;

  void prof_mark(hipStream_t s)
  {
    std::lock_guard<std::mutex> lock(g_prof.mu);
    if (!g_prof.on)
      return;
    if (g_prof.used == g_prof.ev.size())
    {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess)
        return;
      g_prof.ev.push_back(e);
    }
    hipEventRecord(g_prof.ev[g_prof.used++], s);
  }

  int elem_size(int dtype)
  {
    switch (dtype)
    {
    case DTK_U8:
      return 1;
    case DTK_RGB8:
      return 3;
    case DTK_F32:
      return 4;
    default:
      return 0;
    }
  }

  size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

  uint32_t next_pow2(uint32_t v)
  {
    uint32_t p = 1;
    while (p < v)
      p <<= 1;
    return p;
  }

  dtk_status current_ws(Workspace** out, int* dev_out)
  {
    int dev = 0;
    DTK_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
      return fail(DTK_ERR_UNSUPPORTED, "device index out of range");
    Workspace& w = g_ws[dev];
    if (w.cu == 0)
    {
      std::lock_guard<std::mutex> lock(w.mu);
      int cu = 0;
      DTK_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
      if (!w.done)
        DTK_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming), "hipEventCreate");
      if (!w.mailbox)
      {
        void* m = nullptr;
        DTK_HIP(hipHostMalloc(&m, 256, hipHostMallocDefault), "hipHostMalloc(mailbox)");
        w.mailbox                           = static_cast<uint32_t*>(m);
        *static_cast<volatile uint32_t*>(m) = 0u;
      }
      w.cu = cu > 0 ? cu : 1;
    }
    *out = &w;
    if (dev_out)
      *dev_out = dev;
    return DTK_OK;
  }

  dtk_status ws_reserve(Workspace& w, size_t dbytes, size_t hbytes)
  {
    if (dbytes > w.dcap)
    {
      if (w.dbuf)
      {
        DTK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
        DTK_HIP(hipFree(w.dbuf), "hipFree");
        w.dbuf = nullptr;
        w.dcap = 0;
      }
      const size_t cap = align_up(dbytes + dbytes / 4, 1 << 20);
      DTK_HIP(hipMalloc(&w.dbuf, cap), "hipMalloc(workspace)");
      w.dcap = cap;
    }
    if (hbytes > w.hcap)
    {
      if (w.hbuf)
      {
        DTK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
        DTK_HIP(hipHostFree(w.hbuf), "hipHostFree");
        w.hbuf = nullptr;
        w.hcap = 0;
      }
      const size_t cap = align_up(hbytes, 1 << 16);
      DTK_HIP(hipHostMalloc(&w.hbuf, cap, hipHostMallocDefault), "hipHostMalloc(staging)");
      w.hcap = cap;
    }
    return DTK_OK;
  }

  // Device error bits -> status and message.
  dtk_status error_from_bits(uint32_t bits, const char* where)
  {
    if (bits & ERR_MASK_RANGE)
      return fail(DTK_ERR_INVALID_ARGUMENT, std::string("soft mask values must be in [0, 1]") + where);
    if (bits & ERR_QUEUE)
      return fail(DTK_ERR_HIP, std::string("work queue fault: a tile id read back from the queue was out of range") + where);
    if (bits & ERR_WATCHDOG)
      return fail(DTK_ERR_HIP, std::string("device watchdog expired in the work queue") + where);
    return DTK_OK;
  }

  // Errors that the kernels of earlier asynchronous calls on this device
  // left in its mailbox.  Reading it is a plain host read, so a call with no
  // error pending never waits; when one is there, the device is synchronised
  // first (no kernel writes the mailbox any more), then the error is taken
  // and reported once.
  dtk_status take_async_error(Workspace& w)
  {
    volatile uint32_t* mb = w.mailbox;
    if (!mb || *mb == 0u)
      return DTK_OK;
    DTK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const uint32_t bits = *mb;
    *mb                 = 0u;
    return error_from_bits(bits, " (an earlier asynchronous call on this device)");
  }

  // A call's use of the device workspace: holds its lock, orders the call
  // after the previous user of the workspace when that ran on another
  // stream (the stream waits for its event; on one stream the order is
  // already there), reports pending asynchronous errors, and at scope exit
  // records the call's own event.
  class WsUse
  {
  public:
    WsUse(Workspace& w, hipStream_t s)
      : w_(w)
      , s_(s)
      , lock_(w.mu)
    {
    }
    dtk_status begin()
    {
      dtk_status st;
      if ((st = take_async_error(w_)) != DTK_OK)
        return st;
      if (w_.used && w_.last != s_)
        DTK_HIP(hipStreamWaitEvent(s_, w_.done, 0), "hipStreamWaitEvent");
      entered_ = true;
      return DTK_OK;
    }
    ~WsUse()
    {
      if (entered_ && hipEventRecord(w_.done, s_) == hipSuccess)
      {
        w_.last = s_;
        w_.used = true;
      }
    }

  private:
    Workspace&                  w_;
    hipStream_t                 s_;
    std::lock_guard<std::mutex> lock_;
    bool                        entered_ = false;
  };

  // Pinned staging buffer of at least `bytes` for an upload on stream s:
  // the next buffer of the ring, once the last copy that read it is done.
  // Call stage_done after enqueueing the copy.  Caller holds the workspace
  // lock.
  dtk_status stage_get(Workspace& w, size_t bytes, Stage** out)
  {
    Stage& st = w.stage[w.next_stage];
    w.next_stage = (w.next_stage + 1) % kStages;
    if (st.pending)
    {
      DTK_HIP(hipEventSynchronize(st.ev), "hipEventSynchronize");
      st.pending = false;
    }
    if (!st.ev)
      DTK_HIP(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming), "hipEventCreate");
    if (bytes > st.cap)
    {
      if (st.h)
        DTK_HIP(hipHostFree(st.h), "hipHostFree");
      st.h   = nullptr;
      st.cap = 0;
      const size_t cap = align_up(bytes, 1 << 16);
      DTK_HIP(hipHostMalloc(&st.h, cap, hipHostMallocDefault), "hipHostMalloc(staging)");
      st.cap = cap;
    }
    *out = &st;
    return DTK_OK;
  }
  dtk_status stage_done(Stage* st, hipStream_t s)
  {
    DTK_HIP(hipEventRecord(st->ev, s), "hipEventRecord");
    st->pending = true;
    return DTK_OK;
  }

  // Carves aligned pieces out of a base pointer.
  struct Carver
  {
    char*  base;
    size_t off = 0;
    explicit Carver(void* b)
      : base(static_cast<char*>(b))
    {
    }
    template <typename T>
    T* take(size_t count)
    {
      off      = align_up(off, 256);
      T* p     = base ? reinterpret_cast<T*>(base + off) : nullptr;
      off += count * sizeof(T);
      return p;
    }
  };

  dtk_status check_image(const dtk_image* im, const char* name, int need_dtype_mask)
  {
    if (!im)
      return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    const int es = elem_size(im->dtype);
    if (es == 0 || !((need_dtype_mask >> im->dtype) & 1))
      return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + ": unsupported dtype");
    if (im->width < 0 || im->height < 0)
      return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + ": negative size");
    if ((long long)im->width * im->height > 0)
    {
      if (!im->data)
        return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + ": NULL data");
      if (im->pitch < (long long)im->width * es)
        return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + ": pitch smaller than a row");
      if (im->dtype == DTK_F32 && ((uintptr_t)im->data % 4 != 0 || im->pitch % 4 != 0))
        return fail(DTK_ERR_INVALID_ARGUMENT, std::string(name) + ": float rows must be 4-byte aligned");
    }
    return DTK_OK;
  }

  dtk_status check_params(const dtk_params* p)
  {
    if (!p)
      return fail(DTK_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(p->lambda >= 0.f && p->lambda <= 1.f))
      return fail(DTK_ERR_INVALID_ARGUMENT, "lambda must be in [0, 1]");
    if (!(p->v >= 0.f) || std::isinf(p->v))
      return fail(DTK_ERR_INVALID_ARGUMENT, "v must be finite and >= 0");
    if (p->connectivity != DTK_C4 && p->connectivity != DTK_C8)
      return fail(DTK_ERR_INVALID_ARGUMENT, "connectivity must be DTK_C4 or DTK_C8");
    if (p->schedule < DTK_SCHED_RASTER || p->schedule > DTK_SCHED_TASK)
      return fail(DTK_ERR_INVALID_ARGUMENT, "unknown schedule");
    if (p->tile_rows != 0 && p->tile_rows != 32 && p->tile_rows != 16)
      return fail(DTK_ERR_INVALID_ARGUMENT, "tile_rows must be 0, 16 or 32");
    if (!(p->tile_waves == 0 || p->tile_waves == 1 || p->tile_waves == 2 || p->tile_waves == 4))
      return fail(DTK_ERR_INVALID_ARGUMENT, "tile_waves must be 0, 1, 2 or 4");
    if (p->max_iterations < 0 || p->max_sweeps < 0 || p->waves_per_cu < 0)
      return fail(DTK_ERR_INVALID_ARGUMENT, "negative iteration / sweep / wave count");
    if (p->max_iterations > 0 && p->schedule != DTK_SCHED_RASTER)
      return fail(DTK_ERR_INVALID_ARGUMENT, "a fixed iteration count needs the raster schedule");
    if (p->queue_order < DTK_ORDER_DEFAULT || p->queue_order > DTK_ORDER_SEED)
      return fail(DTK_ERR_INVALID_ARGUMENT, "unknown queue order");
    return DTK_OK;
  }

  // Queue order of a task-schedule call over nimg images.  DEFAULT: KEY for
  // batches (256 uneven uint8 images at lambda 0.8: 11.7 ms against 28.8 ms
  // FIFO in one launch), FIFO for one image (4096^2 RGB8 at lambda 0.5:
  // 1.61 ms against 1.83 ms KEY; the 8192^2 spiral maze: 1.52 s against
  // 2.11 s).
  int queue_order(const dtk_params* p, int nimg)
  {
    switch (p->queue_order)
    {
    case DTK_ORDER_FIFO:
      return ORDER_FIFO;
    case DTK_ORDER_KEY:
      return ORDER_KEY;
    case DTK_ORDER_SEED:
      return ORDER_SEED;
    default:
      return nimg > 1 ? ORDER_KEY : ORDER_FIFO;
    }
  }

  // lambda 0 without a sweep cap: the tile kernels use constant chamfer
  // costs and leave the pixels outside the image unmasked.  Those pixels
  // restart at +inf on every visit and settle during it, so a visit must
  // be free to run until it is quiet: under a sweep cap an edge tile could
  // never finish a visit (and would re-queue itself forever), so capped
  // calls keep the masked per-pixel costs.
  bool const_costs(const dtk_params* p) { return p->lambda == 0.f && p->max_sweeps == 0; }

  // Stacked waves per tile, for the task and blocks schedules alike: four
  // 16-row regions (62 x 56 owned pixels), four 32-row regions with constant
  // costs (62 x 120) and two 32-row regions when 32 rows are asked for
  // otherwise (62 x 60), unless told otherwise.  The raster schedule has no
  // tiles.
  int tile_waves(const dtk_params* p, int rh)
  {
    if (p->schedule == DTK_SCHED_RASTER)
      return 1;
    if (p->tile_waves != 0)
      return p->tile_waves;
    return rh == 32 && !const_costs(p) ? 2 : 4;
  }

  KParams make_kparams(const dtk_params* p, int img_dtype, int mask_dtype, int rh)
  {
    KParams k{};
    k.img_kind    = img_dtype == DTK_U8 ? IMG_U8 : (img_dtype == DTK_RGB8 ? IMG_RGB8 : IMG_F32);
    k.mask_kind   = mask_dtype == DTK_U8 ? MASK_U8 : MASK_F32;
    k.l_grad      = p->lambda;
    const float le = 1.f - p->lambda; // dt/src/generalised_distance_transform.cu:108-111
    k.e_axis      = le * 1.f;
    k.e_diag      = le * std::sqrt(2.f);
    k.v           = p->v;
    k.zero_lambda = p->lambda == 0.f;
    k.max_sweeps  = p->max_sweeps;
    k.rh          = rh;
    k.c8          = p->connectivity == DTK_C8;
    k.nw          = tile_waves(p, rh);
    return k;
  }

  // Edge-cost mode of the tile kernels: 2 = constant chamfer costs, 0 = one
  // add (lambda 1, or lambda 0 with masked costs), 1 = two adds.
  int cost_mode(const dtk_params* p)
  {
    if (const_costs(p))
      return 2;
    return p->lambda == 0.f || p->lambda == 1.f ? 0 : 1;
  }

  // Region rows per wave: 16-row waves (the per-pixel cost arrays leave
  // room for two waves per SIMD), and 32-row waves with constant costs
  // (105 VGPRs).  Measured on 4096^2 RGB8 (task) and on 8.5-9.6 Mpx u8
  // images with one seed (task and blocks): 16 x 4 is fastest at lambda 0.5
  // and 1, 32 x 4 at lambda 0; single-wave blocks tiles (62 x 30) took
  // 1.5-2.1x as long.
  int region_rows(const dtk_params* p)
  {
    if (p->tile_rows != 0)
      return p->tile_rows;
    if (p->schedule == DTK_SCHED_RASTER)
      return 32;
    return const_costs(p) ? 32 : kDefaultTaskRows;
  }

  ImgDesc make_desc(const dtk_image* img, const dtk_image* mask, dtk_image* dist, int rh, int nw)
  {
    ImgDesc d{};
    d.img        = static_cast<const unsigned char*>(img->data);
    d.img_pitch  = img->pitch;
    d.mask       = static_cast<const unsigned char*>(mask->data);
    d.mask_pitch = mask->pitch;
    d.D          = static_cast<float*>(dist->data);
    d.D_pitch    = dist->pitch;
    d.pred       = nullptr;
    d.pred_pitch = 0;
    d.W          = img->width;
    d.H          = img->height;
    d.tiles_x    = (d.W + kTileW - 1) / kTileW;
    d.tiles_y    = (d.H + nw * (rh - 2) - 1) / (nw * (rh - 2));
    d.tile0      = 0;
    return d;
  }

  dtk_status check_triplet(const dtk_image* img, const dtk_image* mask, const dtk_image* dist)
  {
    dtk_status s;
    if ((s = check_image(img, "image", (1 << DTK_U8) | (1 << DTK_RGB8) | (1 << DTK_F32))) != DTK_OK)
      return s;
    if ((s = check_image(mask, "mask", (1 << DTK_U8) | (1 << DTK_F32))) != DTK_OK)
      return s;
    if ((s = check_image(dist, "dist", 1 << DTK_F32)) != DTK_OK)
      return s;
    if (img->width != mask->width || img->height != mask->height || img->width != dist->width ||
        img->height != dist->height)
      return fail(DTK_ERR_INVALID_ARGUMENT, "image, mask and dist sizes differ");
    return DTK_OK;
  }

  // Persistent workers (workgroups of nw waves) of a task-schedule call on a
  // device with `cu` compute units.
  int task_workers_cu(int cu, const dtk_params* p, int S, bool c8, int cm, int ntiles)
  {
    const int nw     = tile_waves(p, 64 / S);
    int       per_cu = task_occupancy(S, c8, cm, nw);
    if (per_cu <= 0)
      per_cu = 1;
    if (p->waves_per_cu > 0 && p->waves_per_cu / nw < per_cu)
      per_cu = std::max(1, p->waves_per_cu / nw);
    // Four-wave workers: two per CU by default.  More (the lambda-0 kernels
    // fit three or four) only add redundant visits: 4096^2 RGB8 at lambda 0
    // takes 0.56 ms with two per CU and 0.60 ms with three.
    if (p->waves_per_cu == 0 && nw == 4 && per_cu > 2)
      per_cu = 2;
    long long n = (long long)per_cu * cu;
    if (n > ntiles)
      n = ntiles;
    return (int)(n > 0 ? n : 1);
  }
  int task_workers(const Workspace& w, const dtk_params* p, int S, bool c8, int cm, int ntiles)
  {
    return task_workers_cu(w.cu, p, S, c8, cm, ntiles);
  }

  // Task-schedule buffers of one call, carved from a device block.
  struct TaskBufs
  {
    Queue     q{};
    ImgDesc*  table    = nullptr;
    uint32_t* tile_img = nullptr;
  };

  // Buckets above the front whose wakes still go to the main ring (bucket
  // orders; two buckets per octave of the distance for ORDER_KEY).
#ifndef DTK_QUEUE_WINDOW
#define DTK_QUEUE_WINDOW 0
#endif
  constexpr int kQueueWindow = DTK_QUEUE_WINDOW;

  // The main ring (cap slots, kSlotStride words apart) and, for a bucket
  // order, the later rings (kBuckets x cap slots) with per-tile keys and
  // priorities.  cap is a power of two.
  size_t task_layout(void* base, int ntiles, uint32_t cap, int nimg, int order, TaskBufs* out)
  {
    Carver c(base);
    TaskBufs b;
    b.q.ctrl     = c.take<uint32_t>(CTRL_WORDS);
    b.q.tmin     = c.take<float>(ntiles);
    b.q.tmax     = c.take<float>(ntiles);
    b.q.state    = c.take<uint32_t>((size_t)ntiles * kStateStride);
    b.q.slots    = c.take<unsigned long long>((size_t)cap * kSlotStride);
    b.q.window   = order == ORDER_FIFO ? (int)kBuckets : kQueueWindow;
    if (order != ORDER_FIFO)
    {
      b.q.bslots = c.take<unsigned long long>((size_t)kBuckets * cap);
      b.q.tkey   = c.take<uint32_t>(ntiles);
      b.q.tprio  = c.take<uint8_t>(ntiles);
    }
    b.q.cap_mask = cap - 1;
    b.q.cap_log2 = 0;
    while ((1u << b.q.cap_log2) < cap)
      ++b.q.cap_log2;
    b.q.ntiles = ntiles;
    b.q.order  = order;
    if (nimg > 1)
    {
      b.table    = c.take<ImgDesc>(nimg);
      b.tile_img = c.take<uint32_t>(ntiles);
    }
    if (out)
      *out = b;
    return c.off + 256;
  }

  // Cached edge-cost records of the tile schedules: one pitched plane per
  // image (16 B/px for C8, 8 B/px for C4), written by k_init and read by
  // every tile visit.  With base == NULL only the size is computed.
  // Edge costs from a per-call plane (opt-in), or from the image on every
  // visit: measured on 4096^2 images the per-visit evaluation is 4-7% faster
  // for RGB8, 0-4% for uint8 and 18-24% for float32 images, and it saves the
  // plane (16 B/px) and the pass that writes it.
  bool use_cost_cache(const dtk_params* p) { return p->lambda != 0.f && p->cost_plane; }

  size_t place_costs(ImgDesc* descs, int n, bool c8, char* base)
  {
    size_t off = 0;
    for (int i = 0; i < n; ++i)
    {
      ImgDesc& d    = descs[i];
      d.cost_pitch  = (long long)align_up((size_t)d.W * (c8 ? 16 : 8), 256);
      d.cost        = base ? reinterpret_cast<unsigned char*>(base + off) : nullptr;
      off += (size_t)d.cost_pitch * (size_t)(d.H > 0 ? d.H : 0);
      off = align_up(off, 256);
    }
    return off;
  }

  dtk_status fill_stats(dtk_stats* st, const uint32_t* ctrl_host, int rounds, int tiles, int rh, int workers)
  {
    if (!st)
      return DTK_OK;
    st->rounds    = rounds;
    st->visits    = ctrl_host ? ctrl_host[CTRL_VISITS] : 0;
    st->sweeps    = ctrl_host ? ctrl_host[CTRL_SWEEPS] : 0;
    st->tiles     = tiles;
    st->tile_rows = rh;
    st->workers   = workers;
    st->cycles_wait  = ctrl_host ? ctrl_host[CTRL_T_WAIT] : 0;
    st->cycles_visit = ctrl_host ? ctrl_host[CTRL_T_VISIT] : 0;
    st->cycles_pub   = ctrl_host ? ctrl_host[CTRL_T_PUB] : 0;
    st->cycles_setup = ctrl_host ? ctrl_host[CTRL_T_SETUP] : 0;
    st->cycles_sweep = ctrl_host ? ctrl_host[CTRL_T_SWEEP] : 0;
    st->cycles_store = ctrl_host ? ctrl_host[CTRL_T_STORE] : 0;
    st->cycles_release = ctrl_host ? ctrl_host[CTRL_T_REL] : 0;
    st->cycles_acquire = ctrl_host ? ctrl_host[CTRL_T_ACQ] : 0;
    st->drains         = ctrl_host ? ctrl_host[CTRL_DRAINS] : 0;
    st->moved          = ctrl_host ? ctrl_host[CTRL_MOVED] : 0;
    st->deferred       = ctrl_host ? ctrl_host[CTRL_DEFER] : 0;
    st->cycles_drain   = ctrl_host ? ctrl_host[CTRL_T_DRAIN] : 0;
    st->kept           = ctrl_host ? ctrl_host[CTRL_KEPT] : 0;
    return DTK_OK;
  }

  // Builds the batch descriptors (host) for n images.
  dtk_status build_batch(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists, const dtk_params* p,
                         int rh, std::vector<ImgDesc>& descs, std::vector<uint32_t>& tile_img, int* ntiles)
  {
    descs.resize(n);
    long long total = 0;
    for (int i = 0; i < n; ++i)
    {
      dtk_status s = check_triplet(&imgs[i], &masks[i], &dists[i]);
      if (s != DTK_OK)
        return fail(s, "image " + std::to_string(i) + ": " + g_err);
      if (imgs[i].dtype != imgs[0].dtype || masks[i].dtype != masks[0].dtype)
        return fail(DTK_ERR_INVALID_ARGUMENT, "all images of a batch must share image and mask dtypes");
      descs[i]       = make_desc(&imgs[i], &masks[i], &dists[i], rh, tile_waves(p, rh));
      descs[i].tile0 = (int)total;
      if (descs[i].W > 0 && descs[i].H > 0)
        total += (long long)descs[i].tiles_x * descs[i].tiles_y;
      else
        descs[i].tiles_x = descs[i].tiles_y = 0;
      if (total > (1ll << 30))
        return fail(DTK_ERR_UNSUPPORTED, "too many tiles in one batch");
    }
    tile_img.resize((size_t)total);
    for (int i = 0; i < n; ++i)
      for (int t = 0; t < descs[i].tiles_x * descs[i].tiles_y; ++t)
        tile_img[(size_t)descs[i].tile0 + t] = (uint32_t)i;
    *ntiles = (int)total;
    (void)p;
    return DTK_OK;
  }

  // Enqueue reset + init + seed + persistent solve for a prepared batch.
  dtk_status enqueue_task(const Batch& b, const KParams& kp, const Queue& q, int S, bool c8, int cm, int workers,
                          hipStream_t s, bool profile = true)
  {
    if (profile)
      prof_mark(s);
    DTK_HIP(launch_reset_ctrl(q.ctrl, s), "launch reset");
    DTK_HIP(launch_init(b, kp, q, s), "launch init");
    DTK_HIP(launch_seed(b, kp, q, s), "launch seed");
    if (profile)
      prof_mark(s);
    DTK_HIP(launch_task(S, c8, cm, b, kp, q, workers, s), "launch task");
    if (profile)
      prof_mark(s);
    return DTK_OK;
  }

  // Synchronous read-back of a task call's counters (calls with stats).  An
  // error is reported here, so the copy the kernel left in the mailbox is
  // taken too (the stream is synchronised: no worker writes it any more).
  dtk_status read_ctrl(Workspace& w, const Queue& q, hipStream_t s, uint32_t** host)
  {
    uint32_t* h = static_cast<uint32_t*>(w.hbuf);
    DTK_HIP(hipMemcpyAsync(h, q.ctrl, CTRL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read counters");
    DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    *host = h;
    if (h[CTRL_ERROR] && w.mailbox)
      *static_cast<volatile uint32_t*>(w.mailbox) = 0u;
    return error_from_bits(h[CTRL_ERROR], "");
  }
} // namespace

// ---- display of a distance map -------------------------------------------
namespace
{
  using namespace dtk;

  // Same-size pair check for the colour entry points.
  dtk_status check_pair(const dtk_image* in, int in_mask, const dtk_image* out, int out_mask)
  {
    dtk_status st;
    if ((st = check_image(in, "in", in_mask)) != DTK_OK || (st = check_image(out, "out", out_mask)) != DTK_OK)
      return st;
    if (in->width != out->width || in->height != out->height)
      return fail(DTK_ERR_INVALID_ARGUMENT, "in and out sizes differ");
    return DTK_OK;
  }

  // Min / max keys of a f32 image in the workspace (two u32, stream-ordered
  // like every other use of the workspace), then `then(keys)` enqueues what
  // reads them, all under the workspace lock (another call may grow, and so
  // reallocate, the workspace once the lock is released).
  template <typename Then>
  dtk_status with_minmax_keys(const dtk_image* in, hipStream_t s, Then&& then)
  {
    dtk_status st;
    Workspace* w = nullptr;
    if ((st = current_ws(&w, nullptr)) != DTK_OK)
      return st;
    WsUse use(*w, s);
    if ((st = use.begin()) != DTK_OK)
      return st;
    if ((st = ws_reserve(*w, 256 + sizeof(uint32_t) * kMinmaxPartialWords, 4096)) != DTK_OK)
      return st;
    uint32_t* keys    = static_cast<uint32_t*>(w->dbuf);
    uint32_t* partial = reinterpret_cast<uint32_t*>(static_cast<char*>(w->dbuf) + 256);
    DTK_HIP(launch_minmax(in->data, in->pitch, in->width, in->height, keys, partial, w->cu, s), "launch minmax");
    return then(keys);
  }
  // The adjoint of n images in one launch of each kernel (see BackJob).
  // With have_gm, grad_masks[i] receives dL/dmask (zeros for u8 masks).
  dtk_status backward_impl(int n, const dtk_image* imgs, const dtk_image* masks, const dtk_image* dists,
                           const dtk_image* preds, const dtk_image* gds, dtk_image* gis, dtk_image* gms, bool have_gm,
                           const dtk_params* p, void* stream)
  {
    dtk_status st;
    if ((st = check_params(p)) != DTK_OK)
      return st;
    size_t dbytes = 0;
    int    maxW = 0, maxH = 0;
    for (int i = 0; i < n; ++i)
    {
      if ((st = check_triplet(&imgs[i], &masks[i], dists ? &dists[i] : nullptr)) != DTK_OK)
        return st;
      if (imgs[i].dtype != DTK_F32)
        return fail(DTK_ERR_UNSUPPORTED, "the adjoint needs a float32 image");
      if (masks[i].dtype != masks[0].dtype)
        return fail(DTK_ERR_INVALID_ARGUMENT, "the masks of a batch must share a dtype");
      if ((st = check_image(&preds[i], "pred", 1 << DTK_U8)) != DTK_OK ||
          (st = check_image(&gds[i], "grad_dist", 1 << DTK_F32)) != DTK_OK ||
          (st = check_image(&gis[i], "grad_img", 1 << DTK_F32)) != DTK_OK)
        return st;
      if (have_gm && (st = check_image(&gms[i], "grad_mask", 1 << DTK_F32)) != DTK_OK)
        return st;
      const int W = imgs[i].width, H = imgs[i].height;
      if (preds[i].width != W || preds[i].height != H || gds[i].width != W || gds[i].height != H ||
          gis[i].width != W || gis[i].height != H || (have_gm && (gms[i].width != W || gms[i].height != H)))
        return fail(DTK_ERR_INVALID_ARGUMENT, "gradient image sizes differ");
      if ((unsigned long long)W * (unsigned long long)H >= 0xFFFFFFFFull)
        return fail(DTK_ERR_UNSUPPORTED, "the adjoint takes images of fewer than 2^32 - 1 pixels");
      if (H > 65535 * 64)
        return fail(DTK_ERR_UNSUPPORTED, "the adjoint takes at most 4194240 rows");
      if (n > 65535)
        return fail(DTK_ERR_UNSUPPORTED, "the adjoint takes at most 65535 images per call");
      maxW = std::max(maxW, W);
      maxH = std::max(maxH, H);
      dbytes += 2 * align_up((size_t)W * H * 8, 256) + align_up((size_t)W * H * 4, 256);
    }
    if (n == 0 || maxW == 0 || maxH == 0)
      return DTK_OK;
    Workspace* w = nullptr;
    if ((st = current_ws(&w, nullptr)) != DTK_OK)
      return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    WsUse       use(*w, s);
    if ((st = use.begin()) != DTK_OK)
      return st;
    const size_t jbytes = align_up((size_t)n * sizeof(BackJob), 256);
    if ((st = ws_reserve(*w, dbytes + jbytes + 1024, 4096)) != DTK_OK)
      return st;
    Stage* stg = nullptr;
    if ((st = stage_get(*w, jbytes, &stg)) != DTK_OK)
      return st;
    Carver      c(w->dbuf);
    BackJob*    djobs = c.take<BackJob>(n);
    BackJob*    hjobs = static_cast<BackJob*>(stg->h);
    const bool  soft  = masks[0].dtype == DTK_F32;
    for (int i = 0; i < n; ++i)
    {
      const size_t npx = (size_t)imgs[i].width * imgs[i].height;
      dtk_image    dc  = dists ? dists[i] : dtk_image{};
      BackJob      j{};
      j.d            = make_desc(&imgs[i], &masks[i], &dc, 32, 1);
      j.d.pred       = static_cast<unsigned char*>(preds[i].data);
      j.d.pred_pitch = preds[i].pitch;
      j.gd           = static_cast<const unsigned char*>(gds[i].data);
      j.gd_pitch     = gds[i].pitch;
      j.gi           = static_cast<unsigned char*>(gis[i].data);
      j.gi_pitch     = gis[i].pitch;
      j.gm           = have_gm && soft ? static_cast<unsigned char*>(gms[i].data) : nullptr;
      j.gm_pitch     = have_gm ? gms[i].pitch : 0;
      j.G            = c.take<double>(npx);
      j.cw           = c.take<unsigned long long>(npx);
      j.xp           = c.take<uint32_t>(npx);
      hjobs[i]       = j;
    }
    DTK_HIP(hipMemcpyAsync(djobs, hjobs, (size_t)n * sizeof(BackJob), hipMemcpyHostToDevice, s), "upload jobs");
    if ((st = stage_done(stg, s)) != DTK_OK)
      return st;
    const KParams kp = make_kparams(p, DTK_F32, masks[0].dtype, 32);
    DTK_HIP(launch_backward(djobs, n, maxW, maxH, kp, s), "launch backward");
    if (have_gm && !soft)
      for (int i = 0; i < n; ++i)
        if ((size_t)imgs[i].width * imgs[i].height > 0)
          DTK_HIP(hipMemset2DAsync(gms[i].data, gms[i].pitch, 0, (size_t)imgs[i].width * 4, imgs[i].height, s),
                  "zero grad_mask");
    return DTK_OK;
  }
} // namespace

extern "C" {

void dtk_params_default(dtk_params* p)
{
  if (!p)
    return;
  std::memset(p, 0, sizeof(*p));
  p->lambda       = 0.f;
  p->v            = 1e10f; // dt/include/dt/generalised_distance_transform.hpp:8
  p->connectivity = DTK_C8;
  p->schedule     = DTK_SCHED_TASK;
}

const char* dtk_last_error(void) { return g_err.c_str(); }

const char* dtk_version(void) { return "dtk 0.1 (gfx950)"; }

int dtk_tile_count(int width, int height, int tile_rows)
{
  const int rh = tile_rows == 0 ? 32 : tile_rows;
  if (width <= 0 || height <= 0 || (rh != 32 && rh != 16))
    return 0;
  return ((width + kTileW - 1) / kTileW) * ((height + rh - 3) / (rh - 2));
}

int dtk_tile_count_for(int width, int height, const dtk_params* p)
{
  if (!p || width <= 0 || height <= 0 || check_params(p) != DTK_OK)
    return 0;
  const int rh = region_rows(p), nw = tile_waves(p, rh);
  return ((width + kTileW - 1) / kTileW) * ((height + nw * (rh - 2) - 1) / (nw * (rh - 2)));
}

int dtk_device_cu_count(void)
{
  Workspace* w = nullptr;
  if (current_ws(&w, nullptr) != DTK_OK)
    return 0;
  return w->cu;
}

dtk_status dtk_ggdt(const dtk_image* img, const dtk_image* mask, dtk_image* dist, const dtk_params* p, void* stream,
                    dtk_stats* stats)
{
  dtk_status st;
  if ((st = check_params(p)) != DTK_OK || (st = check_triplet(img, mask, dist)) != DTK_OK)
    return st;
  const int   rh = region_rows(p);
  hipStream_t s  = static_cast<hipStream_t>(stream);
  if (img->width == 0 || img->height == 0)
    return fill_stats(stats, nullptr, 0, 0, rh, 0);
  Workspace* w = nullptr;
  if ((st = current_ws(&w, nullptr)) != DTK_OK)
    return st;
  WsUse use(*w, s);
  if ((st = use.begin()) != DTK_OK)
    return st;
  const bool c8  = p->connectivity == DTK_C8;
  const int  cm  = cost_mode(p);
  const int  S   = 64 / rh;
  KParams    kp  = make_kparams(p, img->dtype, mask->dtype, rh);
  ImgDesc    d   = make_desc(img, mask, dist, rh, tile_waves(p, rh));

  if (p->schedule == DTK_SCHED_RASTER)
  {
    // Strips of a pass wait on each other, so all of them must be resident.
    if ((long long)raster_strips(d, kp) > (long long)w->cu * raster_occupancy(kp, c8))
      return fail(DTK_ERR_UNSUPPORTED, "image too large for the resident raster pipeline");
    // Chunk numbers (16 or 20 steps each) must stay below 2^16 in the stamps.
    if (d.W >= (1 << 19) || d.H >= (1 << 19))
      return fail(DTK_ERR_UNSUPPORTED, "raster schedule: at most 2^19 - 1 pixels per side");
    const size_t mail_bytes = raster_mail_words(d, kp) * sizeof(unsigned long long);
    if ((st = ws_reserve(*w, 4096 + mail_bytes, 4096)) != DTK_OK)
      return st;
    uint32_t*           changed = static_cast<uint32_t*>(w->dbuf);
    unsigned long long* mail    = reinterpret_cast<unsigned long long*>(static_cast<char*>(w->dbuf) + 4096);
    uint32_t*           h       = static_cast<uint32_t*>(w->hbuf);
    DTK_HIP(hipMemsetAsync(changed, 0, 2 * sizeof(uint32_t), s), "memset");
    DTK_HIP(launch_raster_init(d, kp, changed + 1, s), "launch raster init");
    int      iters = 0;
    uint32_t pass0 = 0xffffffffu; // pass index stamped into the mailbox words
    for (;;)
    {
      // Stamps are (pass << 16 | chunk); clear the mailbox before the pass
      // index would wrap, so no word of an earlier pass can match.
      if (pass0 > 65531u)
      {
        DTK_HIP(hipMemsetAsync(mail, 0, mail_bytes, s), "memset mailbox");
        pass0 = 1;
      }
      DTK_HIP(hipMemsetAsync(changed, 0, sizeof(uint32_t), s), "memset");
      DTK_HIP(launch_raster_iteration(d, kp, c8, changed, mail, pass0, s), "launch raster pass");
      pass0 += 4;
      DTK_HIP(hipMemcpyAsync(h, changed, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read flag");
      DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
      ++iters;
      if (h[1] & ERR_MASK_RANGE)
        return fail(DTK_ERR_INVALID_ARGUMENT, "soft mask values must be in [0, 1]");
      if (h[1] & ERR_WATCHDOG)
        return fail(DTK_ERR_HIP, "device watchdog expired in the raster pipeline");
      if (!h[0] || (p->max_iterations > 0 && iters >= p->max_iterations))
        break;
    }
    return fill_stats(stats, nullptr, iters, 0, 0, 0);
  }

  const int ntiles = d.tiles_x * d.tiles_y;
  Batch     b{};
  b.d0     = d;
  b.nimg   = 1;
  b.ntiles = ntiles;

  if (p->schedule == DTK_SCHED_BLOCKS)
  {
    TaskBufs     tb;
    const size_t head   = align_up(task_layout(nullptr, ntiles, 1, 1, ORDER_FIFO, nullptr) + 2 * align_up(ntiles, 256) + 512, 256);
    const size_t cbytes = use_cost_cache(p) ? place_costs(&b.d0, 1, c8, nullptr) : 0;
    if ((st = ws_reserve(*w, head + cbytes, 4096)) != DTK_OK)
      return st;
    task_layout(w->dbuf, ntiles, 1, 1, ORDER_FIFO, &tb);
    if (cbytes)
      place_costs(&b.d0, 1, c8, static_cast<char*>(w->dbuf) + head);
    Queue q     = tb.q;
    q.slots     = nullptr; // blocks schedule: no queue
    q.bslots    = nullptr;
    uint8_t* a0 = static_cast<uint8_t*>(w->dbuf) + align_up(task_layout(nullptr, ntiles, 1, 1, ORDER_FIFO, nullptr), 256);
    uint8_t* a1 = a0 + align_up(ntiles, 256);
    KParams  kb = kp;
    kb.max_sweeps = 0;
    DTK_HIP(hipMemsetAsync(a0, 0, 2 * align_up(ntiles, 256), s), "memset active");
    DTK_HIP(launch_reset_ctrl(q.ctrl, s), "launch reset");
    DTK_HIP(launch_init(b, kb, q, s), "launch init");
    DTK_HIP(launch_blocks_seed(b, q, a0, s), "launch seed");
    uint32_t* h = static_cast<uint32_t*>(w->hbuf);
    // Rounds go out in batches with one read-back each: a round after
    // convergence finds no active tile and costs one empty launch, while a
    // read-back costs a host round trip.  Round r of the call stores r in
    // CTRL_CHANGED when it wakes a tile.  The loop ends once the last waking
    // round is not the batch's final one (the round after it ran and woke
    // nothing), and the call took (last waking round + 1) rounds.
    constexpr int kRoundsPerReadback = 8;
    uint32_t      round = 0, last_wake = 0;
    for (;;)
    {
      DTK_HIP(hipMemsetAsync(q.ctrl + CTRL_CHANGED, 0, sizeof(uint32_t), s), "memset");
      for (int k = 0; k < kRoundsPerReadback; ++k)
      {
        ++round;
        DTK_HIP(launch_blocks(S, c8, cm, b, kb, a0, a1, q.ctrl, p->round_per_blocks, round, s), "launch blocks");
        std::swap(a0, a1);
      }
      DTK_HIP(hipMemcpyAsync(h, q.ctrl, CTRL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read flag");
      DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
      if (h[CTRL_ERROR] & ERR_MASK_RANGE)
        return fail(DTK_ERR_INVALID_ARGUMENT, "soft mask values must be in [0, 1]");
      if (h[CTRL_CHANGED])
        last_wake = h[CTRL_CHANGED];
      if (last_wake < round)
        break;
      // Each round with a wake lowers some stored pixel, so the loop ends;
      // this bound (far above any measured round count) turns a protocol
      // fault into an error instead of an endless loop.
      if (round > 1000 + 64ull * ntiles)
        return fail(DTK_ERR_NOT_CONVERGED, "blocks schedule: no convergence after " + std::to_string(round) + " rounds");
    }
    return fill_stats(stats, h, (int)last_wake + 1, ntiles, rh, 0);
  }

  // Task schedule.
  const int      workers = task_workers(*w, p, S, c8, cm, ntiles);
  const uint32_t cap     = next_pow2((uint32_t)(ntiles + workers + 1));
  const int      order   = queue_order(p, 1);
  TaskBufs       tb;
  const size_t   head   = align_up(task_layout(nullptr, ntiles, cap, 1, order, nullptr), 256);
  const size_t   cbytes = use_cost_cache(p) ? place_costs(&b.d0, 1, c8, nullptr) : 0;
  if ((st = ws_reserve(*w, head + cbytes, 4096)) != DTK_OK)
    return st;
  task_layout(w->dbuf, ntiles, cap, 1, order, &tb);
  if (cbytes)
    place_costs(&b.d0, 1, c8, static_cast<char*>(w->dbuf) + head);
  tb.q.tile_visits = p->round_per_blocks;
  tb.q.mailbox     = w->mailbox;
  if ((st = enqueue_task(b, kp, tb.q, S, c8, cm, workers, s)) != DTK_OK)
    return st;
  if (stats)
  {
    uint32_t* h = nullptr;
    if ((st = read_ctrl(*w, tb.q, s, &h)) != DTK_OK)
      return st;
    return fill_stats(stats, h, 1, ntiles, rh, workers);
  }
  return DTK_OK;
}

dtk_status dtk_ggdt_batched(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                            const dtk_params* p, void* stream, dtk_stats* stats)
{
  dtk_status st;
  if ((st = check_params(p)) != DTK_OK)
    return st;
  if (n < 0 || (n > 0 && (!imgs || !masks || !dists)))
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad batch arrays");
  if (p->schedule != DTK_SCHED_TASK)
    return fail(DTK_ERR_INVALID_ARGUMENT, "batched calls use the task schedule");
  if (n == 0)
    return fill_stats(stats, nullptr, 0, 0, region_rows(p), 0);
  const int             rh = region_rows(p);
  std::vector<ImgDesc>  descs;
  std::vector<uint32_t> tile_img;
  int                   ntiles = 0;
  if ((st = build_batch(n, imgs, masks, dists, p, rh, descs, tile_img, &ntiles)) != DTK_OK)
    return st;
  if (ntiles == 0)
    return fill_stats(stats, nullptr, 0, 0, rh, 0);
  Workspace* w = nullptr;
  if ((st = current_ws(&w, nullptr)) != DTK_OK)
    return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  WsUse       use(*w, s);
  if ((st = use.begin()) != DTK_OK)
    return st;
  const bool     c8  = p->connectivity == DTK_C8;
  const int      cm  = cost_mode(p);
  const int      S   = 64 / rh;
  const int      workers = task_workers(*w, p, S, c8, cm, ntiles);
  const uint32_t cap     = next_pow2((uint32_t)(ntiles + workers + 1));
  const int      order   = queue_order(p, n);
  const size_t   hbytes  = descs.size() * sizeof(ImgDesc) + tile_img.size() * 4 + 1024;
  const size_t   head    = align_up(task_layout(nullptr, ntiles, cap, n > 1 ? n : 2, order, nullptr), 256);
  const size_t   cbytes  = use_cost_cache(p) ? place_costs(descs.data(), n, c8, nullptr) : 0;
  if ((st = ws_reserve(*w, head + cbytes, 4096)) != DTK_OK)
    return st;
  TaskBufs tb;
  task_layout(w->dbuf, ntiles, cap, n > 1 ? n : 2, order, &tb);
  tb.q.mailbox = w->mailbox;
  if (cbytes)
    place_costs(descs.data(), n, c8, static_cast<char*>(w->dbuf) + head);
  Batch b{};
  b.d0     = descs[0];
  b.nimg   = n;
  b.ntiles = ntiles;
  if (n > 1)
  {
    Stage* stg = nullptr;
    if ((st = stage_get(*w, hbytes, &stg)) != DTK_OK)
      return st;
    char* h = static_cast<char*>(stg->h);
    std::memcpy(h, descs.data(), descs.size() * sizeof(ImgDesc));
    std::memcpy(h + descs.size() * sizeof(ImgDesc), tile_img.data(), tile_img.size() * 4);
    DTK_HIP(hipMemcpyAsync(tb.table, h, descs.size() * sizeof(ImgDesc), hipMemcpyHostToDevice, s), "upload table");
    DTK_HIP(hipMemcpyAsync(tb.tile_img, h + descs.size() * sizeof(ImgDesc), tile_img.size() * 4,
                           hipMemcpyHostToDevice, s),
            "upload tile table");
    if ((st = stage_done(stg, s)) != DTK_OK)
      return st;
    b.table    = tb.table;
    b.tile_img = tb.tile_img;
  }
  KParams kp = make_kparams(p, imgs[0].dtype, masks[0].dtype, rh);
  if ((st = enqueue_task(b, kp, tb.q, S, c8, cm, workers, s)) != DTK_OK)
    return st;
  if (stats)
  {
    uint32_t* h = nullptr;
    if ((st = read_ctrl(*w, tb.q, s, &h)) != DTK_OK)
      return st;
    return fill_stats(stats, h, 1, ntiles, rh, workers);
  }
  return DTK_OK;
}

struct dtk_batch_plan
{
  int             device  = 0;
  void*           dbuf    = nullptr;
  hipStream_t     cstream = nullptr;
  hipGraph_t      graph   = nullptr;
  hipGraphExec_t  exec    = nullptr;
  int             ntiles  = 0;
};

dtk_status dtk_batch_plan_create(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                                 const dtk_params* p, void* stream, dtk_batch_plan** plan)
{
  dtk_status st;
  if (!plan)
    return fail(DTK_ERR_INVALID_ARGUMENT, "plan is NULL");
  *plan = nullptr;
  if ((st = check_params(p)) != DTK_OK)
    return st;
  if (n <= 0 || !imgs || !masks || !dists)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad batch arrays");
  if (p->schedule != DTK_SCHED_TASK)
    return fail(DTK_ERR_INVALID_ARGUMENT, "batched plans use the task schedule");
  const int             rh = region_rows(p);
  std::vector<ImgDesc>  descs;
  std::vector<uint32_t> tile_img;
  int                   ntiles = 0;
  if ((st = build_batch(n, imgs, masks, dists, p, rh, descs, tile_img, &ntiles)) != DTK_OK)
    return st;
  if (ntiles == 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "batch has no pixels");
  Workspace* w   = nullptr;
  int        dev = 0;
  if ((st = current_ws(&w, &dev)) != DTK_OK)
    return st;
  const bool     c8      = p->connectivity == DTK_C8;
  const int      cm      = cost_mode(p);
  const int      S       = 64 / rh;
  const int      workers = task_workers(*w, p, S, c8, cm, ntiles);
  const uint32_t cap     = next_pow2((uint32_t)(ntiles + workers + 1));
  dtk_batch_plan* pl     = new dtk_batch_plan();
  pl->device             = dev;
  pl->ntiles             = ntiles;
  const int    order     = queue_order(p, n);
  const size_t head      = align_up(task_layout(nullptr, ntiles, cap, n > 1 ? n : 2, order, nullptr), 256);
  const size_t cbytes    = use_cost_cache(p) ? place_costs(descs.data(), n, c8, nullptr) : 0;
  hipError_t   e         = hipMalloc(&pl->dbuf, head + cbytes);
  if (e != hipSuccess)
  {
    delete pl;
    return fail(DTK_ERR_OUT_OF_MEMORY, std::string("hipMalloc(plan): ") + hipGetErrorString(e));
  }
  TaskBufs tb;
  task_layout(pl->dbuf, ntiles, cap, n > 1 ? n : 2, order, &tb);
  tb.q.mailbox = w->mailbox;
  if (cbytes)
    place_costs(descs.data(), n, c8, static_cast<char*>(pl->dbuf) + head);
  Batch b{};
  b.d0     = descs[0];
  b.nimg   = n;
  b.ntiles = ntiles;
  if (n > 1)
  {
    e = hipMemcpy(tb.table, descs.data(), descs.size() * sizeof(ImgDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(tb.tile_img, tile_img.data(), tile_img.size() * 4, hipMemcpyHostToDevice);
    b.table    = tb.table;
    b.tile_img = tb.tile_img;
  }
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&pl->cstream, hipStreamNonBlocking);
  if (e != hipSuccess)
  {
    dtk_batch_plan_destroy(pl);
    return fail(DTK_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(e));
  }
  KParams kp = make_kparams(p, imgs[0].dtype, masks[0].dtype, rh);
  e          = hipStreamBeginCapture(pl->cstream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess)
  {
    st            = enqueue_task(b, kp, tb.q, S, c8, cm, workers, pl->cstream, false);
    hipError_t e2 = hipStreamEndCapture(pl->cstream, &pl->graph);
    if (st != DTK_OK)
    {
      dtk_batch_plan_destroy(pl);
      return st;
    }
    e = e2;
  }
  if (e == hipSuccess)
    e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
  if (e != hipSuccess)
  {
    dtk_batch_plan_destroy(pl);
    return fail(DTK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  (void)stream;
  *plan = pl;
  return DTK_OK;
}

dtk_status dtk_batch_plan_launch(dtk_batch_plan* plan, void* stream)
{
  if (!plan || !plan->exec)
    return fail(DTK_ERR_INVALID_ARGUMENT, "invalid plan");
  // Errors of earlier launches (this plan's or other calls' on its device).
  if (plan->device >= 0 && plan->device < kMaxDevices)
  {
    Workspace& w = g_ws[plan->device];
    dtk_status st;
    if ((st = take_async_error(w)) != DTK_OK)
      return st;
  }
  DTK_HIP(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  return DTK_OK;
}

dtk_status dtk_batch_plan_destroy(dtk_batch_plan* plan)
{
  if (!plan)
    return DTK_OK;
  hipDeviceSynchronize();
  if (plan->exec)
    hipGraphExecDestroy(plan->exec);
  if (plan->graph)
    hipGraphDestroy(plan->graph);
  if (plan->cstream)
    hipStreamDestroy(plan->cstream);
  if (plan->dbuf)
    hipFree(plan->dbuf);
  delete plan;
  return DTK_OK;
}

namespace
{
  struct TuneEntry
  {
    dtk_params best;
    float      best_ms;
    float      default_ms;
  };
  std::mutex                       g_tune_mu;
  std::map<std::string, TuneEntry> g_tune;

  std::string tune_key(int dev, int n, const dtk_image* imgs, const dtk_image* masks, const dtk_params* p)
  {
    long long total = 0;
    int       side  = 0;
    for (int i = 0; i < n; ++i)
    {
      total += (long long)imgs[i].width * imgs[i].height;
      side = std::max(side, std::max(imgs[i].width, imgs[i].height));
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%d|%d|%lld|%d|%d|%d|%d|%a|%a|%d", dev, n, total, side, imgs[0].dtype,
                  masks[0].dtype, p->connectivity, (double)p->lambda, (double)p->v, p->cost_plane);
    return buf;
  }

  // Mean launch time of one candidate plan (warm-up launch excluded).
  dtk_status time_plan(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists, const dtk_params* p,
                       hipStream_t s, int reps, float* ms)
  {
    dtk_batch_plan* plan = nullptr;
    dtk_status      st   = dtk_batch_plan_create(n, imgs, masks, dists, p, s, &plan);
    if (st != DTK_OK)
      return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e  = hipEventCreate(&e0);
    if (e == hipSuccess)
      e = hipEventCreate(&e1);
    if (e == hipSuccess)
      e = hipGraphLaunch(plan->exec, s);
    if (e == hipSuccess)
      e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r)
      e = hipGraphLaunch(plan->exec, s);
    if (e == hipSuccess)
      e = hipEventRecord(e1, s);
    if (e == hipSuccess)
      e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess)
      e = hipEventElapsedTime(&t, e0, e1);
    if (e0)
      hipEventDestroy(e0);
    if (e1)
      hipEventDestroy(e1);
    dtk_batch_plan_destroy(plan);
    if (e != hipSuccess)
      return fail(DTK_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
    *ms = t / (float)reps;
    return DTK_OK;
  }
} // namespace

dtk_status dtk_batch_autotune(int n, const dtk_image* imgs, const dtk_image* masks, dtk_image* dists,
                              const dtk_params* p, void* stream, int reps, dtk_params* best, dtk_tune_report* report)
{
  dtk_status st;
  if (!best)
    return fail(DTK_ERR_INVALID_ARGUMENT, "best is NULL");
  if ((st = check_params(p)) != DTK_OK)
    return st;
  if (n <= 0 || !imgs || !masks || !dists)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad batch arrays");
  if (p->schedule != DTK_SCHED_TASK)
    return fail(DTK_ERR_INVALID_ARGUMENT, "batched plans use the task schedule");
  if (reps <= 0)
    reps = 2;
  int dev = 0;
  DTK_HIP(hipGetDevice(&dev), "hipGetDevice");
  hipStream_t       s   = static_cast<hipStream_t>(stream);
  const std::string key = tune_key(dev, n, imgs, masks, p);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto                        it = g_tune.find(key);
    if (it != g_tune.end())
    {
      *best = it->second.best;
      if (report)
        *report = dtk_tune_report{0, 1, it->second.best_ms, it->second.default_ms};
      return DTK_OK;
    }
  }
  const bool c8   = p->connectivity == DTK_C8;
  dtk_params cand = *p;
  dtk_params top  = *p;
  float      top_ms = 0.f, default_ms = 0.f;
  int        count  = 0;
  dtk_params dflt   = *p;
  dflt.tile_rows = dflt.tile_waves = dflt.waves_per_cu = dflt.max_sweeps = 0;
  dflt.queue_order = DTK_ORDER_DEFAULT;
  const int drows = region_rows(&dflt), dnw = tile_waves(&dflt, drows);
  // Queue orders first, at the default tile shape; the shapes are then
  // timed under the fastest order.
  {
    float best_order_ms = 0.f;
    int   best_order    = DTK_ORDER_KEY;
    for (int order : {DTK_ORDER_KEY, DTK_ORDER_FIFO, DTK_ORDER_SEED})
    {
      dtk_params oc  = dflt;
      oc.queue_order = order;
      float ms       = 0.f;
      if ((st = time_plan(n, imgs, masks, dists, &oc, s, reps, &ms)) != DTK_OK)
        return st;
      if (queue_order(&oc, n) == queue_order(&dflt, n))
        default_ms = ms;
      if (count == 0 || ms < best_order_ms)
      {
        best_order_ms = ms;
        best_order    = order;
        top           = oc;
        top_ms        = ms;
      }
      ++count;
    }
    cand.queue_order = best_order;
  }
  const int shapes[6][2] = {{16, 4}, {32, 4}, {32, 2}, {16, 2}, {32, 1}, {16, 1}}; // tile_rows, tile_waves
  for (const auto& shp : shapes)
  {
    const int rows = shp[0], nw = shp[1];
    for (int cap : {0, 12})
    {
      cand.tile_rows  = rows;
      cand.tile_waves = nw;
      cand.max_sweeps = cap;
      const int occ   = std::max(1, task_occupancy(64 / rows, c8, cost_mode(&cand), nw)); // workers per CU
      // Worker counts per CU: the default (0), then all, 3/4 and 1/2 of
      // what fits, each tried once.
      const int wpcs[4] = {0, occ * nw, (occ * 3 + 3) / 4 * nw, (occ + 1) / 2 * nw};
      int       seen[4] = {0, 0, 0, 0};
      for (int wi = 0; wi < 4; ++wi)
      {
        const int per_cu = wi == 0 ? (nw == 4 ? std::min(occ, 2) : occ) : std::max(1, std::min(occ, wpcs[wi] / nw));
        bool      dup    = false;
        for (int k = 0; k < wi; ++k)
          dup |= seen[k] == per_cu;
        seen[wi] = per_cu;
        if (dup)
          continue;
        cand.waves_per_cu = wpcs[wi];
        float ms          = 0.f;
        if ((st = time_plan(n, imgs, masks, dists, &cand, s, reps, &ms)) != DTK_OK)
          return st;
        if (ms < top_ms)
        {
          top    = cand;
          top_ms = ms;
        }
        ++count;
      }
    }
  }
  *best = top;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[key] = TuneEntry{top, top_ms, default_ms};
  }
  if (report)
    *report = dtk_tune_report{count, 0, top_ms, default_ms};
  return DTK_OK;
}

dtk_status dtk_ggdt_pred(const dtk_image* img, const dtk_image* mask, const dtk_image* dist, dtk_image* pred,
                         const dtk_params* p, void* stream)
{
  dtk_status st;
  if ((st = check_params(p)) != DTK_OK)
    return st;
  dtk_image dcopy = dist ? *dist : dtk_image{};
  if ((st = check_triplet(img, mask, dist ? &dcopy : nullptr)) != DTK_OK)
    return st;
  if ((st = check_image(pred, "pred", 1 << DTK_U8)) != DTK_OK)
    return st;
  if (pred->width != img->width || pred->height != img->height)
    return fail(DTK_ERR_INVALID_ARGUMENT, "pred size differs");
  Workspace* w = nullptr;
  if ((st = current_ws(&w, nullptr)) != DTK_OK || (st = take_async_error(*w)) != DTK_OK)
    return st;
  ImgDesc d    = make_desc(img, mask, &dcopy, 32, 1);
  d.pred       = static_cast<unsigned char*>(pred->data);
  d.pred_pitch = pred->pitch;
  KParams kp   = make_kparams(p, img->dtype, mask->dtype, 32);
  DTK_HIP(launch_pred(d, kp, p->connectivity == DTK_C8, static_cast<hipStream_t>(stream)), "launch pred");
  return DTK_OK;
}

dtk_status dtk_ggdt_pred_batched(int n, const dtk_image* imgs, const dtk_image* masks, const dtk_image* dists,
                                 dtk_image* preds, const dtk_params* p, void* stream)
{
  dtk_status st;
  if (n < 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "negative image count");
  if (n > 0 && (!imgs || !masks || !dists || !preds))
    return fail(DTK_ERR_INVALID_ARGUMENT, "NULL image array");
  if ((st = check_params(p)) != DTK_OK)
    return st;
  int maxW = 0, maxH = 0;
  for (int i = 0; i < n; ++i)
  {
    if ((st = check_triplet(&imgs[i], &masks[i], &dists[i])) != DTK_OK)
      return st;
    if (imgs[i].dtype != imgs[0].dtype || masks[i].dtype != masks[0].dtype)
      return fail(DTK_ERR_INVALID_ARGUMENT, "all images of a batch must share image and mask dtypes");
    if ((st = check_image(&preds[i], "pred", 1 << DTK_U8)) != DTK_OK)
      return st;
    if (preds[i].width != imgs[i].width || preds[i].height != imgs[i].height)
      return fail(DTK_ERR_INVALID_ARGUMENT, "pred size differs");
    if (imgs[i].height > 65535)
      return fail(DTK_ERR_UNSUPPORTED, "a batch takes at most 65535 rows per image");
    maxW = std::max(maxW, imgs[i].width);
    maxH = std::max(maxH, imgs[i].height);
  }
  if (n == 0 || maxW == 0 || maxH == 0)
    return DTK_OK;
  Workspace* w = nullptr;
  if ((st = current_ws(&w, nullptr)) != DTK_OK)
    return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  WsUse       use(*w, s);
  if ((st = use.begin()) != DTK_OK)
    return st;
  const size_t tbytes = align_up((size_t)n * sizeof(ImgDesc), 256);
  if ((st = ws_reserve(*w, tbytes + 1024, 4096)) != DTK_OK)
    return st;
  Stage* stg = nullptr;
  if ((st = stage_get(*w, tbytes, &stg)) != DTK_OK)
    return st;
  Carver      c(w->dbuf);
  ImgDesc*    dtab = c.take<ImgDesc>(n);
  ImgDesc*    htab = static_cast<ImgDesc*>(stg->h);
  for (int i = 0; i < n; ++i)
  {
    dtk_image dc       = dists[i];
    htab[i]            = make_desc(&imgs[i], &masks[i], &dc, 32, 1);
    htab[i].pred       = static_cast<unsigned char*>(preds[i].data);
    htab[i].pred_pitch = preds[i].pitch;
  }
  DTK_HIP(hipMemcpyAsync(dtab, htab, (size_t)n * sizeof(ImgDesc), hipMemcpyHostToDevice, s), "upload table");
  if ((st = stage_done(stg, s)) != DTK_OK)
    return st;
  const KParams kp = make_kparams(p, imgs[0].dtype, masks[0].dtype, 32);
  DTK_HIP(launch_pred_batched(dtab, n, maxW, maxH, kp, p->connectivity == DTK_C8, s), "launch pred");
  return DTK_OK;
}

dtk_status dtk_ggdt_backward(const dtk_image* img, const dtk_image* mask, const dtk_image* dist,
                             const dtk_image* pred, const dtk_image* grad_dist, dtk_image* grad_img,
                             dtk_image* grad_mask, const dtk_params* p, void* stream)
{
  if (!img || !mask || !pred || !grad_dist || !grad_img)
    return fail(DTK_ERR_INVALID_ARGUMENT, "NULL image");
  dtk_image dcopy = dist ? *dist : dtk_image{};
  return backward_impl(1, img, mask, dist ? &dcopy : nullptr, pred, grad_dist, grad_img, grad_mask,
                       grad_mask != nullptr, p, stream);
}

dtk_status dtk_ggdt_backward_batched(int n, const dtk_image* imgs, const dtk_image* masks, const dtk_image* dists,
                                     const dtk_image* preds, const dtk_image* grad_dists, dtk_image* grad_imgs,
                                     dtk_image* grad_masks, const dtk_params* p, void* stream)
{
  if (n < 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "negative image count");
  if (n > 0 && (!imgs || !masks || !dists || !preds || !grad_dists || !grad_imgs))
    return fail(DTK_ERR_INVALID_ARGUMENT, "NULL image array");
  return backward_impl(n, imgs, masks, dists, preds, grad_dists, grad_imgs, grad_masks, grad_masks != nullptr, p,
                       stream);
}

dtk_status dtk_malloc_pitch(void** ptr, long long* pitch, long long width_bytes, int height)
{
  if (!ptr || !pitch || width_bytes < 0 || height < 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad dtk_malloc_pitch arguments");
  size_t     pp  = 0;
  void*      out = nullptr;
  hipError_t e   = hipMallocPitch(&out, &pp, (size_t)(width_bytes > 0 ? width_bytes : 1), (size_t)(height > 0 ? height : 1));
  if (e != hipSuccess)
    return fail(DTK_ERR_OUT_OF_MEMORY, std::string("Unable to allocate GPU memory: ") + hipGetErrorString(e));
  *ptr   = out;
  *pitch = (long long)pp;
  return DTK_OK;
}

dtk_status dtk_free(void* ptr)
{
  DTK_HIP(hipFree(ptr), "hipFree");
  return DTK_OK;
}

static dtk_status copy2d(const dtk_image* src, dtk_image* dst, hipMemcpyKind kind, const char* what)
{
  if (!src || !dst)
    return fail(DTK_ERR_INVALID_ARGUMENT, "NULL image");
  if (src->width != dst->width || src->height != dst->height || src->dtype != dst->dtype)
    return fail(DTK_ERR_INVALID_ARGUMENT, "transfer between different image shapes");
  const int es = elem_size(src->dtype);
  if (es == 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "unsupported dtype");
  if (src->width == 0 || src->height == 0)
    return DTK_OK;
  hipError_t e = hipMemcpy2D(dst->data, (size_t)dst->pitch, src->data, (size_t)src->pitch,
                             (size_t)src->width * es, (size_t)src->height, kind);
  if (e != hipSuccess)
    return fail(DTK_ERR_HIP, std::string(what) + hipGetErrorString(e));
  return DTK_OK;
}

dtk_status dtk_host_to_device(const dtk_image* src, dtk_image* dst)
{
  return copy2d(src, dst, hipMemcpyHostToDevice, "Unable to transfert host to device: ");
}

dtk_status dtk_device_to_host(const dtk_image* src, dtk_image* dst)
{
  return copy2d(src, dst, hipMemcpyDeviceToHost, "Unable to transfert device to host: ");
}

dtk_status dtk_profile_begin(void)
{
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.on   = true;
  g_prof.used = 0;
  return DTK_OK;
}

dtk_status dtk_profile_end(double* setup_ms, double* solve_ms, int* calls)
{
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.on  = false;
  double a = 0, b = 0;
  int    n = 0;
  for (size_t i = 0; i + 3 <= g_prof.used; i += 3)
  {
    float t0 = 0, t1 = 0;
    DTK_HIP(hipEventSynchronize(g_prof.ev[i + 2]), "hipEventSynchronize");
    DTK_HIP(hipEventElapsedTime(&t0, g_prof.ev[i], g_prof.ev[i + 1]), "hipEventElapsedTime");
    DTK_HIP(hipEventElapsedTime(&t1, g_prof.ev[i + 1], g_prof.ev[i + 2]), "hipEventElapsedTime");
    a += t0;
    b += t1;
    ++n;
  }
  if (setup_ms)
    *setup_ms = a;
  if (solve_ms)
    *solve_ms = b;
  if (calls)
    *calls = n;
  g_prof.used = 0;
  return DTK_OK;
}

dtk_status dtk_copy2d(void* dst, long long dst_pitch, const void* src, long long src_pitch, long long width_bytes,
                      int height, int kind)
{
  if (width_bytes < 0 || height < 0 || kind < 1 || kind > 3)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad dtk_copy2d arguments");
  if (width_bytes == 0 || height == 0)
    return DTK_OK;
  if (!dst || !src || dst_pitch < width_bytes || src_pitch < width_bytes)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad dtk_copy2d buffers");
  const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  const hipError_t    e = hipMemcpy2D(dst, (size_t)dst_pitch, src, (size_t)src_pitch, (size_t)width_bytes, (size_t)height, k);
  if (e != hipSuccess)
    return fail(DTK_ERR_HIP, std::string(kind == 1 ? "Unable to transfert host to device: "
                                                    : (kind == 2 ? "Unable to transfert device to host: "
                                                                 : "Unable to copy on device: ")) +
                                 hipGetErrorString(e));
  return DTK_OK;
}


dtk_status dtk_block_priorities(const dtk_image* mask, uint8_t* priorities, uint32_t* offsets, uint32_t* level0,
                                void* stream)
{
  dtk_status st;
  if ((st = check_image(mask, "mask", (1 << DTK_U8) | (1 << DTK_F32))) != DTK_OK)
    return st;
  if (!level0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "level0 is NULL");
  *level0 = 0;
  if ((long long)mask->width * mask->height == 0)
    return DTK_OK;
  if (!priorities || !offsets)
    return fail(DTK_ERR_INVALID_ARGUMENT, "priorities / offsets is NULL");
  int gw, gh;
  prio_grid(mask->width, mask->height, &gw, &gh);
  if (gh > 65535)
    return fail(DTK_ERR_UNSUPPORTED, "block priorities: at most 65535 tile rows");
  const size_t n = (size_t)gw * gh;
  Workspace*   w = nullptr;
  if ((st = current_ws(&w, nullptr)) != DTK_OK)
    return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  WsUse       use(*w, s);
  if ((st = use.begin()) != DTK_OK)
    return st;
  const size_t dist_bytes = align_up(n * sizeof(uint32_t), 256);
  const size_t hist_bytes = align_up(((size_t)gw + gh) * sizeof(uint32_t), 256);
  if ((st = ws_reserve(*w, 256 + dist_bytes + hist_bytes, 4096)) != DTK_OK)
    return st;
  char*       base = static_cast<char*>(w->dbuf);
  uint32_t*   l0   = reinterpret_cast<uint32_t*>(base);
  uint32_t*   dist = reinterpret_cast<uint32_t*>(base + 256);
  uint32_t*   hist = reinterpret_cast<uint32_t*>(base + 256 + dist_bytes);
  DTK_HIP(launch_block_priorities(mask->data, mask->pitch, mask->dtype == DTK_F32, mask->width, mask->height, dist,
                                  hist, priorities, offsets, l0, s),
          "launch block priorities");
  uint32_t* h = static_cast<uint32_t*>(w->hbuf);
  DTK_HIP(hipMemcpyAsync(h, l0, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read level0");
  DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
  *level0 = h[0];
  return DTK_OK;
}

dtk_status dtk_minmax_f32(const dtk_image* img, float* min_max, void* stream)
{
  dtk_status st;
  if ((st = check_image(img, "img", 1 << DTK_F32)) != DTK_OK)
    return st;
  if (!min_max)
    return fail(DTK_ERR_INVALID_ARGUMENT, "min_max is NULL");
  hipStream_t s    = static_cast<hipStream_t>(stream);
  uint32_t    k[2] = {0x7f7fffffu | 0x80000000u, 0x00800000u | 0x80000000u}; // keys of FLT_MAX, FLT_MIN
  if ((long long)img->width * img->height > 0)
  {
    st = with_minmax_keys(img, s, [&](uint32_t* keys) {
      DTK_HIP(hipMemcpyAsync(k, keys, sizeof(k), hipMemcpyDeviceToHost, s), "read min/max");
      DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
      return DTK_OK;
    });
    if (st != DTK_OK)
      return st;
  }
  min_max[0] = key_to_float(k[0]);
  min_max[1] = key_to_float(k[1]);
  return DTK_OK;
}

dtk_status dtk_normalize_u8(const dtk_image* in, dtk_image* out, void* stream)
{
  dtk_status st;
  if ((st = check_pair(in, 1 << DTK_F32, out, 1 << DTK_U8)) != DTK_OK)
    return st;
  if ((long long)in->width * in->height == 0)
    return DTK_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return with_minmax_keys(in, s, [&](uint32_t* keys) {
    DTK_HIP(launch_normalize_u8(in->data, in->pitch, out->data, out->pitch, in->width, in->height, keys, s),
            "launch normalize");
    return DTK_OK;
  });
}

dtk_status dtk_inferno(const dtk_image* in, dtk_image* out, void* stream)
{
  dtk_status st;
  if ((st = check_pair(in, 1 << DTK_U8, out, 1 << DTK_RGB8)) != DTK_OK)
    return st;
  if ((long long)in->width * in->height == 0)
    return DTK_OK;
  DTK_HIP(launch_inferno(in->data, in->pitch, out->data, out->pitch, in->width, in->height,
                         static_cast<hipStream_t>(stream)),
          "launch inferno");
  return DTK_OK;
}

dtk_status dtk_colorize(const dtk_image* in, dtk_image* out, void* stream)
{
  dtk_status st;
  if ((st = check_pair(in, 1 << DTK_F32, out, 1 << DTK_RGB8)) != DTK_OK)
    return st;
  if ((long long)in->width * in->height == 0)
    return DTK_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return with_minmax_keys(in, s, [&](uint32_t* keys) {
    DTK_HIP(launch_colorize(in->data, in->pitch, out->data, out->pitch, in->width, in->height, keys, s),
            "launch colorize");
    return DTK_OK;
  });
}

int dtk_tile_size_for(const dtk_params* p, int* tile_w, int* tile_h)
{
  if (!p || !tile_w || !tile_h || check_params(p) != DTK_OK)
    return 0;
  const int rh = region_rows(p), nw = tile_waves(p, rh);
  *tile_w      = kTileW;
  *tile_h      = nw * (rh - 2);
  return 1;
}

dtk_status dtk_tile_image(const void* values, int value_bytes, int tiles_x, int tiles_y, int tile_w, int tile_h,
                          dtk_image* out, void* stream)
{
  dtk_status st;
  if ((st = check_image(out, "out", 1 << DTK_F32)) != DTK_OK)
    return st;
  if (value_bytes != 1 && value_bytes != 4)
    return fail(DTK_ERR_INVALID_ARGUMENT, "value_bytes must be 1 (uint8) or 4 (int32)");
  if (tiles_x < 0 || tiles_y < 0 || tile_w <= 0 || tile_h <= 0)
    return fail(DTK_ERR_INVALID_ARGUMENT, "bad tile grid");
  if ((long long)out->width * out->height == 0)
    return DTK_OK;
  if ((long long)tiles_x * tile_w < out->width || (long long)tiles_y * tile_h < out->height)
    return fail(DTK_ERR_INVALID_ARGUMENT, "the tile grid does not cover the image");
  if (out->height > 4 * 65535)
    return fail(DTK_ERR_UNSUPPORTED, "tile image: at most 262140 rows");
  if (!values)
    return fail(DTK_ERR_INVALID_ARGUMENT, "values is NULL");
  DTK_HIP(launch_tile_image(values, value_bytes, tiles_x, tile_w, tile_h, out->data, out->pitch, out->width,
                            out->height, static_cast<hipStream_t>(stream)),
          "launch tile image");
  return DTK_OK;
}

} // extern "C"

// ---- the allocation-free entry point ---------------------------------------
namespace
{
  // Scratch of dtk_ggdt_noalloc: the queue, tile statistics and control
  // words of one call, in a file-scope device array (each device holds its
  // own copy from module load until the process ends).  Host code reaches
  // it only through the address hipGetSymbolAddress returns with that
  // device current.
  constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
  static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
  // Error bits that the kernels of dtk_ggdt_noalloc calls leave on a device,
  // kept until dtk_sync_errors or a call with statistics takes them.
  static __device__ __align__(256) uint32_t scratch_error[64];

  struct NoAllocDevice
  {
    std::once_flag once;
    hipError_t     status = hipSuccess;
    void*          ptr    = nullptr; // device address of `scratch`
    uint32_t*      err    = nullptr; // device address of `scratch_error`
    int            cu     = 0;
  };
  NoAllocDevice g_noalloc[kMaxDevices];

  // Set-up on the device's first call (the device is current): the device
  // addresses of the scratch arrays and the CU count.  Allocates nothing.
  dtk_status noalloc_device(int dev, NoAllocDevice** out)
  {
    NoAllocDevice& nd = g_noalloc[dev];
    std::call_once(nd.once, [&nd, dev] {
      hipError_t e = hipGetSymbolAddress(&nd.ptr, HIP_SYMBOL(scratch));
      if (e == hipSuccess)
        e = hipGetSymbolAddress(reinterpret_cast<void**>(&nd.err), HIP_SYMBOL(scratch_error));
      if (e == hipSuccess)
        e = hipDeviceGetAttribute(&nd.cu, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess && nd.cu <= 0)
        nd.cu = 1;
      nd.status = e;
    });
    if (nd.status != hipSuccess)
      return fail(DTK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(nd.status));
    *out = &nd;
    return DTK_OK;
  }

  // The sticky error bits of the allocation-free calls on device `dev`
  // (synchronous; only once that device has been set up).
  dtk_status take_noalloc_error(int dev)
  {
    NoAllocDevice& nd = g_noalloc[dev];
    if (!nd.err)
      return DTK_OK;
    uint32_t bits = 0u;
    DTK_HIP(hipMemcpy(&bits, nd.err, sizeof(bits), hipMemcpyDeviceToHost), "read scratch error");
    if (!bits)
      return DTK_OK;
    const uint32_t zero = 0u;
    DTK_HIP(hipMemcpy(nd.err, &zero, sizeof(zero), hipMemcpyHostToDevice), "clear scratch error");
    return error_from_bits(bits, " (an earlier dtk_ggdt_noalloc call on this device)");
  }
} // namespace

extern "C" {

dtk_status dtk_ggdt_noalloc(const dtk_image* img, const dtk_image* mask, dtk_image* dist, const dtk_params* p,
                            void* stream, dtk_stats* stats)
{
  dtk_status st;
  if ((st = check_params(p)) != DTK_OK || (st = check_triplet(img, mask, dist)) != DTK_OK)
    return st;
  if (p->schedule != DTK_SCHED_TASK)
    return fail(DTK_ERR_UNSUPPORTED, "dtk_ggdt_noalloc runs the task schedule");
  if (use_cost_cache(p))
    return fail(DTK_ERR_UNSUPPORTED, "dtk_ggdt_noalloc takes no cost plane");
  const int   rh = region_rows(p);
  hipStream_t s  = static_cast<hipStream_t>(stream);
  if (img->width == 0 || img->height == 0)
    return fill_stats(stats, nullptr, 0, 0, rh, 0);
  int dev = 0;
  DTK_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DTK_ERR_UNSUPPORTED, "device index out of range");
  NoAllocDevice* nd = nullptr;
  if ((st = noalloc_device(dev, &nd)) != DTK_OK)
    return st;
  const bool     c8      = p->connectivity == DTK_C8;
  const int      cm      = cost_mode(p);
  const int      S       = 64 / rh;
  const KParams  kp      = make_kparams(p, img->dtype, mask->dtype, rh);
  const ImgDesc  d       = make_desc(img, mask, dist, rh, tile_waves(p, rh));
  const int      ntiles  = d.tiles_x * d.tiles_y;
  const int      workers = task_workers_cu(nd->cu, p, S, c8, cm, ntiles);
  const uint32_t cap     = next_pow2((uint32_t)(ntiles + workers + 1));
  const int      order   = queue_order(p, 1);
  const size_t   need    = task_layout(nullptr, ntiles, cap, 1, order, nullptr);
  if (need > kNoAllocScratchBytes)
    return fail(DTK_ERR_UNSUPPORTED, "dtk_ggdt_noalloc: this call needs " + std::to_string(need) +
                                       " bytes of scratch, at most " + std::to_string(kNoAllocScratchBytes) +
                                       " are available (use dtk_ggdt)");
  TaskBufs tb;
  task_layout(nd->ptr, ntiles, cap, 1, order, &tb);
  tb.q.tile_visits = p->round_per_blocks;
  tb.q.mailbox     = nd->err;
  Batch b{};
  b.d0     = d;
  b.nimg   = 1;
  b.ntiles = ntiles;
  // Every word that must start at zero (control words, ring slots, tile
  // states and keys) is cleared by the call's own kernels.
  if ((st = enqueue_task(b, kp, tb.q, S, c8, cm, workers, s, false)) != DTK_OK)
    return st;
  if (stats)
  {
    uint32_t h[CTRL_WORDS];
    DTK_HIP(hipMemcpyAsync(h, tb.q.ctrl, sizeof(h), hipMemcpyDeviceToHost, s), "read counters");
    DTK_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (h[CTRL_ERROR])
    {
      const uint32_t zero = 0u; // reported here, not again
      DTK_HIP(hipMemcpy(nd->err, &zero, sizeof(zero), hipMemcpyHostToDevice), "clear scratch error");
      return error_from_bits(h[CTRL_ERROR], "");
    }
    return fill_stats(stats, h, 1, ntiles, rh, workers);
  }
  return DTK_OK;
}

dtk_status dtk_sync_errors(void* stream)
{
  DTK_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
  int dev = 0;
  DTK_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DTK_ERR_UNSUPPORTED, "device index out of range");
  dtk_status st = DTK_OK;
  Workspace& w  = g_ws[dev];
  if (w.mailbox)
  {
    std::lock_guard<std::mutex> lock(w.mu);
    st = take_async_error(w);
  }
  const dtk_status st2 = take_noalloc_error(dev);
  return st != DTK_OK ? st : st2;
}

dtk_status dtk_selftest(int* out3)
{
  if (!out3)
    return fail(DTK_ERR_INVALID_ARGUMENT, "out is NULL");
  const int  n   = 3 * 255 * 255 + 1;
  int*       dd  = nullptr;
  float*     ds  = nullptr;
  DTK_HIP(hipMalloc(&dd, sizeof(int) * 4), "hipMalloc");
  hipError_t e = hipMalloc(&ds, sizeof(float) * n);
  if (e != hipSuccess)
  {
    hipFree(dd);
    return fail(DTK_ERR_OUT_OF_MEMORY, "hipMalloc");
  }
  std::vector<float> hs(n);
  int                hd = 0;
  e                     = launch_selftest(dd, ds, nullptr);
  if (e == hipSuccess)
    e = hipMemcpy(&hd, dd, sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(hs.data(), ds, sizeof(float) * n, hipMemcpyDeviceToHost);
  hipFree(dd);
  hipFree(ds);
  if (e != hipSuccess)
    return fail(DTK_ERR_HIP, std::string("selftest: ") + hipGetErrorString(e));
  int bad = 0;
  for (int i = 0; i < n; ++i)
  {
    const float ref = std::sqrt((float)i);
    uint32_t    a, b;
    std::memcpy(&a, &hs[i], 4);
    std::memcpy(&b, &ref, 4);
    bad += a != b;
  }
  out3[0] = hd;
  out3[1] = bad;
  out3[2] = dtk_device_cu_count();
  return DTK_OK;
}

} // extern "C"
