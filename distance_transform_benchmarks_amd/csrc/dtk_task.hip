// GGDT kernels built on the wave64 tile solver (gfx950):
//   k_init    initial map D0 (dt/src/utils.cu:5-13) + per-tile min/max + queue reset
//   k_seed    initial dirty tiles -> work queue
//   k_task    persistent workers over the global dirty-tile queue
//             (replaces dt/src/generalised_distance_transform_task.cu:129-150 and
//              the priority queue of dt/src/task_queue.cuh:20-150)
//   k_blocks  one round of the host-driven block schedule
//             (dt/src/generalised_distance_transform_blocks.cu:8-112)
#include <hip/hip_runtime.h>

#include "dtk_launch.hpp"
#include "dtk_queue.hpp"
#include "dtk_tile.hpp"

namespace dtk
{
  // Image of a global tile id (a single image uses b.d0 without a table read).
  __device__ __forceinline__ void resolve_tile(const Batch& b, int t, ImgDesc& d, int& tx, int& ty)
  {
    if (b.nimg == 1)
      d = b.d0; // kernel argument: stays in scalar registers
    else
    {
      // Table read through the global address space, 8 bytes at a time.
      static_assert(sizeof(ImgDesc) % 8 == 0, "ImgDesc must be a whole number of 8-byte words");
      const uint32_t     i = gload<uint32_t>(b.tile_img, 4ll * t);
      unsigned long long w[sizeof(ImgDesc) / 8];
#pragma unroll
      for (int k = 0; k < (int)(sizeof(ImgDesc) / 8); ++k)
        w[k] = gload<unsigned long long>(b.table, (long long)i * (long long)sizeof(ImgDesc) + 8ll * k);
      __builtin_memcpy(&d, w, sizeof(ImgDesc));
    }
    const int local = t - d.tile0;
    tx              = local % d.tiles_x;
    ty              = local / d.tiles_x;
  }

  __device__ __forceinline__ float wave_min(float v)
  {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      v = fminf(v, __shfl_xor(v, o));
    return v;
  }
  __device__ __forceinline__ float wave_max(float v)
  {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      v = fmaxf(v, __shfl_xor(v, o));
    return v;
  }

  // Tiles queued or running (the kernel's first test: nothing seeded means
  // no publish will ever raise DONE).
  __device__ __forceinline__ uint32_t pending_now(const Queue& q) { return (uint32_t)(at_load(tp_word(q)) >> 32); }
  // k_init: one workgroup of kInitWaves waves per tile.  Each wave takes a
  // contiguous share of the tile's owned rows in groups of kInitRows rows
  // whose loads are all issued before the first use, so a tile costs a few
  // memory latencies instead of one per row.
  constexpr int kInitWaves = 4;
  constexpr int kInitRows  = 8;

  // D0 of rows r0 .. r0+n-1 of the tile (n <= kInitRows, wave-uniform).
  // D0 = v * (mask == 0) for a uint8 mask (mask != 0 marks a seed), or
  // nu * (1 - s) for a soft float mask s in [0, 1].
  __device__ __forceinline__ void init_d0_group(const ImgDesc& d, const KParams& p, int X0, int Y0, int r0, int n,
                                                int lane, float& mn, float& mx, uint32_t& err)
  {
    const int  gx  = X0 + lane;
    const bool own = lane >= 1 && lane <= kTileW && gx < d.W;
    const int  cx  = min(max(gx, 0), d.W - 1); // in-bounds address for every lane
    float      m[kInitRows];
#pragma unroll
    for (int k = 0; k < kInitRows; ++k)
    {
      const int cy = min(max(Y0 + r0 + k, 0), d.H - 1);
      m[k]         = p.mask_kind == MASK_U8 ? (float)gload<unsigned char>(d.mask, (long long)cy * d.mask_pitch + cx)
                                            : gload<float>(d.mask, (long long)cy * d.mask_pitch + 4ll * cx);
    }
#pragma unroll
    for (int k = 0; k < kInitRows; ++k)
    {
      const int  gy = Y0 + r0 + k;
      const bool ok = own && k < n && gy < d.H;
      float      d0;
      if (p.mask_kind == MASK_U8)
        d0 = p.v * (float)(m[k] == 0.f);
      else
      {
        if (ok && !(m[k] >= 0.f && m[k] <= 1.f))
          err |= ERR_MASK_RANGE;
        d0 = p.v * (1.f - m[k]);
      }
      if (ok)
      {
        reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d.D) + (long long)gy * d.D_pitch)[gx] = d0;
        mn = fminf(mn, d0);
        mx = fmaxf(mx, d0);
      }
    }
  }

  template <int KIND, bool C8>
  __device__ void init_tile_rows(const ImgDesc& d, const KParams& p, int X0, int Y0, int r_begin, int r_end, int lane,
                                 float& mn, float& mx, uint32_t& err)
  {
    for (int r0 = r_begin; r0 < r_end; r0 += kInitRows)
    {
      const int n = min(kInitRows, r_end - r0);
      init_d0_group(d, p, X0, Y0, r0, n, lane, mn, mx, err);
      if (d.cost)
        write_cost_group<KIND, C8, kInitRows>(d, p, X0, Y0, r0, n, lane);
    }
  }

  // D0 and (when the call caches them) the edge costs of the owned pixels,
  // the tile's min / max of D0, and the queue reset.
  __global__ void __launch_bounds__(64 * kInitWaves) k_init(Batch b, KParams p, Queue q)
  {
    __shared__ float s_mn[kInitWaves], s_mx[kInitWaves];
    const int        t    = blockIdx.x;
    const int        lane = threadIdx.x & 63;
    const int        wv   = wave_index();
    ImgDesc          d;
    int              tx, ty;
    resolve_tile(b, t, d, tx, ty);
    const int owned   = p.nw * (p.rh - 2); // owned rows of the tile
    const int X0      = tx * kTileW - 1;
    const int Y0      = ty * owned - 1;
    const int chunk   = (owned + kInitWaves - 1) / kInitWaves;
    const int r_begin = 1 + wv * chunk;
    const int r_end   = min(1 + owned, r_begin + chunk);
    float     mn = DTK_INF, mx = -DTK_INF;
    uint32_t  err = 0;
    switch (p.img_kind * 2 + (p.c8 ? 1 : 0))
    {
    case IMG_U8 * 2:
      init_tile_rows<IMG_U8, false>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    case IMG_U8 * 2 + 1:
      init_tile_rows<IMG_U8, true>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    case IMG_RGB8 * 2:
      init_tile_rows<IMG_RGB8, false>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    case IMG_RGB8 * 2 + 1:
      init_tile_rows<IMG_RGB8, true>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    case IMG_F32 * 2:
      init_tile_rows<IMG_F32, false>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    default:
      init_tile_rows<IMG_F32, true>(d, p, X0, Y0, r_begin, r_end, lane, mn, mx, err);
      break;
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if (lane == 0)
    {
      s_mn[wv] = mn;
      s_mx[wv] = mx;
    }
    if (err)
      __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (threadIdx.x == 0 && q.tmin)
    {
      for (int k = 1; k < kInitWaves; ++k)
      {
        mn = fminf(mn, s_mn[k]);
        mx = fmaxf(mx, s_mx[k]);
      }
      q.tmin[t] = mn;
      q.tmax[t] = mx;
    }
    // Queue reset for the next launches of this call: every ring slot free
    // for its first ticket, every tile idle with no pending key.
    const uint32_t cap    = q.cap_mask + 1u;
    const size_t   stride = (size_t)gridDim.x * blockDim.x;
    if (q.slots)
      for (size_t k = (size_t)t * blockDim.x + threadIdx.x; k < cap; k += stride)
        q.slots[k * kSlotStride] = (unsigned long long)(2u * (uint32_t)k) << 32;
    if (q.bslots)
      for (size_t k = (size_t)t * blockDim.x + threadIdx.x; k < (size_t)kBuckets * cap; k += stride)
        q.bslots[k] = (unsigned long long)(2u * ((uint32_t)k & q.cap_mask)) << 32;
    if (threadIdx.x == 0 && (q.slots || q.bslots))
    {
      q.state[(size_t)t * kStateStride] = 0;
      if (q.tkey)
        q.tkey[t] = kInfBits;
    }
  }

  // Zero the counters of a call (a kernel, not a memset, so the launch
  // sequence can be captured in a graph).
  __global__ void k_reset_ctrl(uint32_t* ctrl)
  {
    for (int k = threadIdx.x; k < CTRL_WORDS; k += blockDim.x)
      ctrl[k] = 0;
  }

  // One thread per tile: a tile is initially dirty when some pixel of its
  // 3x3 tile neighbourhood starts below the largest initial value of the
  // tile (otherwise no relaxation into the tile can succeed).  Its key is
  // that neighbourhood minimum.
  __global__ void k_seed(Batch b, Queue q, int ntiles)
  {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
      return;
    ImgDesc d;
    int     tx, ty;
    resolve_tile(b, t, d, tx, ty);
    float nmin = q.tmin[t];
    for (int k = 0; k < 8; ++k)
    {
      const int nx = tx + nbr_dx(k), ny = ty + nbr_dy(k);
      if (nx < 0 || ny < 0 || nx >= d.tiles_x || ny >= d.tiles_y)
        continue;
      nmin = fminf(nmin, q.tmin[d.tile0 + ny * d.tiles_x + nx]);
    }
    if (!(nmin < q.tmax[t]))
      return;
    q.state[(size_t)t * kStateStride] = TILE_DIRTY; // queued
    const uint32_t bucket = q.order == ORDER_KEY    ? key_bucket(__float_as_uint(nmin))
                            : q.order == ORDER_SEED ? (uint32_t)q.tprio[t]
                                                    : 0u;
    if (q.tkey && q.order == ORDER_KEY)
      q.tkey[t] = __float_as_uint(nmin);
    // The front starts at bucket 0.
    if (!deferred(q, bucket, 0u))
    {
      const uint32_t ticket = (uint32_t)at_add(tp_word(q), (unsigned long long)kOnePendingOneTicket);
      q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride] = ((unsigned long long)(2u * ticket + 1u) << 32) | (uint32_t)t;
      return;
    }
    at_add(tp_word(q), 1ull << 32); // pending only
    later_push(q, bucket, (uint32_t)t, DeviceNap{});
  }

  // ORDER_SEED: static priority of every tile, one workgroup per image: the
  // city-block distance in tiles to the nearest tile holding a seed (a pixel
  // whose initial value is below v), then priority = cdf[d] * 64 / (N + 1)
  // with cdf[d] the number of the image's N tiles at distance <= d.  The
  // steps of dt/src/block_priority.cu:74-155 on this build's tiling
  // (distances above kSeedPrioBins - 1 share the last histogram bin).
  constexpr int      kSeedPrioBins = 4096;
  constexpr uint32_t kSeedPrioFar  = 0x7fffffffu;
  __global__ void __launch_bounds__(256) k_seedprio(Batch b, KParams p, Queue q)
  {
    __shared__ uint32_t hist[kSeedPrioBins];
    __shared__ uint32_t s_max;
    ImgDesc             d;
    if (b.nimg == 1)
      d = b.d0;
    else
    {
      unsigned long long w[sizeof(ImgDesc) / 8];
#pragma unroll
      for (int k = 0; k < (int)(sizeof(ImgDesc) / 8); ++k)
        w[k] = gload<unsigned long long>(b.table, (long long)blockIdx.x * (long long)sizeof(ImgDesc) + 8ll * k);
      __builtin_memcpy(&d, w, sizeof(ImgDesc));
    }
    const int gw = d.tiles_x, gh = d.tiles_y, n = gw * gh;
    if (n <= 0)
      return;
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    uint32_t* dist = q.tkey + d.tile0;
    if (tid == 0)
      s_max = 0u;
    for (int i = tid; i < n; i += nt)
      dist[i] = q.tmin[d.tile0 + i] < p.v ? 0u : kSeedPrioFar;
    __syncthreads();
    for (int y = tid; y < gh; y += nt)
    {
      uint32_t* r = dist + (size_t)y * gw;
      for (int x = 1; x < gw; ++x)
        r[x] = min(r[x - 1] + 1u, r[x]);
      for (int x = gw - 2; x >= 0; --x)
        r[x] = min(r[x + 1] + 1u, r[x]);
    }
    __syncthreads();
    for (int x = tid; x < gw; x += nt)
    {
      for (int y = 1; y < gh; ++y)
        dist[(size_t)y * gw + x] = min(dist[(size_t)(y - 1) * gw + x] + 1u, dist[(size_t)y * gw + x]);
      for (int y = gh - 2; y >= 0; --y)
        dist[(size_t)y * gw + x] = min(dist[(size_t)(y + 1) * gw + x] + 1u, dist[(size_t)y * gw + x]);
    }
    __syncthreads();
    uint32_t m = 0u;
    for (int i = tid; i < n; i += nt)
      m = max(m, dist[i]);
    atomicMax(&s_max, m);
    __syncthreads();
    const bool none = s_max >= kSeedPrioFar; // no seed tile in this image
    const int  top  = none ? 0 : (int)min(s_max, (uint32_t)(kSeedPrioBins - 1));
    for (int i = tid; i <= top; i += nt)
      hist[i] = 0u;
    __syncthreads();
    if (!none)
      for (int i = tid; i < n; i += nt)
        atomicAdd(&hist[min(dist[i], (uint32_t)(kSeedPrioBins - 1))], 1u);
    __syncthreads();
    if (tid == 0 && !none)
    {
      uint32_t acc = 0u;
      for (int i = 0; i <= top; ++i)
      {
        acc += hist[i];
        hist[i] = acc;
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt)
    {
      const unsigned long long c = none ? (unsigned long long)n : hist[min(dist[i], (uint32_t)(kSeedPrioBins - 1))];
      q.tprio[d.tile0 + i]       = (uint8_t)(c * 64ull / (unsigned long long)(n + 1));
    }
  }

  // A worker leaving the persistent kernel copies the call's error bits (if
  // any) into the host mailbox with a plain system-scope store; the host
  // reports them at its next call on this device.  Every worker stores the
  // word it reads; the host only needs it non-zero.
  __device__ __forceinline__ void report_error(const Queue& q)
  {
    if (threadIdx.x != 0 || !q.mailbox)
      return;
    const uint32_t e = at_load(&q.ctrl[CTRL_ERROR]);
    if (e)
      __hip_atomic_store(q.mailbox, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }

  // ---- queue operations of one wave -----------------------------------------
  // Bucket orders: move the lowest non-empty later rings into the main ring
  // (all lanes of the wave; the caller holds CTRL_DRAIN): the lowest
  // non-empty bucket and the `window` buckets above it, then further
  // non-empty buckets in order while fewer tiles than `want` (the claims
  // waiting for a tile) have moved; the front moves to the last bucket
  // drained.  Lane b reads later ring b's head and tail; then each lane moves
  // one item per pass: it waits for the item's slot (its pusher fills it
  // right after reserving the ticket), frees it, and fills a main-ring slot
  // reserved for the whole pass with one atomic.  Items stay counted as
  // pending.  Returns the number of tiles moved (the same in every lane).
  static_assert(kBuckets == 64, "one wave lane per later ring");
  __device__ uint32_t drain_later(const Queue& q, uint32_t want)
  {
    const int      lane = threadIdx.x & 63;
    const uint32_t h    = at_load(&q.ctrl[CTRL_LHEAD + lane]);
    const uint32_t t    = at_load(&q.ctrl[CTRL_LTAIL + lane]);
    uint64_t       m    = __ballot((int)(t - h) > 0);
    uint32_t       moved = 0;
    if (m == 0)
      return 0;
    const int b0    = __builtin_ctzll(m);
    const int first = min(b0 + q.window, (int)kBuckets - 1);
    while (m)
    {
      const int b = __builtin_ctzll(m);
      if (b > first && moved >= want)
        break;
      m &= m - 1ull;
      if (lane == 0)
        at_store(&q.ctrl[CTRL_FRONT], (uint32_t)b);
      const uint32_t hb = (uint32_t)__shfl((int)h, b), tb = (uint32_t)__shfl((int)t, b);
      // Passes of up to kPer * 64 items, kPer per lane, their slot loads all
      // in flight together.
      constexpr int kPer = 4;
      for (uint32_t base = hb; (int)(tb - base) > 0; base += 64u * kPer)
      {
        const uint32_t n = min(64u * kPer, tb - base);
        unsigned long long word[kPer];
        int                tile[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          word[k] = (uint32_t)(lane + 64 * k) < n ? at_load(later_slot(q, (uint32_t)b, base + (uint32_t)(lane + 64 * k)))
                                                  : 0ull;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          tile[k] = (uint32_t)(lane + 64 * k) < n ? later_take(q, (uint32_t)b, base + (uint32_t)(lane + 64 * k), word[k])
                                                  : -1;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (tile[k] == -2) // not filled yet: its pusher is about to
            for (uint32_t spin = 0;; ++spin)
            {
              __builtin_amdgcn_s_sleep(1);
              tile[k] = later_poll(q, (uint32_t)b, base + (uint32_t)(lane + 64 * k));
              if (tile[k] != -2)
                break;
              if (spin >= kSlotWaitLimit)
              {
                at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
                tile[k] = -1;
                break;
              }
            }
        uint64_t ok[kPer];
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
          ok[k] = __ballot(tile[k] >= 0);
          total += (uint32_t)__popcll(ok[k]);
        }
        uint32_t mb = 0;
        if (lane == 0 && total)
          mb = (uint32_t)at_add(tp_word(q), (unsigned long long)total); // tickets only
        mb = (uint32_t)__shfl((int)mb, 0);
        uint32_t before = 0; // moved items of the earlier groups of this pass
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
          if (tile[k] >= 0)
          {
            const uint32_t ticket = mb + before + (uint32_t)__popcll(ok[k] & ((1ull << lane) - 1ull));
            ring_fill(q, &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride], ticket, (uint32_t)tile[k],
                      DeviceNap{});
          }
          before += (uint32_t)__popcll(ok[k]);
        }
        moved += total;
      }
      if (lane == 0)
        at_store(&q.ctrl[CTRL_LHEAD + b], tb);
    }
    return moved;
  }

  // Phases whose shader-clock cycles k_task sums per worker.
  enum CycPhase : int
  {
    CY_WAIT = 0, CY_VISIT, CY_PUB, CY_SETUP, CY_SWEEP, CY_STORE, CY_REL, CY_ACQ, CY_COUNT
  };

#ifndef DTK_DRAIN_SPIN
#define DTK_DRAIN_SPIN 32u
#endif
  // Pop: wait until the claimed ticket's slot is filled, or no work is
  // pending (-1), all lanes of one wave (the same result in every lane).
  // The popped tile goes from queued to RUNNING; *key receives the key of
  // its queued entry.  A worker claims its next ticket when a visit starts.
  // That reorders the work (a tile queued during the visit goes to this
  // worker after it), which cuts the visits of 4096^2 RGB8 at lambda 0.5 by
  // 9%, and the ticket atomic's latency overlaps the visit.  Each worker
  // holds at most one claim, so the ring bound (capacity > tiles + workers)
  // holds.  Bucket orders: a waiter whose ticket is not yet issued to any
  // pusher (the main ring is empty for it) drains the later rings: the
  // first waiter in line at once, any other after about 16 us.
  //   pre: the claimed tile when the last publish already took it (see
  //   fill_main), else -1; reset here.
  __device__ int pop(const Queue& q, uint32_t ticket, uint32_t* key, int& pre, uint64_t* cyc)
  {
    const int lane    = threadIdx.x & 63;
    uint64_t  t_start = 0;
    for (uint32_t spin = 0;; ++spin)
    {
      int t = pre;
      pre   = -1;
      if (t < 0)
      {
        t = -2;
        if (lane == 0)
          t = fifo_poll(q, ticket);
        t = __shfl(t, 0);
      }
      if (t >= 0)
      {
        uint32_t k = kInfBits;
        if (lane == 0)
        {
          const uint64_t a0 = __builtin_amdgcn_s_memtime();
          k                 = begin_visit(q, t);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          cyc[CY_ACQ] += __builtin_amdgcn_s_memtime() - a0;
        }
        *key = (uint32_t)__shfl((int)k, 0);
        return t;
      }
      if (t == -1)
        return -1;
      // Termination: the wave whose publish took pending to 0 raised DONE.
      // It is polled on its own line, every 8th poll, so idle waves do not
      // contend with the publishers' atomics on CTRL_TP.  (It comes before
      // the drain below, which would otherwise take every 8th poll.)
      if ((spin & 7u) == 7u && is_done(q))
        return -1;
      if (q.order != ORDER_FIFO && (spin & 3u) == 3u)
      {
        int drain = 0;
        if (lane == 0)
        {
          const uint32_t tail = (uint32_t)at_load(tp_word(q));
          if ((int)(tail - ticket) <= 0 && (tail == ticket || spin > DTK_DRAIN_SPIN))
            drain = drain_try_lock(q) ? 1 : 0;
        }
        if (__shfl(drain, 0))
        {
          // Every lane reads and writes later-ring heads: the whole wave
          // takes part in the lock's acquire and release.
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          const uint64_t d0 = __builtin_amdgcn_s_memtime();
          // Claims not yet served: tickets taken beyond the main ring's tail.
          uint32_t want = 0;
          if (lane == 0)
            want = at_load(&q.ctrl[CTRL_HEAD]) - (uint32_t)at_load(tp_word(q));
          want                 = (uint32_t)__shfl((int)want, 0);
          const uint32_t moved = drain_later(q, (int)want > 0 ? want : 1u);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          if (lane == 0)
          {
            drain_unlock(q);
            if (moved)
            {
              at_add(&q.ctrl[CTRL_DRAINS], 1u);
              at_add(&q.ctrl[CTRL_MOVED], moved);
            }
            at_add(&q.ctrl[CTRL_T_DRAIN], (uint32_t)((__builtin_amdgcn_s_memtime() - d0) >> 10));
          }
          if (moved)
            continue; // poll the own slot again at once
        }
      }
      // Watchdog on the constant 100 MHz clock: an idle wave legitimately
      // waits as long as the whole solve runs, so only a wait far beyond any
      // real solve (kIdleLimitTicks) is taken as a hang.
      if ((spin & 63u) == 0u)
      {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (spin == 0u)
          t_start = now;
        else if (now - t_start > kIdleLimitTicks)
        {
          if (lane == 0)
            at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
          return -1;
        }
      }
      // Short backoff (at most kPollNaps x 512 cycles between polls of the own
      // slot): a tile queued on a serial chain is picked up within about 0.5 us.
      const uint32_t naps = spin < 16 ? 1u : kPollNaps;
      for (uint32_t i = 0; i < naps; ++i)
        __builtin_amdgcn_s_sleep(8);
    }
  }

  // Fills the main-ring slots of the lanes with `need` (tickets the caller
  // reserved).  A slot is free for its ticket once the claimer of the ticket
  // one lap earlier has taken its tile.  Claims are taken out of order (a
  // worker claims its next ticket when a visit starts and takes it after
  // the visit), so that claimer may be this very worker, or another one
  // that is itself waiting here: while any lane waits, lane 0 takes the
  // worker's own claimed tile as soon as it is filled and keeps it in `pre`
  // for the next pop.  The oldest filled claim is then always taken, and no
  // wait can close a cycle.
  __device__ void fill_main(const Queue& q, bool need, uint32_t ticket, uint32_t tile, uint32_t claim, int& pre)
  {
    const int                lane = threadIdx.x & 63;
    unsigned long long*      slot = &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride];
    const unsigned long long full = ((unsigned long long)(2u * ticket + 1u) << 32) | tile;
    bool                     done = !need;
    for (uint32_t spin = 0;; ++spin)
    {
      if (!done)
      {
        unsigned long long expect = (unsigned long long)(2u * ticket) << 32;
        done                      = at_cas(slot, expect, full);
      }
      if (__ballot(!done) == 0)
        return;
      int p = pre;
      if (lane == 0 && p < 0)
      {
        const int t = fifo_poll(q, claim);
        p           = t >= 0 ? t : -1;
      }
      pre = __shfl(p, 0);
      if (spin >= kSlotWaitLimit)
      {
        if (!done)
          at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
        return;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }

  // Publish a finished visit, wave-wide (called by all 64 lanes after a
  // release fence): lanes 0..7 wake the neighbours whose pixels this visit
  // improved, lane 8 ends the tile's own visit (idle, or queued again when
  // woken meanwhile or stopped at its sweep cap).  Tiles that became queued
  // are inserted: into the main ring, or (bucket orders) into the later ring
  // of their bucket when it lies beyond the front read at the visit's start.
  // One atomic reserves the main-ring tickets and counts every new item as
  // pending, retiring this visit's item; the wave that takes pending to 0
  // raises DONE.  Kw: the visit's wake keys per wave.
  // claim: the ticket this worker claimed when the visit started (lane 0);
  // pre: see fill_main.
  // keep_ok: a tile woken during this visit may be kept (end_visit's keep):
  // then nothing is inserted for it, its item stays pending, and the return
  // value is true (the same in every lane) after an acquire fence that orders
  // the next visit's loads after the wakers' stores; own_key then holds the
  // absorbed wakes' key.  (Handing a newly woken idle neighbour to the same
  // worker in the same way, instead of queueing it, was measured and not
  // kept: 11% more visits on 4096^2 RGB8 and no shorter gap between visits.)
  template <int NW>
  __device__ bool publish(const Queue& q, const ImgDesc& d, int t, int tx, int ty, uint32_t border, bool requeue,
                          uint32_t& own_key, uint32_t front, const uint32_t* Kw, uint32_t claim, int& pre,
                          bool keep_ok)
  {
    const int lane   = threadIdx.x & 63;
    int       target = -1;
    bool      insert = false;
    bool      kept   = false;
    uint32_t  nkey   = kInfBits;
    uint32_t  bucket = 0;
    if (lane < 8)
    {
      const int nx = tx + nbr_dx(lane), ny = ty + nbr_dy(lane);
      if (((border >> lane) & 1u) && nx >= 0 && ny >= 0 && nx < d.tiles_x && ny < d.tiles_y)
      {
        target      = d.tile0 + ny * d.tiles_x + nx;
        uint32_t kb = kInfBits; // wake keys are formed for the KEY order only
        if (q.order == ORDER_KEY)
        {
          kb = Kw[lane];
#pragma unroll
          for (int w = 1; w < NW; ++w)
            kb = min(kb, Kw[8 * w + lane]);
        }
        insert = wake_tile(q, target, kb, &bucket);
      }
    }
    else if (lane == 8)
    {
      target = t;
      insert = end_visit(q, t, requeue, own_key, &bucket, keep_ok, &kept, &nkey);
    }
    const bool     keep  = __shfl((int)kept, 8) != 0;
    const bool     later = insert && deferred(q, bucket, front);
    const uint64_t all   = __ballot(insert);
    const uint64_t mains = __ballot(insert && !later);
    uint32_t       base  = 0;
    if (lane == 0)
    {
      if (!keep)
      {
        const unsigned long long old = publish_count(q, (uint32_t)__popcll(mains), (uint32_t)__popcll(all));
        base                         = (uint32_t)old;
        if (all == 0u && (uint32_t)(old >> 32) == 1u)
          raise_done(q);
      }
      else if (all != 0u)
        base = (uint32_t)publish_count_keep(q, (uint32_t)__popcll(mains), (uint32_t)__popcll(all));
    }
    base = (uint32_t)__shfl((int)base, 0);
    // No insertion is issued before that count has returned: the new items
    // are pending before any of them can be popped.
    __asm__ volatile("" ::: "memory");
    const uint32_t ticket = base + (uint32_t)__popcll(mains & ((1ull << lane) - 1ull));
    fill_main(q, insert && !later, ticket, (uint32_t)target, claim, pre);
    if (later)
      later_push(q, bucket, (uint32_t)target, DeviceNap{});
    const uint64_t lat = __ballot(later);
    if (lane == 0 && lat)
      at_add(&q.ctrl[CTRL_DEFER], (uint32_t)__popcll(lat));
    if (keep)
    {
      own_key = (uint32_t)__shfl((int)nkey, 8);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    return keep;
  }

#ifdef DTK_TRACE
  // Visit trace of experimental builds (not part of the library): word 0
  // counts records; record i is words 1 + 4i .. 4 + 4i = worker, visit start
  // and end (s_memrealtime, 100 MHz, low 32 bits), tile | kept << 31.
  constexpr uint32_t kTraceRecords = 1u << 20;
  __device__ uint32_t g_trace[1 + 4 * kTraceRecords];
  __device__ __forceinline__ void trace_visit(uint32_t worker, uint64_t t0, uint64_t t1, uint32_t tile_kept)
  {
    const uint32_t i = atomicAdd(&g_trace[0], 1u);
    if (i < kTraceRecords)
    {
      g_trace[1 + 4 * i] = worker;
      g_trace[2 + 4 * i] = (uint32_t)t0;
      g_trace[3 + 4 * i] = (uint32_t)t1;
      g_trace[4 + 4 * i] = tile_kept;
    }
  }
#endif

#ifndef DTK_TASK_WAVES_PER_EU_16
#define DTK_TASK_WAVES_PER_EU_16 2
#endif
// Whether lambda-0 tiles (constant edge costs, CM == 2) are kept for a visit
// at once when woken during their own visit; their setup is nearly free, so
// keeping saves only the queue round trip.
#ifndef DTK_KEEP_CONST
#define DTK_KEEP_CONST 1
#endif
  // One workgroup of NW waves per worker; NW stacked waves solve one tile.
  // Wave 0 pops and publishes for the tile.
  template <int S, bool C8, int CM, int NW>
  __global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(S == 4 ? DTK_TASK_WAVES_PER_EU_16 : 1)))
  k_task(Batch b, KParams p, Queue q)
  {
    __shared__ float    T[NW][(64 / S) * 65];
    __shared__ float    Sv[NW][2 * (64 / S)];
    __shared__ float    X[NW > 1 ? 2 * NW * 4 * 64 : 1];
    __shared__ uint32_t F[4 * NW];
    __shared__ uint32_t Kw[8 * NW];
    __shared__ int      s_tile;
    __shared__ int      s_keep;
    const int           w    = NW > 1 ? wave_index() : 0;
    const int           lane = threadIdx.x & 63;
    uint32_t         visits = 0, sweeps = 0, kept = 0;
    // Cycles per phase, kept in LDS by thread 0 (64-bit registers here
    // would add to the spills of the visit's peak register pressure).
    __shared__ uint64_t cyc[CY_COUNT];
    if (threadIdx.x == 0)
      for (int k = 0; k < CY_COUNT; ++k)
        cyc[k] = 0;
    // Nothing seeded: no publish will ever raise DONE.
    if (pending_now(q) == 0u)
    {
      report_error(q);
      return;
    }
    uint32_t ticket = 0, own_key = kInfBits, front = 0;
    int      pre    = -1; // wave 0: claimed tile taken early by fill_main
    if (threadIdx.x == 0)
      ticket = fifo_claim(q);
    for (;;)
    {
      const uint64_t c0 = __builtin_amdgcn_s_memtime();
      if (w == 0)
      {
        const int t = pop(q, ticket, &own_key, pre, cyc);
        if (lane == 0)
          s_tile = t;
      }
      __syncthreads();
      const int t = s_tile;
      __syncthreads();
      const uint64_t c1 = __builtin_amdgcn_s_memtime();
      if (threadIdx.x == 0)
        cyc[CY_WAIT] += c1 - c0;
      if (t < 0)
        break;
      if (threadIdx.x == 0)
      {
        ticket = fifo_claim(q);
        if (q.order != ORDER_FIFO)
          front = at_load(&q.ctrl[CTRL_FRONT]);
      }
      ImgDesc d;
      int     tx, ty;
      resolve_tile(b, t, d, tx, ty);
      // Runs after each visit of tile t: accounting, then wave 0 publishes.
      // True when the tile was woken during the visit and is kept for a
      // visit at once (same in every thread).
      uint64_t cv    = c1;
      uint32_t keeps = 0;
#ifdef DTK_TRACE
      uint64_t tv0 = __builtin_amdgcn_s_memrealtime();
#endif
      auto after = [&](const VisitResult& vr) -> bool {
#ifdef DTK_TRACE
        if (threadIdx.x == 0)
          trace_visit(blockIdx.x, tv0, __builtin_amdgcn_s_memrealtime(), (uint32_t)t | (keeps ? 0x80000000u : 0u));
#endif
        ++visits;
        if (q.tile_visits && threadIdx.x == 0)
          at_add(&q.tile_visits[t], 1);
        sweeps += (uint32_t)vr.sweeps;
        if (threadIdx.x == 0)
        {
          cyc[CY_SETUP] += vr.cyc_setup;
          cyc[CY_SWEEP] += vr.cyc_sweep;
          cyc[CY_STORE] += vr.cyc_store;
        }
        // Every wave's D stores are ordered (workgroup barrier) before the
        // release below; wave 0 publishes for the whole tile.
        __syncthreads();
        const uint64_t c2 = __builtin_amdgcn_s_memtime();
        if (threadIdx.x == 0)
          cyc[CY_VISIT] += c2 - cv;
        if (w == 0)
        {
          if (vr.changed)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          if (threadIdx.x == 0)
            cyc[CY_REL] += __builtin_amdgcn_s_memtime() - c2;
          const bool k = publish<NW>(q, d, t, tx, ty, vr.border, !vr.converged, own_key,
                                     (uint32_t)__shfl((int)front, 0), Kw, ticket, pre,
                                     keeps < kMaxKeeps && (CM != 2 || DTK_KEEP_CONST));
          if (lane == 0)
            s_keep = k ? 1 : 0;
        }
        // The keep decision, and (wave 0's acquire fence) the wakers' stores,
        // reach every wave through this barrier.  (s_keep, not s_tile: wave 0
        // writes s_tile at its next pop with no barrier after this read.)
        __syncthreads();
        const bool k = s_keep != 0;
        cv           = __builtin_amdgcn_s_memtime();
        if (threadIdx.x == 0)
          cyc[CY_PUB] += cv - c2;
        keeps += k ? 1u : 0u;
#ifdef DTK_TRACE
        tv0 = __builtin_amdgcn_s_memrealtime();
#endif
        return k;
      };
      visit_tile<S, C8, CM, NW, true, true>(d, p, tx, ty, T[w], Sv[w], X, F, q.order == ORDER_KEY ? Kw : nullptr,
                                             after);
      kept += keeps;
    }
    if (threadIdx.x == 0 && visits)
    {
      at_add(&q.ctrl[CTRL_VISITS], visits);
      at_add(&q.ctrl[CTRL_SWEEPS], sweeps);
      if (kept)
        at_add(&q.ctrl[CTRL_KEPT], kept);
    }
    if (threadIdx.x == 0)
    {
      at_add(&q.ctrl[CTRL_T_WAIT], (uint32_t)(cyc[CY_WAIT] >> 10));
      at_add(&q.ctrl[CTRL_T_VISIT], (uint32_t)(cyc[CY_VISIT] >> 10));
      at_add(&q.ctrl[CTRL_T_PUB], (uint32_t)(cyc[CY_PUB] >> 10));
      at_add(&q.ctrl[CTRL_T_SETUP], (uint32_t)(cyc[CY_SETUP] >> 10));
      at_add(&q.ctrl[CTRL_T_SWEEP], (uint32_t)(cyc[CY_SWEEP] >> 10));
      at_add(&q.ctrl[CTRL_T_STORE], (uint32_t)(cyc[CY_STORE] >> 10));
      at_add(&q.ctrl[CTRL_T_REL], (uint32_t)(cyc[CY_REL] >> 10));
      at_add(&q.ctrl[CTRL_T_ACQ], (uint32_t)(cyc[CY_ACQ] >> 10));
    }
    report_error(q);
  }

  // One host-driven round: every active tile solves to its local fixed point
  // and activates the neighbours whose halo it changed.  A round that woke a
  // tile stores its number (>= 1) in CTRL_CHANGED, so the host can enqueue
  // several rounds and read back once which was the last to wake anything.
  template <int S, bool C8, int CM, int NW>
  __global__ void __launch_bounds__(64 * NW) k_blocks(Batch b, KParams p, uint8_t* active, uint8_t* next,
                                                      uint32_t* ctrl, int* rounds_per_tile, uint32_t round)
  {
    __shared__ float    T[NW][(64 / S) * 65];
    __shared__ float    Sv[NW][2 * (64 / S)];
    __shared__ float    X[NW > 1 ? 2 * NW * 4 * 64 : 1];
    __shared__ uint32_t F[4 * NW];
    const int           w = NW > 1 ? wave_index() : 0;
    const int           t = blockIdx.x;
    if (!active[t])
      return;
    __syncthreads();
    if (threadIdx.x == 0)
    {
      active[t] = 0;
      if (rounds_per_tile)
        rounds_per_tile[t] += 1;
    }
    ImgDesc d;
    int     tx, ty;
    resolve_tile(b, t, d, tx, ty);
    const VisitResult vr = visit_tile<S, C8, CM, NW>(d, p, tx, ty, T[w], Sv[w], X, F);
    if (threadIdx.x == 0)
    {
      __hip_atomic_fetch_add(&ctrl[CTRL_VISITS], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&ctrl[CTRL_SWEEPS], (uint32_t)vr.sweeps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // The next round has work iff some tile woke a neighbour.
      bool woke = false;
      for (int k = 0; k < 8; ++k)
      {
        if (!((vr.border >> k) & 1u))
          continue;
        const int nx = tx + nbr_dx(k), ny = ty + nbr_dy(k);
        if (nx < 0 || ny < 0 || nx >= d.tiles_x || ny >= d.tiles_y)
          continue;
        next[d.tile0 + ny * d.tiles_x + nx] = 1;
        woke                                = true;
      }
      if (woke)
        __hip_atomic_store(&ctrl[CTRL_CHANGED], round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }

  // First round of the block schedule: activate the initially dirty tiles.
  __global__ void k_blocks_seed(Batch b, Queue q, uint8_t* active, int ntiles)
  {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
      return;
    ImgDesc d;
    int     tx, ty;
    resolve_tile(b, t, d, tx, ty);
    float nmin = q.tmin[t];
    for (int k = 0; k < 8; ++k)
    {
      const int nx = tx + nbr_dx(k), ny = ty + nbr_dy(k);
      if (nx < 0 || ny < 0 || nx >= d.tiles_x || ny >= d.tiles_y)
        continue;
      nmin = fminf(nmin, q.tmin[d.tile0 + ny * d.tiles_x + nx]);
    }
    active[t] = nmin < q.tmax[t] ? 1 : 0;
  }

  // ---- host launchers --------------------------------------------------------
  template <int S, bool C8, int CM>
  static hipError_t launch_task_t(const Batch& b, const KParams& p, const Queue& q, int workers, hipStream_t s)
  {
    if (p.nw == 2)
    {
      hipLaunchKernelGGL((k_task<S, C8, CM, 2>), dim3(workers), dim3(128), 0, s, b, p, q);
      return hipGetLastError();
    }
    if (p.nw == 4)
    {
      hipLaunchKernelGGL((k_task<S, C8, CM, 4>), dim3(workers), dim3(256), 0, s, b, p, q);
      return hipGetLastError();
    }
    hipLaunchKernelGGL((k_task<S, C8, CM, 1>), dim3(workers), dim3(64), 0, s, b, p, q);
    return hipGetLastError();
  }

  template <int S, bool C8, int CM>
  static hipError_t launch_blocks_t(const Batch& b, const KParams& p, uint8_t* active, uint8_t* next, uint32_t* ctrl,
                                    int* rpt, uint32_t round, hipStream_t s)
  {
    if (p.nw == 4)
      hipLaunchKernelGGL((k_blocks<S, C8, CM, 4>), dim3(b.ntiles), dim3(256), 0, s, b, p, active, next, ctrl, rpt,
                         round);
    else if (p.nw == 2)
      hipLaunchKernelGGL((k_blocks<S, C8, CM, 2>), dim3(b.ntiles), dim3(128), 0, s, b, p, active, next, ctrl, rpt,
                         round);
    else
      hipLaunchKernelGGL((k_blocks<S, C8, CM, 1>), dim3(b.ntiles), dim3(64), 0, s, b, p, active, next, ctrl, rpt,
                         round);
    return hipGetLastError();
  }

  template <int S, bool C8, int CM>
  static int occupancy_t(int nw)
  {
    int n = 0;
    if (nw == 2)
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_task<S, C8, CM, 2>, 128, 0) == hipSuccess ? n : 0;
    if (nw == 4)
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_task<S, C8, CM, 4>, 256, 0) == hipSuccess ? n : 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_task<S, C8, CM, 1>, 64, 0) == hipSuccess ? n : 0;
  }

  hipError_t launch_reset_ctrl(uint32_t* ctrl, hipStream_t s)
  {
    hipLaunchKernelGGL(k_reset_ctrl, dim3(1), dim3(128), 0, s, ctrl);
    return hipGetLastError();
  }

  hipError_t launch_init(const Batch& b, const KParams& p, const Queue& q, hipStream_t s)
  {
    hipLaunchKernelGGL(k_init, dim3(b.ntiles), dim3(64 * kInitWaves), 0, s, b, p, q);
    return hipGetLastError();
  }

  hipError_t launch_seed(const Batch& b, const KParams& p, const Queue& q, hipStream_t s)
  {
    if (q.order == ORDER_SEED)
    {
      hipLaunchKernelGGL(k_seedprio, dim3(b.nimg), dim3(256), 0, s, b, p, q);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess)
        return e;
    }
    hipLaunchKernelGGL(k_seed, dim3((b.ntiles + 255) / 256), dim3(256), 0, s, b, q, b.ntiles);
    return hipGetLastError();
  }

  hipError_t launch_blocks_seed(const Batch& b, const Queue& q, uint8_t* active, hipStream_t s)
  {
    hipLaunchKernelGGL(k_blocks_seed, dim3((b.ntiles + 255) / 256), dim3(256), 0, s, b, q, active, b.ntiles);
    return hipGetLastError();
  }

  // Run F<S, C8, CM>() for the runtime (S, c8, cm); cm: 0 one add, 1 two
  // adds, 2 constant lambda-0 costs.
  template <template <int, bool, int> class F, class... A>
  static auto dispatch(int S, bool c8, int cm, A&&... a)
  {
    if (S == 2)
    {
      if (c8)
        return cm == 2 ? F<2, true, 2>::run(a...) : cm == 1 ? F<2, true, 1>::run(a...) : F<2, true, 0>::run(a...);
      return cm == 2 ? F<2, false, 2>::run(a...) : cm == 1 ? F<2, false, 1>::run(a...) : F<2, false, 0>::run(a...);
    }
    if (c8)
      return cm == 2 ? F<4, true, 2>::run(a...) : cm == 1 ? F<4, true, 1>::run(a...) : F<4, true, 0>::run(a...);
    return cm == 2 ? F<4, false, 2>::run(a...) : cm == 1 ? F<4, false, 1>::run(a...) : F<4, false, 0>::run(a...);
  }
  template <int S, bool C8, int CM>
  struct TaskLauncher
  {
    static hipError_t run(const Batch& b, const KParams& p, const Queue& q, int workers, hipStream_t s)
    {
      return launch_task_t<S, C8, CM>(b, p, q, workers, s);
    }
  };
  template <int S, bool C8, int CM>
  struct BlocksLauncher
  {
    static hipError_t run(const Batch& b, const KParams& p, uint8_t* active, uint8_t* next, uint32_t* ctrl, int* rpt,
                          uint32_t round, hipStream_t s)
    {
      return launch_blocks_t<S, C8, CM>(b, p, active, next, ctrl, rpt, round, s);
    }
  };
  template <int S, bool C8, int CM>
  struct Occupancy
  {
    static int run(int nw) { return occupancy_t<S, C8, CM>(nw); }
  };

  hipError_t launch_task(int S, bool c8, int cm, const Batch& b, const KParams& p, const Queue& q, int workers,
                         hipStream_t s)
  {
    return dispatch<TaskLauncher>(S, c8, cm, b, p, q, workers, s);
  }

  hipError_t launch_blocks(int S, bool c8, int cm, const Batch& b, const KParams& p, uint8_t* active, uint8_t* next,
                           uint32_t* ctrl, int* rpt, uint32_t round, hipStream_t s)
  {
    return dispatch<BlocksLauncher>(S, c8, cm, b, p, active, next, ctrl, rpt, round, s);
  }

  int task_occupancy(int S, bool c8, int cm, int nw) { return dispatch<Occupancy>(S, c8, cm, nw); }
} // namespace dtk

#ifdef DTK_TRACE
// Copies the visit trace to `host` (at most `words` words) and clears its
// record count; returns the words copied, or -1.
extern "C" int dtk_trace_read(uint32_t* host, int words)
{
  if (hipDeviceSynchronize() != hipSuccess)
    return -1;
  uint32_t n = 0;
  if (hipMemcpyFromSymbol(&n, HIP_SYMBOL(dtk::g_trace), 4) != hipSuccess)
    return -1;
  if (n > dtk::kTraceRecords)
    n = dtk::kTraceRecords;
  int w = (int)(1 + 4 * n);
  if (w > words)
    w = words;
  if (hipMemcpyFromSymbol(host, HIP_SYMBOL(dtk::g_trace), (size_t)w * 4) != hipSuccess)
    return -1;
  const uint32_t zero = 0;
  if (hipMemcpyToSymbol(HIP_SYMBOL(dtk::g_trace), &zero, 4) != hipSuccess)
    return -1;
  return w;
}
#endif
