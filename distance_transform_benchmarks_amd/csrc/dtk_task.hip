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

  // tail and pending share one 64-bit word, so one atomic both reserves
  // tickets and counts the new items as pending.
  constexpr unsigned long long kOnePendingOneTicket = (1ull << 32) | 1ull;
  constexpr unsigned long long kMinusOnePending     = ~0ull << 32; // adds -1 to the high half

  __device__ __forceinline__ unsigned long long* tp_word(const Queue& q)
  {
    return reinterpret_cast<unsigned long long*>(q.ctrl + CTRL_TP);
  }
  __device__ __forceinline__ uint32_t pending_now(const Queue& q)
  {
    return (uint32_t)(__hip_atomic_load(tp_word(q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
  }
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
    const int        wv   = threadIdx.x >> 6;
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
    // Queue reset for the next launches of this call.
    if (q.slots)
    {
      const uint32_t cap = q.cap_mask + 1u;
      for (uint32_t k = (uint32_t)t * blockDim.x + threadIdx.x; k < cap; k += gridDim.x * blockDim.x)
        q.slots[(size_t)k * kSlotStride] = (unsigned long long)(2u * k) << 32;
      if (threadIdx.x == 0)
        q.state[(size_t)t * kStateStride] = 0;
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
  // tile (otherwise no relaxation into the tile can succeed).
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
      const int nx = tx + kNbrDX[k], ny = ty + kNbrDY[k];
      if (nx < 0 || ny < 0 || nx >= d.tiles_x || ny >= d.tiles_y)
        continue;
      nmin = fminf(nmin, q.tmin[d.tile0 + ny * d.tiles_x + nx]);
    }
    if (nmin < q.tmax[t])
    {
      q.state[(size_t)t * kStateStride] = TILE_DIRTY; // queued
      const uint32_t ticket = (uint32_t)__hip_atomic_fetch_add(tp_word(q), kOnePendingOneTicket, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
      q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride] = ((unsigned long long)(2u * ticket + 1u) << 32) | (uint32_t)t;
    }
  }

  // ---- queue operations (lane 0 of a wave) --------------------------------
  // Pop in two steps.  Claim: take the next ticket.  Take: wait until the
  // ticket's slot is filled or no work is pending; the popped tile goes from
  // QUEUED to RUNNING.  A worker claims its next ticket when a visit starts.
  // That reorders the work (a tile queued during the visit goes to this
  // worker after it), which cuts the visits of 4096^2 RGB8 at lambda 0.5 by
  // 9%, and the ticket atomic's latency overlaps the visit.  Each worker
  // holds at most one claim, as before, so the ring bound holds.
  __device__ uint32_t queue_claim(const Queue& q)
  {
    return __hip_atomic_fetch_add(&q.ctrl[CTRL_HEAD], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  __device__ int queue_take(const Queue& q, uint32_t ticket, uint64_t& t_acq)
  {
    unsigned long long* slot   = &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride];
    const uint32_t      want   = 2u * ticket + 1u;
    uint64_t            t_start = 0;
    for (uint32_t spin = 0;; ++spin)
    {
      const unsigned long long s = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(s >> 32) == want)
      {
        const int      t  = (int)(uint32_t)s;
        const uint64_t a0 = __builtin_amdgcn_s_memtime();
        __hip_atomic_store(slot, (unsigned long long)(2u * (ticket + q.cap_mask + 1u)) << 32, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // queued -> RUNNING.  This RMW continues the release sequence of every
        // earlier state RMW (the wakers' and the previous visit's), so the
        // acquire fence makes their D stores visible to the loads of this visit.
        __hip_atomic_exchange(&q.state[(size_t)t * kStateStride], (uint32_t)TILE_RUNNING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        t_acq += __builtin_amdgcn_s_memtime() - a0;
        return t;
      }
      // Termination: the wave whose publish took pending to 0 raised DONE.
      // It is polled on its own line, every 8th poll, so idle waves do not
      // contend with the publishers' atomics on CTRL_TP.
      if ((spin & 7u) == 7u &&
          __hip_atomic_load(&q.ctrl[CTRL_DONE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        return -1;
      // Watchdog on the constant 100 MHz clock: an idle wave legitimately
      // waits as long as the whole solve runs, so only a wait far beyond any
      // real solve (kIdleLimitTicks) is taken as a hang.
      if ((spin & 63u) == 0u)
      {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (spin == 0u)
          t_start = now;
        else if (now - t_start > kIdleLimitTicks)
          spin = kSpinLimit;
      }
      if (spin >= kSpinLimit)
      {
        __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return -1;
      }
      // Short backoff (at most kPollNaps x 512 cycles between polls of the own
      // slot): a tile queued on a serial chain is picked up within about 0.5 us.
      const uint32_t naps = spin < 16 ? 1u : kPollNaps;
      for (uint32_t i = 0; i < naps; ++i)
        __builtin_amdgcn_s_sleep(8);
    }
  }

  // Publish a finished visit, wave-wide (called by all 64 lanes after a
  // release fence): lanes 0..7 wake the neighbours whose pixels this visit
  // improved (fetch_or DIRTY: idle -> queued, running -> running+dirty),
  // lane 8 moves the tile itself to idle or, when a wake-up arrived while it
  // ran (or the visit hit its sweep cap), back to queued.  Tiles that became
  // queued go into the ring with one reservation of tickets and pending
  // count for all of them.
  __device__ void publish(const Queue& q, const ImgDesc& d, int t, int tx, int ty, uint32_t border, bool requeue)
  {
    const int lane   = threadIdx.x;
    int       target = -1;
    if (lane < 8)
    {
      const int nx = tx + kNbrDX[lane], ny = ty + kNbrDY[lane];
      if (((border >> lane) & 1u) && nx >= 0 && ny >= 0 && nx < d.tiles_x && ny < d.tiles_y)
        target = d.tile0 + ny * d.tiles_x + nx;
    }
    else if (lane == 8)
      target = t;
    bool insert = false;
    if (target >= 0)
    {
      if (lane == 8)
      {
        // RUNNING -> idle, or -> queued when woken meanwhile (DIRTY stays set)
        // or when the visit hit its sweep cap.
        const uint32_t old = __hip_atomic_fetch_and(&q.state[(size_t)target * kStateStride], ~(uint32_t)TILE_RUNNING, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old & TILE_DIRTY)
          insert = true;
        else if (requeue)
          insert = __hip_atomic_fetch_or(&q.state[(size_t)target * kStateStride], (uint32_t)TILE_DIRTY, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == 0u;
      }
      else
        insert = __hip_atomic_fetch_or(&q.state[(size_t)target * kStateStride], (uint32_t)TILE_DIRTY, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT) == 0u;
    }
    const uint64_t ins  = __ballot(insert);
    const uint32_t n    = (uint32_t)__popcll(ins);
    uint32_t       base = 0;
    if (lane == 0)
    {
      // One atomic: n tickets, n more pending items and this visit's item
      // retired.  Pending stays >= n until the new items are popped and
      // finished, and they become poppable only after their slots are
      // written below.  The wave that takes pending to 0 raises DONE.
      const unsigned long long old = __hip_atomic_fetch_add(
        tp_word(q), (unsigned long long)n * kOnePendingOneTicket + kMinusOnePending, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT);
      base = (uint32_t)old;
      if (n == 0u && (uint32_t)(old >> 32) == 1u)
        __hip_atomic_store(&q.ctrl[CTRL_DONE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    base = __shfl(base, 0);
    if (insert)
    {
      const uint32_t      ticket   = base + (uint32_t)__popcll(ins & ((1ull << lane) - 1ull));
      unsigned long long* slot     = &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride];
      const uint32_t      free_seq = 2u * ticket;
      const unsigned long long full = ((unsigned long long)(free_seq + 1u) << 32) | (uint32_t)target;
      for (uint32_t spin = 0;; ++spin)
      {
        unsigned long long expect = (unsigned long long)free_seq << 32;
        if (__hip_atomic_compare_exchange_strong(slot, &expect, full, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
          break;
        if (spin >= kSpinLimit)
        {
          __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }

#ifndef DTK_TASK_WAVES_PER_EU_16
#define DTK_TASK_WAVES_PER_EU_16 2
#endif
  // One workgroup of NW waves per worker; NW stacked waves solve one tile.
  template <int S, bool C8, int CM, int NW>
  __global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(S == 4 ? DTK_TASK_WAVES_PER_EU_16 : 1)))
  k_task(Batch b, KParams p, Queue q)
  {
    __shared__ float    T[NW][(64 / S) * 65];
    __shared__ float    Sv[NW][2 * (64 / S)];
    __shared__ float    X[NW > 1 ? 2 * NW * 4 * 64 : 1];
    __shared__ uint32_t F[4 * NW];
    __shared__ int      s_tile;
    const int           w = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
    uint32_t         visits = 0, sweeps = 0;
    uint64_t         t_wait = 0, t_visit = 0, t_pub = 0, t_setup = 0, t_sweep = 0, t_store = 0;
    uint64_t         t_rel = 0, t_acq = 0;
    // Nothing seeded: no publish will ever raise DONE.
    if (pending_now(q) == 0u)
      return;
    uint32_t ticket = 0;
    if (threadIdx.x == 0)
      ticket = queue_claim(q);
    for (;;)
    {
      const uint64_t c0 = __builtin_amdgcn_s_memtime();
      if (threadIdx.x == 0)
        s_tile = queue_take(q, ticket, t_acq);
      __syncthreads();
      const int t = s_tile;
      __syncthreads();
      const uint64_t c1 = __builtin_amdgcn_s_memtime();
      t_wait += c1 - c0;
      if (t < 0)
        break;
      if (threadIdx.x == 0)
        ticket = queue_claim(q);
      ImgDesc d;
      int     tx, ty;
      resolve_tile(b, t, d, tx, ty);
      const VisitResult vr = visit_tile<S, C8, CM, NW>(d, p, tx, ty, T[w], Sv[w], X, F);
      ++visits;
      if (q.tile_visits && threadIdx.x == 0)
        __hip_atomic_fetch_add(&q.tile_visits[t], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sweeps += (uint32_t)vr.sweeps;
      t_setup += vr.cyc_setup;
      t_sweep += vr.cyc_sweep;
      t_store += vr.cyc_store;
      // Every wave's D stores are ordered (workgroup barrier) before the
      // release below; wave 0 publishes for the whole tile.
      __syncthreads();
      const uint64_t c2 = __builtin_amdgcn_s_memtime();
      t_visit += c2 - c1;
      if (w == 0)
      {
        if (vr.changed)
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        t_rel += __builtin_amdgcn_s_memtime() - c2;
        publish(q, d, t, tx, ty, vr.border, !vr.converged);
      }
      t_pub += __builtin_amdgcn_s_memtime() - c2;
    }
    if (threadIdx.x == 0 && visits)
    {
      __hip_atomic_fetch_add(&q.ctrl[CTRL_VISITS], visits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_SWEEPS], sweeps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0)
    {
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_WAIT], (uint32_t)(t_wait >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_VISIT], (uint32_t)(t_visit >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_PUB], (uint32_t)(t_pub >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_SETUP], (uint32_t)(t_setup >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_SWEEP], (uint32_t)(t_sweep >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_STORE], (uint32_t)(t_store >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_REL], (uint32_t)(t_rel >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_T_ACQ], (uint32_t)(t_acq >> 10), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
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
    const int           w = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
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
        const int nx = tx + kNbrDX[k], ny = ty + kNbrDY[k];
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
      const int nx = tx + kNbrDX[k], ny = ty + kNbrDY[k];
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

  hipError_t launch_seed(const Batch& b, const Queue& q, hipStream_t s)
  {
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
