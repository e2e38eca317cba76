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
      d = b.d0;
    else
      d = b.table[b.tile_img[t]];
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

  // One wave per tile: D0 of the owned pixels, tile min / max of D0.
  // D0 = v * (mask == 0) for a uint8 mask (mask != 0 marks a seed), or
  // nu * (1 - s) for a soft float mask s in [0, 1].
  __global__ void __launch_bounds__(64) k_init(Batch b, KParams p, Queue q)
  {
    const int t    = blockIdx.x;
    const int lane = threadIdx.x;
    ImgDesc   d;
    int       tx, ty;
    resolve_tile(b, t, d, tx, ty);
    const int  RH  = p.rh;
    const int  X0  = tx * kTileW - 1;
    const int  Y0  = ty * (RH - 2) - 1;
    const int  gx  = X0 + lane;
    const bool own = lane >= 1 && lane <= kTileW && gx < d.W;
    float      mn = DTK_INF, mx = -DTK_INF;
    uint32_t   err = 0;
    if (own)
    {
      for (int r = 1; r < RH - 1; ++r)
      {
        const int gy = Y0 + r;
        if (gy >= d.H)
          break;
        float d0;
        if (p.mask_kind == MASK_U8)
        {
          const unsigned char m = d.mask[(long long)gy * d.mask_pitch + gx];
          d0                    = p.v * (float)(m == 0);
        }
        else
        {
          const float s = reinterpret_cast<const float*>(d.mask + (long long)gy * d.mask_pitch)[gx];
          if (!(s >= 0.f && s <= 1.f))
            err |= ERR_MASK_RANGE;
          d0 = p.v * (1.f - s);
        }
        reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d.D) + (long long)gy * d.D_pitch)[gx] = d0;
        mn = fminf(mn, d0);
        mx = fmaxf(mx, d0);
      }
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if (err)
      __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lane == 0 && q.tmin)
    {
      q.tmin[t] = mn;
      q.tmax[t] = mx;
    }
    // Queue reset for the next launches of this call.
    if (q.slots)
    {
      const uint32_t cap = q.cap_mask + 1u;
      for (uint32_t k = (uint32_t)t * 64u + lane; k < cap; k += gridDim.x * 64u)
        q.slots[k] = (unsigned long long)(2u * k) << 32;
      if (lane == 0)
        q.state[t] = 0;
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
      q.state[t] = TILE_QUEUED;
      __hip_atomic_fetch_add(&q.ctrl[CTRL_PENDING], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t ticket = __hip_atomic_fetch_add(&q.ctrl[CTRL_TAIL], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      q.slots[ticket & q.cap_mask] = ((unsigned long long)(2u * ticket + 1u) << 32) | (uint32_t)t;
    }
  }

  // ---- queue operations (lane 0 of a wave) --------------------------------
  // Pop: take a ticket, wait until its slot is filled or no work is pending.
  // The popped tile goes from QUEUED to RUNNING.
  __device__ int queue_pop(const Queue& q)
  {
    const uint32_t      ticket = __hip_atomic_fetch_add(&q.ctrl[CTRL_HEAD], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long* slot   = &q.slots[ticket & q.cap_mask];
    const uint32_t      want   = 2u * ticket + 1u;
    for (uint32_t spin = 0;; ++spin)
    {
      const unsigned long long s = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(s >> 32) == want)
      {
        const int t = (int)(uint32_t)s;
        __hip_atomic_store(slot, (unsigned long long)(2u * (ticket + q.cap_mask + 1u)) << 32, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // QUEUED -> RUNNING.  This RMW continues the release sequence of every
        // earlier state RMW (the wakers' and the previous visit's), so the
        // acquire fence makes their D stores visible to the loads of this visit.
        __hip_atomic_exchange(&q.state[t], (uint32_t)TILE_RUNNING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return t;
      }
      if (__hip_atomic_load(&q.ctrl[CTRL_PENDING], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
        return -1;
      if (spin >= (kSpinLimit >> 2))
      {
        __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return -1;
      }
      // Exponential backoff (up to ~16 x 512 cycles) keeps idle workers from
      // flooding the L2 with polls while others work.
      const uint32_t naps = spin < 4 ? 1u : (spin < 64 ? 4u : 16u);
      for (uint32_t i = 0; i < naps; ++i)
        __builtin_amdgcn_s_sleep(8);
    }
  }

  // Put a tile whose state is already QUEUED into the ring.
  __device__ void ring_insert(const Queue& q, int t)
  {
    __hip_atomic_fetch_add(&q.ctrl[CTRL_PENDING], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t      ticket   = __hip_atomic_fetch_add(&q.ctrl[CTRL_TAIL], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long* slot     = &q.slots[ticket & q.cap_mask];
    const uint32_t      free_seq = 2u * ticket;
    for (uint32_t spin = 0;
         (uint32_t)(__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) != free_seq; ++spin)
    {
      if (spin >= kSpinLimit)
      {
        __hip_atomic_fetch_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(slot, ((unsigned long long)(free_seq + 1u) << 32) | (uint32_t)t, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }

  // Wake a tile (caller issued a release fence after its D stores):
  //   idle    -> QUEUED and into the ring
  //   queued  -> unchanged (it will load the new halo when popped)
  //   running -> RUNNING|DIRTY (its wave re-queues it when the visit ends)
  __device__ void queue_push(const Queue& q, int t)
  {
    uint32_t s = __hip_atomic_load(&q.state[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;)
    {
      if (s & TILE_QUEUED)
        return;
      const uint32_t n = (s & TILE_RUNNING) ? (s | TILE_DIRTY) : (uint32_t)TILE_QUEUED;
      if (n == s)
        return;
      if (__hip_atomic_compare_exchange_weak(&q.state[t], &s, n, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      {
        if (n == TILE_QUEUED)
          ring_insert(q, t);
        return;
      }
    }
  }

  // End of a visit (caller issued a release fence after its D stores):
  // RUNNING -> idle, or RUNNING|DIRTY (or an unconverged visit) -> QUEUED.
  __device__ void queue_finish(const Queue& q, int t, bool requeue)
  {
    uint32_t s = __hip_atomic_load(&q.state[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;)
    {
      const uint32_t n = (requeue || (s & TILE_DIRTY)) ? (uint32_t)TILE_QUEUED : 0u;
      if (__hip_atomic_compare_exchange_weak(&q.state[t], &s, n, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      {
        if (n == TILE_QUEUED)
          ring_insert(q, t);
        return;
      }
    }
  }

  __device__ __forceinline__ void wake_neighbours(const Queue& q, const ImgDesc& d, int tx, int ty, uint32_t border)
  {
    for (int k = 0; k < 8; ++k)
    {
      if (!((border >> k) & 1u))
        continue;
      const int nx = tx + kNbrDX[k], ny = ty + kNbrDY[k];
      if (nx < 0 || ny < 0 || nx >= d.tiles_x || ny >= d.tiles_y)
        continue;
      queue_push(q, d.tile0 + ny * d.tiles_x + nx);
    }
  }

  template <int S, bool C8, bool TWO>
  __global__ void __launch_bounds__(64) k_task(Batch b, KParams p, Queue q)
  {
    __shared__ float T[(64 / S) * 65];
    __shared__ float Sv[2 * (64 / S)];
    __shared__ int   s_tile;
    uint32_t         visits = 0, sweeps = 0;
    for (;;)
    {
      if (threadIdx.x == 0)
        s_tile = queue_pop(q);
      __syncthreads();
      const int t = s_tile;
      __syncthreads();
      if (t < 0)
        break;
      ImgDesc d;
      int     tx, ty;
      resolve_tile(b, t, d, tx, ty);
      const VisitResult vr = visit_tile<S, C8, TWO>(d, p, tx, ty, T, Sv);
      ++visits;
      sweeps += (uint32_t)vr.sweeps;
      // Every lane's D stores are ordered before the release below.
      __syncthreads();
      if (threadIdx.x == 0)
      {
        if (vr.changed)
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        wake_neighbours(q, d, tx, ty, vr.border);
        queue_finish(q, t, !vr.converged);
        __hip_atomic_fetch_sub(&q.ctrl[CTRL_PENDING], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (threadIdx.x == 0 && visits)
    {
      __hip_atomic_fetch_add(&q.ctrl[CTRL_VISITS], visits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&q.ctrl[CTRL_SWEEPS], sweeps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }

  // One host-driven round: every active tile solves to its local fixed point
  // and activates the neighbours whose halo it changed.
  template <int S, bool C8, bool TWO>
  __global__ void __launch_bounds__(64) k_blocks(Batch b, KParams p, uint8_t* active, uint8_t* next, uint32_t* ctrl,
                                                 int* rounds_per_tile)
  {
    __shared__ float T[(64 / S) * 65];
    __shared__ float Sv[2 * (64 / S)];
    const int        t = blockIdx.x;
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
    const VisitResult vr = visit_tile<S, C8, TWO>(d, p, tx, ty, T, Sv);
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
        __hip_atomic_store(&ctrl[CTRL_CHANGED], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  template <int S, bool C8, bool TWO>
  static hipError_t launch_task_t(const Batch& b, const KParams& p, const Queue& q, int workers, hipStream_t s)
  {
    hipLaunchKernelGGL((k_task<S, C8, TWO>), dim3(workers), dim3(64), 0, s, b, p, q);
    return hipGetLastError();
  }

  template <int S, bool C8, bool TWO>
  static hipError_t launch_blocks_t(const Batch& b, const KParams& p, uint8_t* active, uint8_t* next, uint32_t* ctrl,
                                    int* rpt, hipStream_t s)
  {
    hipLaunchKernelGGL((k_blocks<S, C8, TWO>), dim3(b.ntiles), dim3(64), 0, s, b, p, active, next, ctrl, rpt);
    return hipGetLastError();
  }

  template <int S, bool C8, bool TWO>
  static int occupancy_t()
  {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_task<S, C8, TWO>, 64, 0) != hipSuccess)
      return 0;
    return n;
  }

  hipError_t launch_reset_ctrl(uint32_t* ctrl, hipStream_t s)
  {
    hipLaunchKernelGGL(k_reset_ctrl, dim3(1), dim3(128), 0, s, ctrl);
    return hipGetLastError();
  }

  hipError_t launch_init(const Batch& b, const KParams& p, const Queue& q, hipStream_t s)
  {
    hipLaunchKernelGGL(k_init, dim3(b.ntiles), dim3(64), 0, s, b, p, q);
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

  hipError_t launch_task(int S, bool c8, bool two, const Batch& b, const KParams& p, const Queue& q, int workers,
                         hipStream_t s)
  {
    if (S == 2)
    {
      if (c8)
        return two ? launch_task_t<2, true, true>(b, p, q, workers, s) : launch_task_t<2, true, false>(b, p, q, workers, s);
      return two ? launch_task_t<2, false, true>(b, p, q, workers, s) : launch_task_t<2, false, false>(b, p, q, workers, s);
    }
    if (c8)
      return two ? launch_task_t<4, true, true>(b, p, q, workers, s) : launch_task_t<4, true, false>(b, p, q, workers, s);
    return two ? launch_task_t<4, false, true>(b, p, q, workers, s) : launch_task_t<4, false, false>(b, p, q, workers, s);
  }

  hipError_t launch_blocks(int S, bool c8, bool two, const Batch& b, const KParams& p, uint8_t* active, uint8_t* next,
                           uint32_t* ctrl, int* rpt, hipStream_t s)
  {
    if (S == 2)
    {
      if (c8)
        return two ? launch_blocks_t<2, true, true>(b, p, active, next, ctrl, rpt, s)
                   : launch_blocks_t<2, true, false>(b, p, active, next, ctrl, rpt, s);
      return two ? launch_blocks_t<2, false, true>(b, p, active, next, ctrl, rpt, s)
                 : launch_blocks_t<2, false, false>(b, p, active, next, ctrl, rpt, s);
    }
    if (c8)
      return two ? launch_blocks_t<4, true, true>(b, p, active, next, ctrl, rpt, s)
                 : launch_blocks_t<4, true, false>(b, p, active, next, ctrl, rpt, s);
    return two ? launch_blocks_t<4, false, true>(b, p, active, next, ctrl, rpt, s)
               : launch_blocks_t<4, false, false>(b, p, active, next, ctrl, rpt, s);
  }

  int task_occupancy(int S, bool c8, bool two)
  {
    if (S == 2)
    {
      if (c8)
        return two ? occupancy_t<2, true, true>() : occupancy_t<2, true, false>();
      return two ? occupancy_t<2, false, true>() : occupancy_t<2, false, false>();
    }
    if (c8)
      return two ? occupancy_t<4, true, true>() : occupancy_t<4, true, false>();
    return two ? occupancy_t<4, false, true>() : occupancy_t<4, false, false>();
  }
} // namespace dtk
