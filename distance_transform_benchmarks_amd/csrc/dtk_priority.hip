// Tile priorities of the reference's task queue: dt::compute_block_priorities
// (dt/include/dt/block_priority.hpp:7-8, dt/src/block_priority.cu:74-155).
// The image is cut into 32 x 32 tiles (BLOCK_SIZE, dt/src/utils.cuh:11),
// gw x gh of them, N = gw * gh.
//
//   k_prio_seed   one workgroup per tile: distance 0 when any mask pixel of
//                 the tile is nonzero, else 2^31 - 1 (uint32 max / 2,
//                 block_priority.cu:84-91)
//   k_prio_solve  one workgroup for the whole grid:
//                 - city-block distance to the nearest seed tile: each row
//                   forward then backward, then each column forward then
//                   backward, d = min(d(prev) + 1, d) (block_priority.cu:37-71,
//                   94-106; exact for the L1 metric);
//                 - cdf[d] = number of tiles at distance <= d
//                   (block_priority.cu:109-115), level0 = cdf[0];
//                 - priority = cdf[d] * 64 / (N + 1) in uint32 arithmetic
//                   (block_priority.cu:118-121), in [0, 63];
//                 - offsets[k] = number of tiles of priority < k
//                   (block_priority.cu:125-148).
//
// With no seed tile at all every distance stays 2^31 - 1 and every cdf
// value is N, so each priority is N * 64 / (N + 1) and level0 is 0; the
// histogram is then skipped (the reference would size it 2^31 entries).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dtk_launch.hpp"

namespace dtk
{
  namespace
  {
    constexpr int      kPrioTile = 32;
    constexpr uint32_t kPrioFar  = 0x7fffffffu;
  } // namespace

  // Block (tx, ty) = blockIdx; 256 threads, 32 columns x 8 rows per step.
  template <bool F32>
  __global__ void __launch_bounds__(256) k_prio_seed(const void* mask, long long pitch, int W, int H, int gw,
                                                     uint32_t* dist)
  {
    const int x   = blockIdx.x * kPrioTile + (int)(threadIdx.x & 31u);
    bool      any = false;
    for (int r = (int)(threadIdx.x >> 5); r < kPrioTile; r += 8)
    {
      const int y = blockIdx.y * kPrioTile + r;
      if (x < W && y < H)
      {
        const unsigned char* row = static_cast<const unsigned char*>(mask) + (long long)y * pitch;
        if constexpr (F32)
          any |= reinterpret_cast<const float*>(row)[x] != 0.f;
        else
          any |= row[x] != 0;
      }
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0)
      dist[(size_t)blockIdx.y * gw + blockIdx.x] = any ? 0u : kPrioFar;
  }

  // One workgroup of 1024 threads.  hist: room for gw + gh words (the
  // largest finite distance is gw + gh - 2).
  __global__ void __launch_bounds__(1024) k_prio_solve(uint32_t* dist, int gw, int gh, uint32_t* hist,
                                                       uint8_t* prio, uint32_t* offsets, uint32_t* level0)
  {
    __shared__ uint32_t s_max;
    __shared__ uint32_t s_cnt[64];
    const int           tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const int           n   = gw * gh;
    if (tid == 0)
      s_max = 0u;
    if (tid < 64)
      s_cnt[tid] = 0u;

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
    const uint32_t maxd = s_max;
    // The tile grid is connected, so one seed tile makes every distance
    // finite; a far maximum means there is none.
    const bool none = maxd >= kPrioFar;
    if (!none)
    {
      for (uint32_t i = (uint32_t)tid; i <= maxd; i += (uint32_t)nt)
        hist[i] = 0u;
      __syncthreads();
      for (int i = tid; i < n; i += nt)
        atomicAdd(&hist[dist[i]], 1u);
      __syncthreads();
      if (tid == 0)
      {
        uint32_t acc = 0u;
        for (uint32_t i = 0; i <= maxd; ++i)
        {
          acc += hist[i];
          hist[i] = acc;
        }
      }
      __syncthreads();
    }

    const uint32_t total = (uint32_t)n + 1u;
    for (int i = tid; i < n; i += nt)
    {
      const uint32_t c  = none ? (uint32_t)n : hist[dist[i]];
      const uint32_t pr = c * 64u / total;
      prio[i]           = (uint8_t)pr;
      atomicAdd(&s_cnt[pr & 63u], 1u);
    }
    __syncthreads();
    if (tid == 0)
    {
      uint32_t acc = 0u;
      for (int k = 0; k < 64; ++k)
      {
        offsets[k] = acc;
        acc += s_cnt[k];
      }
      *level0 = none ? 0u : hist[0];
    }
  }

  void prio_grid(int W, int H, int* gw, int* gh)
  {
    *gw = (W + kPrioTile - 1) / kPrioTile;
    *gh = (H + kPrioTile - 1) / kPrioTile;
  }

  hipError_t launch_block_priorities(const void* mask, long long pitch, bool f32, int W, int H, uint32_t* dist,
                                     uint32_t* hist, uint8_t* prio, uint32_t* offsets, uint32_t* level0,
                                     hipStream_t s)
  {
    int gw, gh;
    prio_grid(W, H, &gw, &gh);
    if (f32)
      hipLaunchKernelGGL(k_prio_seed<true>, dim3(gw, gh), dim3(256), 0, s, mask, pitch, W, H, gw, dist);
    else
      hipLaunchKernelGGL(k_prio_seed<false>, dim3(gw, gh), dim3(256), 0, s, mask, pitch, W, H, gw, dist);
    hipLaunchKernelGGL(k_prio_solve, dim3(1), dim3(1024), 0, s, dist, gw, gh, hist, prio, offsets, level0);
    return hipGetLastError();
  }
} // namespace dtk
