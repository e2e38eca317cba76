// Display of a distance map, the step after the transform in the reference
// CLI (bin/distance_transform.cpp:66-70): min / max and normalisation to
// uint8 (dt/include/dt/normalize.hpp:28-62) and the inferno colour map
// (dt/src/inferno.cpp:263-274), on device images.
//
//   k_minmax / k_minmax_final  min and max of a pitched float image as two
//                              order-preserving uint32 keys (NaNs skipped,
//                              as std::min / std::max skip them when the
//                              running value comes first; max starts at the
//                              smallest positive float, as in the
//                              reference): one partial pair per block, then
//                              one block reduces the partials (a single
//                              address takes ~12 ns per atomic, so per-wave
//                              atomics would serialise)
//   k_normalize_u8             (uint8)(double(x - min) / double(max - min) * 255)
//   k_inferno                  rgb8 = table[u8]
//   k_colorize                 both in one pass (float in, rgb8 out), four
//                              pixels per thread, three dword stores when
//                              the rows allow it, colours from a packed
//                              copy of the table in LDS
//
// The other kernels run one thread per pixel over 64 x 4 blocks, so a wave
// reads 64 consecutive pixels of one row.
#include "dtk_internal.hpp"
#include "dtk_launch.hpp"

#include <algorithm>
#include <cfloat>
#include <cstdint>

namespace dtk
{
  namespace
  {
    // 256 inferno colours, the same table as the C++ API's dt::inferno.
    __constant__ uint8_t c_inferno[256 * 3] = {
#include "dt_inferno_table.inc"
    };

    // uint32 key with the order of the floats it encodes.
    __device__ __forceinline__ uint32_t fkey(float f)
    {
      const uint32_t u = __float_as_uint(f);
      return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
    __device__ __forceinline__ float funkey(uint32_t k)
    {
      return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
    }

    __device__ __forceinline__ uint8_t norm_u8(float x, float mn, float mx)
    {
      // The same operations, in the same precisions, as the host version.
      const double denom = (double)(mx - mn);
      const double v     = (double)(x - mn) / denom * 255.0;
      return v == v ? (uint8_t)v : (uint8_t)0;
    }

    __device__ __forceinline__ const unsigned char* row(const void* base, long long pitch, int y)
    {
      return static_cast<const unsigned char*>(base) + (long long)y * pitch;
    }
  } // namespace

  constexpr int kMinmaxThreads = 256;

  // Block b covers rows [b * rows, (b + 1) * rows); each thread takes
  // columns tid, tid + 256, ... of every row, four loads in flight at a time.
  __global__ void __launch_bounds__(kMinmaxThreads) k_minmax(const void* img, long long pitch, int W, int H,
                                                              int rows, uint32_t* partial)
  {
    __shared__ uint32_t s_lo[kMinmaxThreads / 64], s_hi[kMinmaxThreads / 64];
    uint32_t            lo = 0xffffffffu, hi = 0u;
    const int           y0 = blockIdx.x * rows, y1 = min(H, y0 + rows);
    for (int y = y0; y < y1; ++y)
    {
      const float* r = reinterpret_cast<const float*>(row(img, pitch, y));
      for (int x = threadIdx.x; x < W; x += 4 * kMinmaxThreads)
      {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
          const int xx = x + k * kMinmaxThreads;
          v[k]         = xx < W ? r[xx] : __uint_as_float(0x7fc00000u); // NaN: skipped
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (v[k] == v[k])
          {
            lo = min(lo, fkey(v[k]));
            hi = max(hi, fkey(v[k]));
          }
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
      lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
      hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    if ((threadIdx.x & 63) == 0)
    {
      s_lo[threadIdx.x >> 6] = lo;
      s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
      for (int k = 1; k < kMinmaxThreads / 64; ++k)
      {
        lo = min(lo, s_lo[k]);
        hi = max(hi, s_hi[k]);
      }
      partial[2 * blockIdx.x]     = lo;
      partial[2 * blockIdx.x + 1] = hi;
    }
  }

  // keys[0] = min(key(FLT_MAX), partial lows), keys[1] = max(key(FLT_MIN), partial highs).
  __global__ void __launch_bounds__(256) k_minmax_final(const uint32_t* partial, int n, uint32_t* keys)
  {
    __shared__ uint32_t s_lo[4], s_hi[4];
    uint32_t            lo = fkey(FLT_MAX), hi = fkey(FLT_MIN);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
    {
      lo = min(lo, partial[2 * i]);
      hi = max(hi, partial[2 * i + 1]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
      lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
      hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    if ((threadIdx.x & 63) == 0)
    {
      s_lo[threadIdx.x >> 6] = lo;
      s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
      for (int k = 1; k < 4; ++k)
      {
        lo = min(lo, s_lo[k]);
        hi = max(hi, s_hi[k]);
      }
      keys[0] = lo;
      keys[1] = hi;
    }
  }

  __global__ void __launch_bounds__(256) k_normalize_u8(const void* in, long long in_pitch, void* out,
                                                        long long out_pitch, int W, int H, const uint32_t* keys)
  {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H)
      return;
    const float mn = funkey(keys[0]), mx = funkey(keys[1]);
    const float v  = reinterpret_cast<const float*>(row(in, in_pitch, y))[x];
    static_cast<uint8_t*>(out)[(long long)y * out_pitch + x] = norm_u8(v, mn, mx);
  }

  __global__ void __launch_bounds__(256) k_inferno(const void* in, long long in_pitch, void* out, long long out_pitch,
                                                   int W, int H)
  {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H)
      return;
    const int i = row(in, in_pitch, y)[x];
    uint8_t*  o = static_cast<uint8_t*>(out) + (long long)y * out_pitch + 3ll * x;
    o[0]        = c_inferno[3 * i];
    o[1]        = c_inferno[3 * i + 1];
    o[2]        = c_inferno[3 * i + 2];
  }

  // Four pixels per thread.  VEC: the input rows are 16-byte aligned and
  // the output rows 4-byte aligned, so a group of four is one 16-byte load
  // and three dword stores; otherwise (and for a row's last partial group)
  // scalar accesses.
  template <bool VEC>
  __global__ void __launch_bounds__(256) k_colorize(const void* in, long long in_pitch, void* out, long long out_pitch,
                                                    int W, int H, const uint32_t* keys)
  {
    // The colour table packed one colour per word (r | g << 8 | b << 16) in
    // LDS: one LDS read per pixel instead of three byte loads.
    __shared__ uint32_t s_lut[256];
    {
      const int t = (int)(threadIdx.y * 64 + threadIdx.x);
      s_lut[t]    = (uint32_t)c_inferno[3 * t] | ((uint32_t)c_inferno[3 * t + 1] << 8) |
                 ((uint32_t)c_inferno[3 * t + 2] << 16);
    }
    __syncthreads();
    const int x0 = (blockIdx.x * 64 + threadIdx.x) * 4;
    const int y  = blockIdx.y * 4 + threadIdx.y;
    if (x0 >= W || y >= H)
      return;
    const float    mn = funkey(keys[0]), mx = funkey(keys[1]);
    const float*   r  = reinterpret_cast<const float*>(row(in, in_pitch, y));
    uint8_t*       o  = static_cast<uint8_t*>(out) + (long long)y * out_pitch + 3ll * x0;
    if (VEC && x0 + 4 <= W)
    {
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v v = *reinterpret_cast<const f4v*>(r + x0);
      uint32_t  c[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        c[k] = s_lut[norm_u8(v[k], mn, mx)];
      uint32_t* o32 = reinterpret_cast<uint32_t*>(o);
      o32[0]        = c[0] | (c[1] << 24);
      o32[1]        = (c[1] >> 8) | (c[2] << 16);
      o32[2]        = (c[2] >> 16) | (c[3] << 8);
      return;
    }
    for (int k = 0; k < 4 && x0 + k < W; ++k)
    {
      const uint32_t c = s_lut[norm_u8(r[x0 + k], mn, mx)];
      o[3 * k]         = (uint8_t)c;
      o[3 * k + 1]     = (uint8_t)(c >> 8);
      o[3 * k + 2]     = (uint8_t)(c >> 16);
    }
  }

  // ---- launchers ------------------------------------------------------------
  static dim3 pixel_grid(int W, int H) { return dim3((unsigned)((W + 63) / 64), (unsigned)((H + 3) / 4)); }

  // partial: 2 u32 per block of k_minmax, kMinmaxMaxBlocks blocks at most.
  constexpr int kMinmaxMaxBlocks = 4096;

  hipError_t launch_minmax(const void* img, long long pitch, int W, int H, uint32_t* keys, uint32_t* partial, int cu,
                           hipStream_t s)
  {
    // About four blocks per CU, each over a band of whole rows.
    const int want   = std::max(1, std::min(kMinmaxMaxBlocks, 4 * cu));
    const int rows   = std::max(1, (H + want - 1) / want);
    const int blocks = (H + rows - 1) / rows;
    hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(kMinmaxThreads), 0, s, img, pitch, W, H, rows, partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, partial, blocks, keys);
    return hipGetLastError();
  }

  hipError_t launch_normalize_u8(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                                 const uint32_t* keys, hipStream_t s)
  {
    hipLaunchKernelGGL(k_normalize_u8, pixel_grid(W, H), dim3(64, 4), 0, s, in, in_pitch, out, out_pitch, W, H, keys);
    return hipGetLastError();
  }

  hipError_t launch_inferno(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                            hipStream_t s)
  {
    hipLaunchKernelGGL(k_inferno, pixel_grid(W, H), dim3(64, 4), 0, s, in, in_pitch, out, out_pitch, W, H);
    return hipGetLastError();
  }

  hipError_t launch_colorize(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                             const uint32_t* keys, hipStream_t s)
  {
    const dim3 grid((unsigned)((W + 255) / 256), (unsigned)((H + 3) / 4));
    const bool vec = (uintptr_t)in % 16 == 0 && in_pitch % 16 == 0 && (uintptr_t)out % 4 == 0 && out_pitch % 4 == 0;
    if (vec)
      hipLaunchKernelGGL(k_colorize<true>, grid, dim3(64, 4), 0, s, in, in_pitch, out, out_pitch, W, H, keys);
    else
      hipLaunchKernelGGL(k_colorize<false>, grid, dim3(64, 4), 0, s, in, in_pitch, out, out_pitch, W, H, keys);
    return hipGetLastError();
  }

  // Per-tile values spread over the pixels of their tiles, as float: the
  // image the reference's block_attribute_to_image builds
  // (bin/rounds_per_blocks.cu:13-23, bin/priority_block.cu:12-22), here for
  // any tile size.  Tile (tx, ty) covers x in [tx * tw, (tx + 1) * tw),
  // y in [ty * th, (ty + 1) * th); values are row-major over tiles_x tiles a
  // row, int32 (BYTES 4) or uint8 (BYTES 1).
  template <int BYTES>
  __global__ void __launch_bounds__(256) k_tile_image(const void* values, int tiles_x, int tw, int th, void* out,
                                                       long long out_pitch, int W, int H)
  {
    const int x = blockIdx.x * 64 + (int)threadIdx.x, y = blockIdx.y * 4 + (int)threadIdx.y;
    if (x >= W || y >= H)
      return;
    const long long t = (long long)(y / th) * tiles_x + x / tw;
    float           v;
    if constexpr (BYTES == 4)
      v = (float)static_cast<const int32_t*>(values)[t];
    else
      v = (float)static_cast<const uint8_t*>(values)[t];
    reinterpret_cast<float*>(static_cast<unsigned char*>(out) + (long long)y * out_pitch)[x] = v;
  }

  hipError_t launch_tile_image(const void* values, int value_bytes, int tiles_x, int tw, int th, void* out,
                               long long out_pitch, int W, int H, hipStream_t s)
  {
    if (value_bytes == 4)
      hipLaunchKernelGGL(k_tile_image<4>, pixel_grid(W, H), dim3(64, 4), 0, s, values, tiles_x, tw, th, out,
                         out_pitch, W, H);
    else
      hipLaunchKernelGGL(k_tile_image<1>, pixel_grid(W, H), dim3(64, 4), 0, s, values, tiles_x, tw, th, out,
                         out_pitch, W, H);
    return hipGetLastError();
  }

  float key_to_float(uint32_t k)
  {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float    f;
    __builtin_memcpy(&f, &u, 4);
    return f;
  }
} // namespace dtk
