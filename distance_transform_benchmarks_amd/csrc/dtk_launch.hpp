// Host-side launchers of the GGDT kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include "dtk_internal.hpp"

namespace dtk
{
  // dtk_task.hip
  hipError_t launch_reset_ctrl(uint32_t* ctrl, hipStream_t s);
  hipError_t launch_init(const Batch& b, const KParams& p, const Queue& q, hipStream_t s);
  hipError_t launch_seed(const Batch& b, const KParams& p, const Queue& q, hipStream_t s);
  hipError_t launch_blocks_seed(const Batch& b, const Queue& q, uint8_t* active, hipStream_t s);
  hipError_t launch_task(int S, bool c8, int cm, const Batch& b, const KParams& p, const Queue& q, int workers,
                         hipStream_t s);
  hipError_t launch_blocks(int S, bool c8, int cm, const Batch& b, const KParams& p, uint8_t* active, uint8_t* next,
                           uint32_t* ctrl, int* rpt, uint32_t round, hipStream_t s);
  int        task_occupancy(int S, bool c8, int cm, int nw);

  // dtk_raster.hip: one full raster iteration (four passes) with the exact
  // reference schedule; `changed` (device u32) is set when a pixel decreased.
  hipError_t launch_raster_init(const ImgDesc& d, const KParams& p, uint32_t* err, hipStream_t s);
  hipError_t launch_raster_iteration(const ImgDesc& d, const KParams& p, bool c8, uint32_t* changed,
                                     unsigned long long* mail, uint32_t pass0, hipStream_t s);
  // Strips of the longest pass (all resident at once) and the mailbox words
  // of their chunk handoff.
  size_t raster_strips(const ImgDesc& d, const KParams& p);
  size_t raster_mail_words(const ImgDesc& d, const KParams& p);
  // Strips (one 64-lane workgroup each) of one raster pass kernel that can be
  // resident on one CU at once; every strip of a pass must be resident.
  int        raster_occupancy(const KParams& p, bool c8);

  // dtk_adjoint.hip
  hipError_t launch_pred(const ImgDesc& d, const KParams& p, bool c8, hipStream_t s);
  hipError_t launch_pred_batched(const ImgDesc* tab, int n, int maxW, int maxH, const KParams& p, bool c8,
                                 hipStream_t s);
  hipError_t launch_backward(const BackJob* jobs, int n, int maxW, int maxH, const KParams& p, hipStream_t s);

  // Grid of the adjoint's tile kernels for n images of at most maxW x maxH
  // pixels: one 256-thread block per tile x tile pixels, gx x gy tiles per
  // image, images along z.  False when HIP cannot launch it (z and y at
  // most 65535 blocks; blocks times threads along x below 2^32).
  inline bool backward_tiles(int n, int maxW, int maxH, int tile, unsigned* gx, unsigned* gy)
  {
    if (n < 1 || maxW < 1 || maxH < 1 || n > 65535 || tile < 1)
      return false;
    const unsigned long long x = ((unsigned long long)maxW + (unsigned)tile - 1ull) / (unsigned)tile;
    const unsigned long long y = ((unsigned long long)maxH + (unsigned)tile - 1ull) / (unsigned)tile;
    if (y > 65535ull || x * 256ull > 0xffffffffull)
      return false;
    *gx = (unsigned)x;
    *gy = (unsigned)y;
    return true;
  }

  // dtk_selftest.hip
  hipError_t launch_selftest(int* dpp_out, float* sqrt_out, hipStream_t s);

  // dtk_color.hip: min / max into two order-preserving keys, normalisation
  // and the inferno colour map of pitched device images.
  constexpr int kMinmaxPartialWords = 2 * 4096; // partial keys of launch_minmax's blocks
  hipError_t launch_minmax(const void* img, long long pitch, int W, int H, uint32_t* keys, uint32_t* partial, int cu,
                           hipStream_t s);
  hipError_t launch_normalize_u8(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                                 const uint32_t* keys, hipStream_t s);
  hipError_t launch_inferno(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                            hipStream_t s);
  hipError_t launch_colorize(const void* in, long long in_pitch, void* out, long long out_pitch, int W, int H,
                             const uint32_t* keys, hipStream_t s);
  float      key_to_float(uint32_t k);
  // Per-tile values (int32 or uint8) spread over their tiles' pixels (f32).
  hipError_t launch_tile_image(const void* values, int value_bytes, int tiles_x, int tw, int th, void* out,
                               long long out_pitch, int W, int H, hipStream_t s);

  // dtk_priority.hip: tile priorities of 32 x 32 tiles (gw x gh of them).
  // dist: gw * gh words; hist: gw + gh words; all device buffers.
  void       prio_grid(int W, int H, int* gw, int* gh);
  hipError_t launch_block_priorities(const void* mask, long long pitch, bool f32, int W, int H, uint32_t* dist,
                                     uint32_t* hist, uint8_t* prio, uint32_t* offsets, uint32_t* level0,
                                     hipStream_t s);
} // namespace dtk
