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
  size_t raster_strips(const ImgDesc& d);
  size_t raster_mail_words(const ImgDesc& d);
  // Strips (one 64-lane workgroup each) of one raster pass kernel that can be
  // resident on one CU at once; every strip of a pass must be resident.
  int        raster_occupancy(const KParams& p, bool c8);

  // dtk_adjoint.hip
  hipError_t launch_pred(const ImgDesc& d, const KParams& p, bool c8, hipStream_t s);
  hipError_t launch_pred_batched(const ImgDesc* tab, int n, int maxW, int maxH, const KParams& p, bool c8,
                                 hipStream_t s);
  hipError_t launch_backward(const BackJob* jobs, int n, int maxW, int maxH, const KParams& p, hipStream_t s);

  // Grid of the adjoint's subtree-sum kernel for n images of at most maxW x
  // maxH pixels (256-thread blocks, gx = ceil(maxW / 256) by maxH blocks per
  // image).  From eight images on it is one-dimensional, image z on XCD
  // z % 8: 8 * ceil(n / 8) * gx * maxH blocks; HIP bounds a dimension's
  // blocks times threads below 2^32, so beyond that (and below eight
  // images) the grid is gx x maxH x n.  Returns the x extent of the 1-D grid,
  // or 0 for the 3-D grid.
  inline unsigned long long subtree_grid_1d(int n, int maxW, int maxH)
  {
    if (n < 8 || maxW <= 0 || maxH <= 0)
      return 0ull;
    const unsigned long long gx = (unsigned long long)(maxW + 255) / 256ull, slots = (unsigned long long)(n + 7) / 8ull;
    const unsigned long long blocks = 8ull * slots * gx * (unsigned long long)maxH;
    return blocks * 256ull <= 0xffffffffull ? blocks : 0ull;
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
