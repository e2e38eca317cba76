// Generalised geodesic distance transform (GGDT), C++ API.
//
// The six declarations marked [reference] keep the signatures of
// dt/include/dt/generalised_distance_transform.hpp:7-27 (uint8 image, uint8
// seed mask with mask != 0 marking a seed, C8, lambda in [0, 1], non-seed
// value v).  They run on MI355X through the C ABI of <dtk.h>; all three
// schedules return the same, bit-exact converged map.  Like the reference
// they are synchronous.  Unlike the reference (asserts only) they check
// their preconditions and throw std::invalid_argument; device failures
// throw std::runtime_error (dt/src/generalised_distance_transform.cu:136-137).
//
// The remaining overloads are extensions: RGB8 and float32 images, soft
// float masks (D0 = nu * (1 - s)), C4 connectivity, a fixed iteration count
// (raster schedule), batched calls and predecessor output.
#pragma once

#include <cstdint>
#include <vector>

#include <dt/image2d.hpp>
#include <dt/neighborhood.hpp>
#include <dt/rgb.hpp>

namespace dt
{
  // [reference] raster schedule
  void generalised_distance_transform(const image2d_view<std::uint8_t>& img, const image2d_view<std::uint8_t>& mask,
                                      image2d_view<float>& dist, float lambda, float v = 1e10);
  image2d<float> generalised_distance_transform(const image2d_view<std::uint8_t>& img,
                                                const image2d_view<std::uint8_t>& mask, float lambda, float v = 1e10);

  // [reference] blocks schedule.  round_per_blocks (device int array, or
  // nullptr) counts visits per tile of this library's tiling, row-major:
  // 62 x 56 owned pixels by default (62 x 120 at lambda 0), so
  // dtk_tile_count_for() entries; dtk_tile_count(W, H, 0) entries (62 x 30
  // tiles) are enough for any default tiling.
  void generalised_distance_transform_blocks(const image2d_view<std::uint8_t>& img,
                                             const image2d_view<std::uint8_t>& mask, image2d_view<float>& dist,
                                             float lambda, float v = 1e10, int* round_per_blocks = nullptr);
  image2d<float> generalised_distance_transform_blocks(const image2d_view<std::uint8_t>& img,
                                                       const image2d_view<std::uint8_t>& mask, float lambda,
                                                       float v = 1e10, int* round_per_blocks = nullptr);

  // [reference] task schedule (persistent kernel over a device work queue)
  void generalised_distance_transform_task(const image2d_view<std::uint8_t>& img,
                                           const image2d_view<std::uint8_t>& mask, image2d_view<float>& D, float lambda,
                                           float v = 1e10);
  image2d<float> generalised_distance_transform_task(const image2d_view<std::uint8_t>& img,
                                                     const image2d_view<std::uint8_t>& mask, float lambda,
                                                     float v = 1e10);

  // ---- extensions -------------------------------------------------------
  enum class e_schedule
  {
    raster,
    blocks,
    task
  };

  struct ggdt_options
  {
    int        connectivity   = 8;                // 4 or 8
    e_schedule schedule       = e_schedule::task;
    int        max_iterations = 0;                // raster only: 0 = converge
    int        tile_rows      = 0;                // 0 (default), 16 or 32
    int        max_sweeps     = 0;                // per tile visit, 0 = until stable
    int        waves_per_cu   = 0;                // 0 = occupancy limit
    int*       round_per_blocks = nullptr;        // blocks schedule instrumentation
  };

  ggdt_options with_connectivity(ggdt_options o, const c4_t&);
  ggdt_options with_connectivity(ggdt_options o, const c8_t&);

  // Image pixel types: std::uint8_t, rgb8, float.  Mask types: std::uint8_t
  // (binary), float (soft, in [0, 1]).
  template <typename TI, typename TM>
  void generalised_distance_transform(const image2d_view<TI>& img, const image2d_view<TM>& mask,
                                      image2d_view<float>& dist, float lambda, float v, const ggdt_options& opt);

  template <typename TI, typename TM>
  image2d<float> generalised_distance_transform(const image2d_view<TI>& img, const image2d_view<TM>& mask,
                                                float lambda, float v, const ggdt_options& opt)
  {
    image2d<float> D(img.width(), img.height(), e_memory_kind::GPU);
    generalised_distance_transform(img, mask, D, lambda, v, opt);
    return D;
  }

  // Many images of any sizes in one persistent launch (task schedule).
  template <typename TI, typename TM>
  void generalised_distance_transform_batched(const std::vector<image2d_view<TI>>& imgs,
                                              const std::vector<image2d_view<TM>>& masks,
                                              std::vector<image2d_view<float>>& dists, float lambda, float v,
                                              const ggdt_options& opt = {});

  // Argmin predecessor codes (uint8, 0..7 in dt::c8 order, 8 = root) of a
  // converged map.
  template <typename TI, typename TM>
  void generalised_distance_transform_predecessors(const image2d_view<TI>& img, const image2d_view<TM>& mask,
                                                   const image2d_view<float>& dist, image2d_view<std::uint8_t>& pred,
                                                   float lambda, float v, const ggdt_options& opt = {});
} // namespace dt
