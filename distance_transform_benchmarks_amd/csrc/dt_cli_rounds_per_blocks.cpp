// dt_rounds_per_blocks: heat map of how many rounds each tile was solved in
// by the blocks schedule, the counterpart of the reference tool
// bin/rounds_per_blocks.cu (same arguments and usage line, same output file
// rounds_per_block.png in the working directory).
//
//   dt_rounds_per_blocks img.png mask.png lmd
//
// The counts come from round_per_blocks of the blocks schedule, one per
// tile of this library's tiling (dtk_tile_size_for); the map is spread over
// the tiles (dtk_tile_image), normalised and coloured on the device
// (dtk_colorize).  The reference passes host images to the blocks
// schedule; here they are uploaded first.
#include <dt/generalised_distance_transform.hpp>
#include <dt/imsave.hpp>
#include <dt/rgb.hpp>
#include <dt/transfert.hpp>
#include <dtk.h>

#include <stdexcept>

#include "dt_cli_common.hpp"

namespace
{
  void ok(dtk_status st)
  {
    if (st != DTK_OK)
      throw std::runtime_error(dtk_last_error());
  }
  dtk_image as_dtk(const dt::image2d_view<void>& v, int dtype)
  {
    return dtk_image{const_cast<void*>(static_cast<const void*>(v.buffer())), v.width(), v.height(),
                     (long long)v.pitch(), dtype};
  }
} // namespace

int main(int argc, char* argv[])
{
  return dtcli::run(argc, argv, 4, "img.png mask.png lmd", [&] {
    const auto  image  = dt::host_to_device(dt::imread<std::uint8_t>(argv[1]));
    const float lambda = (float)std::atof(argv[3]);
    const auto  seeds  = dt::host_to_device(dt::imread<std::uint8_t>(argv[2]));
    const int   W = image.width(), H = image.height();
    dtk_params  p;
    dtk_params_default(&p);
    p.lambda   = lambda;
    p.schedule = DTK_SCHED_BLOCKS;
    int tw = 0, th = 0;
    if (!dtk_tile_size_for(&p, &tw, &th))
      throw std::invalid_argument("invalid lambda");
    const int tiles_x = (W + tw - 1) / tw, tiles_y = (H + th - 1) / th;
    // Zeroed device counters, one per tile, row-major (a one-row image is packed).
    const int        n = tiles_x * tiles_y;
    dt::image2d<int> counts(n > 0 ? n : 1, 1, dt::e_memory_kind::GPU);
    {
      dt::image2d<int> zero(n > 0 ? n : 1, 1);
      for (int k = 0; k < zero.width(); ++k)
        zero(k, 0) = 0;
      dt::host_to_device(zero, counts);
    }
    int* rounds = reinterpret_cast<int*>(counts.buffer());
    dt::generalised_distance_transform_blocks(image, seeds, lambda, 1e10f, rounds);
    dt::image2d<float>    d_map(W, H, dt::e_memory_kind::GPU);
    dt::image2d<dt::rgb8> d_rgb(W, H, dt::e_memory_kind::GPU);
    dtk_image             map = as_dtk(d_map, DTK_F32), rgb = as_dtk(d_rgb, DTK_RGB8);
    ok(dtk_tile_image(rounds, 4, tiles_x, tiles_y, tw, th, &map, nullptr));
    ok(dtk_colorize(&map, &rgb, nullptr));
    ok(dtk_sync_errors(nullptr));
    dt::imsave("rounds_per_block.png", dt::device_to_host(d_rgb));
    return 0;
  });
}
