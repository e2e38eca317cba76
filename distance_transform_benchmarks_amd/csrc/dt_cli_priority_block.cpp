// dt_priority_block: the reference task queue's tile priorities of a mask,
// the counterpart of the reference tool bin/priority_block.cu (same
// argument, usage line and output: "Grid with N blocks", the 64 bucket
// offsets on one line, and the heat map priorities.png in the working
// directory).
//
//   dt_priority_block mask.png
//
// dt::compute_block_priorities on the device (32 x 32 tiles), then the
// priorities spread over their tiles (dtk_tile_image), normalised and
// coloured on the device (dtk_colorize).
#include <dt/block_priority.hpp>
#include <dt/imsave.hpp>
#include <dt/rgb.hpp>
#include <dt/transfert.hpp>
#include <dtk.h>

#include <iostream>
#include <stdexcept>
#include <vector>

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
  constexpr int kTile = 32; // the reference's BLOCK_SIZE (dt/src/utils.cuh:11)
} // namespace

int main(int argc, char* argv[])
{
  return dtcli::run(argc, argv, 2, "mask.png", [&] {
    const auto d_mask  = dt::host_to_device(dt::imread<std::uint8_t>(argv[1]));
    const int  W = d_mask.width(), H = d_mask.height();
    const int  tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + kTile - 1) / kTile, N = tiles_x * tiles_y;
    std::cout << "Grid with " << N << " blocks\n";
    // Device arrays: N priorities and 64 offsets (packed single-row images).
    dt::image2d<std::uint8_t>  prio(N > 0 ? N : 1, 1, dt::e_memory_kind::GPU);
    dt::image2d<std::uint32_t> offsets(64, 1, dt::e_memory_kind::GPU);
    dt::compute_block_priorities(d_mask, reinterpret_cast<std::uint8_t*>(prio.buffer()),
                                 reinterpret_cast<std::uint32_t*>(offsets.buffer()));
    const auto h_off = dt::device_to_host(offsets);
    for (int k = 0; k < 64; ++k)
      std::cout << h_off(k, 0) << (k < 63 ? " " : "\n");
    dt::image2d<float>    d_map(W, H, dt::e_memory_kind::GPU);
    dt::image2d<dt::rgb8> d_rgb(W, H, dt::e_memory_kind::GPU);
    dtk_image             map = as_dtk(d_map, DTK_F32), rgb = as_dtk(d_rgb, DTK_RGB8);
    ok(dtk_tile_image(prio.buffer(), 1, tiles_x, tiles_y, kTile, kTile, &map, nullptr));
    ok(dtk_colorize(&map, &rgb, nullptr));
    ok(dtk_sync_errors(nullptr));
    dt::imsave("priorities.png", dt::device_to_host(d_rgb));
    return 0;
  });
}
