#pragma once

// Tile priorities of the reference's task queue
// (dt/include/dt/block_priority.hpp:7-8, dt/src/block_priority.cu:74-155):
// same name, arguments and result.  32 x 32 tiles, N = ceil(W/32) *
// ceil(H/32); `priorities` (N bytes) and `cdf` (64 words, the offset of each
// priority bucket) are device buffers the caller allocates; the result is
// the number of tiles that hold a seed.  The mask must be a GPU image
// (std::invalid_argument otherwise).  The task schedule of this library
// does not need the priorities: its queue starts from the tiles next to the
// seeds and grows outward from them.

#include <cstdint>

#include <dt/image2d_view.hpp>

namespace dt
{
  std::uint32_t compute_block_priorities(const dt::image2d_view<std::uint8_t>& mask, std::uint8_t* priorities,
                                         std::uint32_t* cdf);
} // namespace dt
