// Mask inversion (dt/include/dt/invert_mask.hpp:5-8, dt/src/invert_mask.cpp:5-15
// of the reference API): res(x, y) = (mask(x, y) == 0), on CPU images.
#pragma once

#include <dt/image2d.hpp>

#include <cstdint>

namespace dt
{
  image2d<std::uint8_t> invert_mask(const image2d_view<std::uint8_t>& mask);
} // namespace dt
