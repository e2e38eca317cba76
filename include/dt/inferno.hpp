// Inferno colour map (dt/include/dt/inferno.hpp:6-9, dt/src/inferno.cpp:263-274
// of the reference API): out(x, y) = inferno[normalized(x, y)], 256 entries.
// CPU images are mapped on the host, GPU images by a device kernel
// (dtk_inferno); the result lives where the input does.
#pragma once

#include <dt/image2d.hpp>
#include <dt/rgb.hpp>

#include <cstdint>

namespace dt
{
  image2d<rgb8> inferno(const image2d_view<std::uint8_t>& normalized);

  // The 256 colours, inferno_table()[v] for v = 0..255.
  const rgb8* inferno_table() noexcept;
} // namespace dt
