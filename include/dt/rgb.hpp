// 3-byte RGB pixel (dt/include/dt/rgb.hpp:7-12 of the reference API).  Also
// the pixel type of the RGB8 GGDT (colour cost sqrtf(dr^2 + dg^2 + db^2)).
#pragma once

#include <cstdint>

namespace dt
{
  struct rgb8
  {
    std::uint8_t r;
    std::uint8_t g;
    std::uint8_t b;
  };
  static_assert(sizeof(rgb8) == 3, "rgb8 must be 3 bytes");
} // namespace dt
