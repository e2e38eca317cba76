// C4 / C8 neighbourhoods (dt/include/dt/c4.hpp:9-27, dt/include/dt/c8.hpp:9-31
// of the reference API, same offset orders).  dt::c4(p) / dt::c8(p) return
// the neighbours of p; the objects also select the GGDT connectivity.
#pragma once

#include <array>

#include <dt/point2d.hpp>

namespace dt
{
  class c4_t
  {
  public:
    static constexpr int connectivity = 4;
    constexpr c4_t() noexcept = default;
    constexpr std::array<point2d, 4> operator()(const point2d& p) const noexcept
    {
      return {p + point2d{-1, 0}, p + point2d{0, -1}, p + point2d{1, 0}, p + point2d{0, 1}};
    }
  };

  class c8_t
  {
  public:
    static constexpr int connectivity = 8;
    constexpr c8_t() noexcept = default;
    constexpr std::array<point2d, 8> operator()(const point2d& p) const noexcept
    {
      return {p + point2d{-1, 0}, p + point2d{-1, -1}, p + point2d{0, -1}, p + point2d{1, -1},
              p + point2d{1, 0},  p + point2d{1, 1},   p + point2d{0, 1},  p + point2d{-1, 1}};
    }
  };

  static constexpr c4_t c4 = {};
  static constexpr c8_t c8 = {};
} // namespace dt
