// Min / max and normalisation (dt/include/dt/normalize.hpp:11-70 of the
// reference API), for float images normalised to uint8 (what the reference
// CLI does with a distance map):
//   min starts at the type's max and max at numeric_limits<T>::min() (the
//   smallest positive float for T = float, as in the reference);
//   out = (uint8) (double(in - min) / double(max - min) * 255), truncated,
//   with the subtractions rounded in T.
// The reference leaves a 0/0 quotient (a constant image) undefined; here it
// gives 0.  CPU images are normalised on the host, GPU images by device
// kernels (dtk_minmax_f32 + dtk_normalize_u8); the result lives where the
// input does.
#pragma once

#include <dt/image2d.hpp>

#include <cstdint>
#include <utility>

namespace dt
{
  std::pair<float, float> minmax(const image2d_view<float>& in);

  void normalize(const image2d_view<float>& in, image2d_view<std::uint8_t>& out);

  template <typename O>
  image2d<O> normalize(const image2d_view<float>& in);

  template <>
  image2d<std::uint8_t> normalize<std::uint8_t>(const image2d_view<float>& in);
} // namespace dt
