// Image reading (dt/include/dt/imread.hpp:5-9 of the reference API).
// imread<std::uint8_t>(file) reads a one-component 8-bit PNG into a CPU
// image; it raises std::runtime_error ("Error while loading <file>") when the
// file cannot be read or decoded and std::invalid_argument when the image
// has more than one component.
#pragma once

#include <dt/image2d.hpp>

namespace dt
{
  template <typename T>
  image2d<T> imread(const char* filename);
} // namespace dt
