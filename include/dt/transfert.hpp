// Host <-> device pitched copies (reference: dt/include/dt/transfert.hpp:9-39,
// dt/src/transfert.cpp:10-28).  Throw std::runtime_error on failure.
#pragma once

#include <cassert>

#include <dt/image2d.hpp>

namespace dt
{
  void host_to_device(const image2d_view<void>& src, image2d_view<void>& dst);
  void device_to_host(const image2d_view<void>& src, image2d_view<void>& dst);

  template <typename T>
  image2d<T> host_to_device(const image2d_view<T>& src)
  {
    image2d<T> dst(src.width(), src.height(), e_memory_kind::GPU);
    host_to_device(src, dst);
    return dst;
  }

  template <typename T>
  image2d<T> device_to_host(const image2d_view<T>& src)
  {
    image2d<T> dst(src.width(), src.height(), e_memory_kind::CPU);
    device_to_host(src, dst);
    return dst;
  }
} // namespace dt
