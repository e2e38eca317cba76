// Image writing (dt/include/dt/imsave.hpp:9-38 of the reference API): 8-bit
// grey or rgb8 CPU images to PNG; std::invalid_argument when the name does
// not end in ".png", std::runtime_error when writing fails.
#pragma once

#include <dt/image2d_view.hpp>
#include <dt/rgb.hpp>

#include <cstdint>
#include <type_traits>

namespace dt
{
  namespace impl
  {
    template <typename T>
    void imsave(const char* filename, const T* buffer, int width, int height, int pitch);
  }

  template <typename T>
  void imsave(const char* filename, const image2d_view<T>& img)
  {
    using V = std::remove_cv_t<std::remove_reference_t<T>>;
    static_assert(std::is_same<V, std::uint8_t>::value || std::is_same<V, rgb8>::value,
                  "imsave writes uint8 or rgb8 images");
    impl::imsave<V>(filename, reinterpret_cast<const V*>(img.buffer()), img.width(), img.height(), (int)img.pitch());
  }
} // namespace dt
