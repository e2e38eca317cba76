// Owning image (reference: dt/include/dt/image2d.hpp:13-49).  Copies share
// the storage (std::shared_ptr), like the reference.
#pragma once

#include <memory>
#include <stdexcept>

#include <dt/details/image2d_data.hpp>
#include <dt/image2d_view.hpp>

namespace dt
{
  template <typename T>
  class image2d final : public image2d_view<T>
  {
  public:
    image2d()                          = default;
    image2d(const image2d&)            = default;
    image2d(image2d&&)                 = default;
    image2d& operator=(const image2d&) = default;
    image2d& operator=(image2d&&)      = default;

    image2d(int width, int height, e_memory_kind memory_kind = e_memory_kind::CPU)
      : image2d_view<T>(nullptr, width, height, 0, memory_kind)
    {
      switch (memory_kind)
      {
      case e_memory_kind::CPU:
        m_data = std::make_shared<details::image2d_data_cpu<T>>(width, height);
        break;
      case e_memory_kind::GPU:
        m_data = std::make_shared<details::image2d_data_gpu<T>>(width, height);
        break;
      default:
        throw std::invalid_argument("Invalid memory kind for image2d allocation");
      }
      this->m_buffer = m_data->buffer();
      this->m_pitch  = m_data->pitch();
    }

  private:
    std::shared_ptr<details::image2d_data> m_data;
  };
} // namespace dt
