// Owned pixel storage behind dt::image2d (reference:
// dt/include/dt/details/image2d_data*.hpp).  CPU storage is packed
// (pitch = width * sizeof(T)); GPU storage is pitched device memory from
// dtk_malloc_pitch (hipMallocPitch).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include <dtk.h>

namespace dt::details
{
  class image2d_data
  {
  public:
    image2d_data(const image2d_data&)            = delete;
    image2d_data& operator=(const image2d_data&) = delete;
    virtual ~image2d_data()                      = default;

    std::uint8_t*       buffer() noexcept { return m_buffer; }
    const std::uint8_t* buffer() const noexcept { return m_buffer; }
    int                 pitch() const noexcept { return m_pitch; }

  protected:
    image2d_data() = default;
    std::uint8_t* m_buffer = nullptr;
    int           m_pitch  = 0;
  };

  template <typename T>
  struct image2d_data_cpu final : public image2d_data
  {
    image2d_data_cpu(int width, int height)
    {
      const std::size_t n = (std::size_t)width * (std::size_t)height * sizeof(T);
      m_buffer            = static_cast<std::uint8_t*>(std::malloc(n > 0 ? n : 1));
      if (!m_buffer)
        throw std::runtime_error("Unable to allocate buffer");
      m_pitch = width * (int)sizeof(T);
    }
    ~image2d_data_cpu() override { std::free(m_buffer); }
  };

  template <typename T>
  struct image2d_data_gpu final : public image2d_data
  {
    image2d_data_gpu(int width, int height)
    {
      void*     p     = nullptr;
      long long pitch = 0;
      if (dtk_malloc_pitch(&p, &pitch, (long long)width * (long long)sizeof(T), height) != DTK_OK)
        throw std::runtime_error(std::string(dtk_last_error()));
      m_buffer = static_cast<std::uint8_t*>(p);
      m_pitch  = (int)pitch;
    }
    ~image2d_data_gpu() override { dtk_free(m_buffer); }
  };
} // namespace dt::details
