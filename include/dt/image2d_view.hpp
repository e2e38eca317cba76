// Non-owning pitched 2-D image views (API of dt/include/dt/image2d_view.hpp
// in the reference: e_memory_kind :11-15, image2d_view<void> :20-97,
// image2d_view<T> :99-126, addressing :194-199).
//
// address(x, y) = buffer + y * pitch + x * elem_size, pitch in bytes.
// Views never own memory; a moved-from view has a null buffer (the
// reference's move constructor, dt/src/image2d_view.cpp:28-37).
#pragma once

#include <cassert>
#include <cstdint>
#include <type_traits>
#include <utility>

#include <dt/point2d.hpp>

#if defined(__HIPCC__)
#define DT_HD __host__ __device__
#else
#define DT_HD
#endif

namespace dt
{
  enum class e_memory_kind
  {
    CPU,
    GPU
  };

  template <typename T>
  class image2d_view;

  template <>
  class image2d_view<void>
  {
  public:
    image2d_view() noexcept = default;
    image2d_view(std::uint8_t* buffer, int width, int height, int pitch, int elem_size,
                 e_memory_kind memory_kind = e_memory_kind::CPU) noexcept
      : m_buffer(buffer)
      , m_width(width)
      , m_height(height)
      , m_pitch(pitch)
      , m_elem_size(elem_size)
      , m_memory_kind(memory_kind)
    {
    }
    DT_HD image2d_view(const image2d_view<void>& o) noexcept
      : m_buffer(o.m_buffer)
      , m_width(o.m_width)
      , m_height(o.m_height)
      , m_pitch(o.m_pitch)
      , m_elem_size(o.m_elem_size)
      , m_memory_kind(o.m_memory_kind)
    {
    }
    image2d_view(image2d_view<void>&& o) noexcept
      : m_buffer(nullptr)
      , m_width(o.m_width)
      , m_height(o.m_height)
      , m_pitch(o.m_pitch)
      , m_elem_size(o.m_elem_size)
      , m_memory_kind(o.m_memory_kind)
    {
      std::swap(m_buffer, o.m_buffer);
    }
    image2d_view<void>& operator=(const image2d_view<void>& o) noexcept
    {
      m_buffer      = o.m_buffer;
      m_width       = o.m_width;
      m_height      = o.m_height;
      m_pitch       = o.m_pitch;
      m_elem_size   = o.m_elem_size;
      m_memory_kind = o.m_memory_kind;
      return *this;
    }
    image2d_view<void>& operator=(image2d_view<void>&& o) noexcept
    {
      std::swap(m_buffer, o.m_buffer);
      m_width       = o.m_width;
      m_height      = o.m_height;
      m_pitch       = o.m_pitch;
      m_elem_size   = o.m_elem_size;
      m_memory_kind = o.m_memory_kind;
      return *this;
    }

    DT_HD int           width() const noexcept { return m_width; }
    DT_HD int           height() const noexcept { return m_height; }
    DT_HD int           pitch() const noexcept { return m_pitch; }
    int                 elem_size() const noexcept { return m_elem_size; }
    std::uint8_t*       buffer() noexcept { return m_buffer; }
    const std::uint8_t* buffer() const noexcept { return m_buffer; }
    e_memory_kind       memory_kind() const noexcept { return m_memory_kind; }

    bool       valid() const noexcept { return m_buffer != nullptr; }
    DT_HD bool in_domain(int x, int y) const noexcept { return x >= 0 && y >= 0 && x < m_width && y < m_height; }
    bool       in_domain(const point2d& p) const noexcept { return in_domain(p.x(), p.y()); }

    DT_HD std::uint8_t* operator()(int x, int y) noexcept
    {
      assert(in_domain(x, y) && m_buffer);
      return m_buffer + ((long long)y * m_pitch + (long long)x * m_elem_size);
    }
    DT_HD const std::uint8_t* operator()(int x, int y) const noexcept
    {
      assert(in_domain(x, y) && m_buffer);
      return m_buffer + ((long long)y * m_pitch + (long long)x * m_elem_size);
    }
    std::uint8_t*       operator()(const point2d& p) noexcept { return (*this)(p.x(), p.y()); }
    const std::uint8_t* operator()(const point2d& p) const noexcept { return (*this)(p.x(), p.y()); }

  protected:
    std::uint8_t* m_buffer      = nullptr;
    int           m_width       = 0;
    int           m_height      = 0;
    int           m_pitch       = 0;
    int           m_elem_size   = 0;
    e_memory_kind m_memory_kind = e_memory_kind::CPU;
  };

  template <typename T>
  class image2d_view : public image2d_view<void>
  {
  public:
    image2d_view() noexcept = default;
    // T may be const (a view of a read-only buffer, as in the reference's
    // tests/io.cpp:10-11); the untyped base stores a plain byte pointer.
    image2d_view(T* buffer, int width, int height, int pitch, e_memory_kind memory_kind = e_memory_kind::CPU) noexcept
      : image2d_view<void>(reinterpret_cast<std::uint8_t*>(const_cast<std::remove_const_t<T>*>(buffer)), width, height,
                           pitch, (int)sizeof(T), memory_kind)
    {
    }
    image2d_view(const image2d_view<T>&) noexcept            = default;
    image2d_view(image2d_view<T>&&) noexcept                 = default;
    image2d_view<T>& operator=(const image2d_view<T>&) noexcept = default;
    image2d_view<T>& operator=(image2d_view<T>&&) noexcept      = default;

    DT_HD const T& operator()(int x, int y) const noexcept
    {
      return *reinterpret_cast<const T*>(image2d_view<void>::operator()(x, y));
    }
    DT_HD T& operator()(int x, int y) noexcept { return *reinterpret_cast<T*>(image2d_view<void>::operator()(x, y)); }
    const T& operator()(const point2d& p) const noexcept { return (*this)(p.x(), p.y()); }
    T&       operator()(const point2d& p) noexcept { return (*this)(p.x(), p.y()); }
  };
} // namespace dt
