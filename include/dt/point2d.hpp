// 2-D integer point (dt/include/dt/point2d.hpp:1-98 of the reference API).
#pragma once

namespace dt
{
  class point2d
  {
  public:
    constexpr point2d() noexcept = default;
    constexpr point2d(int x, int y) noexcept
      : m_x(x)
      , m_y(y)
    {
    }

    constexpr int& x() noexcept { return m_x; }
    constexpr int  x() const noexcept { return m_x; }
    constexpr int& y() noexcept { return m_y; }
    constexpr int  y() const noexcept { return m_y; }

    constexpr bool operator==(const point2d& o) const noexcept { return m_x == o.m_x && m_y == o.m_y; }
    constexpr bool operator!=(const point2d& o) const noexcept { return !(*this == o); }

  private:
    int m_x = 0;
    int m_y = 0;
  };

  constexpr point2d operator+(const point2d& a, const point2d& b) noexcept { return {a.x() + b.x(), a.y() + b.y()}; }
  constexpr point2d operator-(const point2d& a, const point2d& b) noexcept { return {a.x() - b.x(), a.y() - b.y()}; }
  constexpr point2d operator*(const point2d& p, int v) noexcept { return {p.x() * v, p.y() * v}; }
  constexpr point2d operator*(int v, const point2d& p) noexcept { return p * v; }
} // namespace dt
