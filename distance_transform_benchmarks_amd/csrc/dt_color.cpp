// Display helpers of the C++ API: dt::minmax, dt::normalize, dt::inferno
// and dt::invert_mask.  CPU images are processed on the host with the
// reference's arithmetic (dt/include/dt/normalize.hpp:28-62,
// dt/src/inferno.cpp:263-274, dt/src/invert_mask.cpp:5-15); GPU images go
// through the device kernels of the C ABI (dtk_minmax_f32, dtk_normalize_u8,
// dtk_inferno), which compute the same values.  GPU calls are synchronous,
// like the rest of the C++ API.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <stdexcept>
#include <string>

#include <dt/inferno.hpp>
#include <dt/invert_mask.hpp>
#include <dt/normalize.hpp>
#include <dtk.h>

namespace dt
{
  namespace
  {
    // The table of the device kernel (dt_inferno_table.inc: r, g, b per
    // entry), here as rgb8 values.
    const rgb8 kInferno[256] = {
#include "dt_inferno_table.inc"
    };

    void check(dtk_status s, const char* what)
    {
      if (s == DTK_OK)
        return;
      const std::string msg = std::string(what) + ": " + dtk_last_error();
      if (s == DTK_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
      throw std::runtime_error(msg);
    }

    template <typename T>
    dtk_image to_dtk(const image2d_view<T>& v, int dtype)
    {
      dtk_image d;
      d.data   = const_cast<std::uint8_t*>(v.buffer());
      d.width  = v.width();
      d.height = v.height();
      d.pitch  = v.pitch();
      d.dtype  = dtype;
      return d;
    }

    // (uint8) of the reference's double expression; 0/0 (undefined in the
    // reference) gives 0.
    std::uint8_t norm_u8(float x, float mn, float mx)
    {
      const double denom = static_cast<double>(mx - mn);
      const double v     = static_cast<double>(x - mn) / denom * 255.0;
      return v == v ? static_cast<std::uint8_t>(v) : std::uint8_t(0);
    }

    void same_size(int w0, int h0, int w1, int h1)
    {
      if (w0 != w1 || h0 != h1)
        throw std::invalid_argument("normalize / inferno: input and output sizes differ");
    }
  } // namespace

  const rgb8* inferno_table() noexcept { return kInferno; }

  std::pair<float, float> minmax(const image2d_view<float>& in)
  {
    if (in.memory_kind() == e_memory_kind::GPU)
    {
      dtk_image d = to_dtk(in, DTK_F32);
      float     r[2];
      check(dtk_minmax_f32(&d, r, nullptr), "minmax");
      return {r[0], r[1]};
    }
    float mn = FLT_MAX, mx = FLT_MIN; // std::numeric_limits<float>::max() / min(), as in the reference
    for (int y = 0; y < in.height(); y++)
      for (int x = 0; x < in.width(); x++)
      {
        const float v = in(x, y);
        mn            = v < mn ? v : mn;
        mx            = mx < v ? v : mx;
      }
    return {mn, mx};
  }

  void normalize(const image2d_view<float>& in, image2d_view<std::uint8_t>& out)
  {
    same_size(in.width(), in.height(), out.width(), out.height());
    if (in.memory_kind() != out.memory_kind())
      throw std::invalid_argument("normalize: input and output must live on the same side");
    if (in.memory_kind() == e_memory_kind::GPU)
    {
      dtk_image di = to_dtk(in, DTK_F32), doo = to_dtk(out, DTK_U8);
      check(dtk_normalize_u8(&di, &doo, nullptr), "normalize");
      check(hipDeviceSynchronize() == hipSuccess ? DTK_OK : DTK_ERR_HIP, "normalize");
      return;
    }
    const auto [mn, mx] = minmax(in);
    for (int y = 0; y < in.height(); y++)
      for (int x = 0; x < in.width(); x++)
        out(x, y) = norm_u8(in(x, y), mn, mx);
  }

  template <>
  image2d<std::uint8_t> normalize<std::uint8_t>(const image2d_view<float>& in)
  {
    image2d<std::uint8_t> out(in.width(), in.height(), in.memory_kind());
    normalize(in, out);
    return out;
  }

  image2d<rgb8> inferno(const image2d_view<std::uint8_t>& normalized)
  {
    image2d<rgb8> out(normalized.width(), normalized.height(), normalized.memory_kind());
    if (normalized.memory_kind() == e_memory_kind::GPU)
    {
      dtk_image di = to_dtk(normalized, DTK_U8), doo = to_dtk(out, DTK_RGB8);
      check(dtk_inferno(&di, &doo, nullptr), "inferno");
      check(hipDeviceSynchronize() == hipSuccess ? DTK_OK : DTK_ERR_HIP, "inferno");
      return out;
    }
    for (int y = 0; y < out.height(); y++)
      for (int x = 0; x < out.width(); x++)
        out(x, y) = kInferno[normalized(x, y)];
    return out;
  }

  image2d<std::uint8_t> invert_mask(const image2d_view<std::uint8_t>& mask)
  {
    if (mask.memory_kind() != e_memory_kind::CPU)
      throw std::invalid_argument("invert_mask: CPU images only");
    image2d<std::uint8_t> res(mask.width(), mask.height());
    for (int y = 0; y < res.height(); ++y)
      for (int x = 0; x < res.width(); x++)
        res(x, y) = (mask(x, y) == 0);
    return res;
  }
} // namespace dt
