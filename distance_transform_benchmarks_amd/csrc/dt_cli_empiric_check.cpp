// dt_empiric_check: the raster and the task schedules must give the same
// map of one image.  The counterpart of the reference tool
// bin/empiric_check.cpp (same arguments, usage line, checks and messages:
// see dt_cli_common.hpp; the difference lines and the total are the
// reference's).
//
//   dt_empiric_check input_filename lambda mask_filename
//
// Exit status 0 when the maps are identical; otherwise up to five
// differing pixels and the count on stderr, exit status 1.
#include <dt/generalised_distance_transform.hpp>
#include <dt/transfert.hpp>

#include <cstring>
#include <vector>

#include "dt_cli_common.hpp"

namespace
{
  struct Mismatch
  {
    int   x, y;
    float got, want;
  };

  // Pixels where `got` and `want` (host maps of one size) differ: the count
  // and the first `keep` of them in row order.  Rows are compared whole
  // first; only a differing row is walked pixel by pixel.
  struct Comparison
  {
    long long             count = 0;
    std::vector<Mismatch> first;
  };

  Comparison compare(const dt::image2d<float>& got, const dt::image2d<float>& want, std::size_t keep)
  {
    Comparison c;
    const int  W = got.width(), H = got.height();
    for (int y = 0; y < H; ++y)
    {
      const float* g = &got(0, y);
      const float* w = &want(0, y);
      if (std::memcmp(g, w, sizeof(float) * (std::size_t)W) == 0)
        continue;
      for (int x = 0; x < W; ++x)
        if (g[x] != w[x])
        {
          if (c.first.size() < keep)
            c.first.push_back({x, y, g[x], w[x]});
          ++c.count;
        }
    }
    return c;
  }
} // namespace

int main(int argc, char* argv[])
{
  return dtcli::run(argc, argv, 4, "input_filename lambda mask_filename", [&] {
    const auto  image  = dt::imread<std::uint8_t>(argv[1]);
    const auto  seeds  = dtcli::mask_file(argv[3], image.width(), image.height());
    const float lambda = dtcli::lambda_arg(argv[2]);
    const auto  d_img  = dt::host_to_device(image);
    const auto  d_seed = dt::host_to_device(seeds);
    const auto  raster = dt::device_to_host(dt::generalised_distance_transform(d_img, d_seed, lambda));
    const auto  task   = dt::device_to_host(dt::generalised_distance_transform_task(d_img, d_seed, lambda));
    const Comparison c = compare(task, raster, 5);
    for (const Mismatch& m : c.first)
      std::cerr << "Different value at (" << m.x << ", " << m.y << "): got " << m.got << ", expected " << m.want
                << "\n";
    if (c.count == 0)
      return 0;
    std::cerr << "Total differences: " << c.count << " / " << (long long)image.width() * image.height() << "\n";
    return 1;
  });
}
