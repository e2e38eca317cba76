// Schedule-versus-schedule check on one image (bin/empiric_check.cpp:13-72 of
// the reference): the raster and task schedules must give the same map.
//
//   dt_empiric_check input_filename lambda mask_filename
//
// Prints the first five differing pixels and the total on stderr and exits
// with status 1 when the maps differ, 0 when they are identical.  Same
// arguments, checks and messages as the reference; an exception is
// reported on stderr with exit status 1.
#include <dt/generalised_distance_transform.hpp>
#include <dt/imread.hpp>
#include <dt/transfert.hpp>

#include <cstdint>
#include <cstdlib>
#include <exception>
#include <iostream>

int main(int argc, char* argv[])
{
  if (argc < 4)
  {
    std::cerr << "Usage: " << argv[0] << " input_filename lambda mask_filename\n";
    return 1;
  }
  try
  {
    auto                      img  = dt::imread<std::uint8_t>(argv[1]);
    dt::image2d<std::uint8_t> mask = dt::image2d<std::uint8_t>(img.width(), img.height());

    const auto temp_mask = dt::imread<std::uint8_t>(argv[3]);
    if (temp_mask.width() != img.width() || temp_mask.height() != img.height())
    {
      std::cerr << "Invalid mask shape (Got (w: " << temp_mask.width() << " h: " << temp_mask.height()
                << "), expected (w: " << img.width() << " h: " << img.height() << "))";
      return 1;
    }
    for (int y = 0; y < img.height(); ++y)
      for (int x = 0; x < img.width(); ++x)
        mask(x, y) = temp_mask(x, y);

    const float lambda = (float)std::atof(argv[2]);
    if (lambda < 0 || lambda > 1)
    {
      std::cerr << "Lambda argument must be in the range [0.0 - 1.0] (Got " << lambda << ")";
      return 1;
    }

    const auto d_img      = dt::host_to_device(img);
    const auto d_mask     = dt::host_to_device(mask);
    const auto d_dist_ref = dt::generalised_distance_transform(d_img, d_mask, lambda);
    const auto d_dist     = dt::generalised_distance_transform_task(d_img, d_mask, lambda);
    const auto dist_ref   = dt::device_to_host(d_dist_ref);
    const auto dist       = dt::device_to_host(d_dist);

    long long diff_count = 0;
    for (int y = 0; y < img.height(); y++)
      for (int x = 0; x < img.width(); x++)
        if (dist(x, y) != dist_ref(x, y))
        {
          if (diff_count < 5)
            std::cerr << "Different value at (" << x << ", " << y << "): got " << dist(x, y) << ", expected "
                      << dist_ref(x, y) << "\n";
          diff_count++;
        }
    if (diff_count > 0)
    {
      std::cerr << "Total differences: " << diff_count << " / " << (long long)img.width() * img.height() << "\n";
      return 1;
    }
  }
  catch (const std::exception& e)
  {
    std::cerr << e.what() << "\n";
    return 1;
  }
  return 0;
}
