// Command-line GGDT of a grey PNG (bin/distance_transform.cpp:27-71 of the
// reference): reads the image and an optional u8 mask (default: one seed at
// the centre), runs the transform on the GPU, normalises the map to uint8,
// colours it with inferno and writes an RGB PNG.
//
//   dt_distance_transform input_filename lambda output_filename [mask_filename[u8]]
//
// Same arguments, checks and messages as the reference.  Differences: the
// transform runs the task schedule (every schedule gives the same map bit
// for bit; the reference tool calls the raster one), normalisation and
// colouring run on the device (same values as the host versions), and an
// exception is reported on stderr with exit status 1 instead of terminating.
#include <dt/fill.hpp>
#include <dt/generalised_distance_transform.hpp>
#include <dt/imread.hpp>
#include <dt/imsave.hpp>
#include <dt/inferno.hpp>
#include <dt/normalize.hpp>
#include <dt/transfert.hpp>

#include <cstdint>
#include <cstdlib>
#include <exception>
#include <iostream>

int main(int argc, char* argv[])
{
  if (argc < 4)
  {
    std::cerr << "Usage: " << argv[0] << " input_filename lambda output_filename [mask_filename[u8]]\n";
    return 1;
  }
  try
  {
    auto                      img  = dt::imread<std::uint8_t>(argv[1]);
    dt::image2d<std::uint8_t> mask = dt::image2d<std::uint8_t>(img.width(), img.height());
    if (argc > 4)
    {
      const auto temp_mask = dt::imread<std::uint8_t>(argv[4]);
      if (temp_mask.width() != img.width() || temp_mask.height() != img.height())
      {
        std::cerr << "Invalid mask shape (Got (w: " << temp_mask.width() << " h: " << temp_mask.height()
                  << "), expected (w: " << img.width() << " h: " << img.height() << "))";
        return 1;
      }
      for (int y = 0; y < img.height(); ++y)
        for (int x = 0; x < img.width(); ++x)
          mask(x, y) = temp_mask(x, y);
    }
    else
    {
      dt::fill(mask, std::uint8_t(0));
      mask(img.width() / 2, img.height() / 2) = 1;
    }
    const float lambda = (float)std::atof(argv[2]);
    if (lambda < 0 || lambda > 1)
    {
      std::cerr << "Lambda argument must be in the range [0.0 - 1.0] (Got " << lambda << ")";
      return 1;
    }

    const auto d_img     = dt::host_to_device(img);
    const auto d_mask    = dt::host_to_device(mask);
    const auto d_dist    = dt::generalised_distance_transform_task(d_img, d_mask, lambda);
    const auto d_norm    = dt::normalize<std::uint8_t>(d_dist);
    const auto d_colored = dt::inferno(d_norm);
    const auto colored   = dt::device_to_host(d_colored);
    dt::imsave(argv[3], colored);
  }
  catch (const std::exception& e)
  {
    std::cerr << e.what() << "\n";
    return 1;
  }
  return 0;
}
