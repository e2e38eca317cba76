// dt_distance_transform: GGDT of a grey PNG, written as an inferno-coloured
// RGB PNG.  The counterpart of the reference tool bin/distance_transform.cpp
// (same arguments, usage line, checks and messages: see dt_cli_common.hpp).
//
//   dt_distance_transform input_filename lambda output_filename [mask_filename[u8]]
//
// Without a mask file the seed is the pixel (width / 2, height / 2).  The
// whole pipeline after the upload stays on the device: the converged
// transform (task schedule; every schedule gives the same map bit for bit,
// the reference tool calls the raster one), then min / max, normalisation
// to uint8 and the colour map in one pass (dtk_colorize, the same values as
// dt::normalize and dt::inferno).  Only the coloured image comes back.
#include <dt/generalised_distance_transform.hpp>
#include <dt/imsave.hpp>
#include <dt/rgb.hpp>
#include <dt/transfert.hpp>
#include <dtk.h>

#include <stdexcept>
#include <string>

#include "dt_cli_common.hpp"

namespace
{
  dtk_image as_dtk(const dt::image2d_view<void>& v, int dtype)
  {
    return dtk_image{const_cast<void*>(static_cast<const void*>(v.buffer())), v.width(), v.height(),
                     (long long)v.pitch(), dtype};
  }

  // Device map -> device inferno image (dtk_colorize), then to the host.
  dt::image2d<dt::rgb8> colour_map(const dt::image2d<float>& d_dist)
  {
    dt::image2d<dt::rgb8> d_rgb(d_dist.width(), d_dist.height(), dt::e_memory_kind::GPU);
    dtk_image             in = as_dtk(d_dist, DTK_F32), out = as_dtk(d_rgb, DTK_RGB8);
    if (dtk_colorize(&in, &out, nullptr) != DTK_OK || dtk_sync_errors(nullptr) != DTK_OK)
      throw std::runtime_error(dtk_last_error());
    return dt::device_to_host(d_rgb);
  }
} // namespace

int main(int argc, char* argv[])
{
  return dtcli::run(argc, argv, 4, "input_filename lambda output_filename [mask_filename[u8]]", [&] {
    // The reference's order: read the image, then the mask, then check lambda.
    const auto  image  = dt::imread<std::uint8_t>(argv[1]);
    const auto  seeds  = argc > 4 ? dtcli::mask_file(argv[4], image.width(), image.height())
                                  : dtcli::centre_seed(image.width(), image.height());
    const float lambda = dtcli::lambda_arg(argv[2]);
    const auto  d_dist = dt::generalised_distance_transform_task(dt::host_to_device(image), dt::host_to_device(seeds), lambda);
    dt::imsave(argv[3], colour_map(d_dist));
    return 0;
  });
}
