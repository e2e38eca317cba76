// Image decoding behind dt::imread (dt_png.cpp) and the C ABI of
// include/dt_io.h.  The reference reads every input through stb_image 2.30
// (dt/src/imread.cpp:12, stbi_load(filename, &w, &h, &n, 0)) and accepts
// the result only when stb reports one component (dt/src/imread.cpp:17-21).
// These decoders return what stbi_load returns for the formats that can
// yield one 8-bit component (PNG, JPEG, PGM, TGA), and for the formats that
// never can (BMP, GIF, PSD, Softimage PIC, Radiance HDR, PPM) they report
// the component count stb would, so that imread raises the same error.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dtio
{
  struct Decoded
  {
    bool                       header_only = false; // set by the caller: size and components only
    int                        width = 0, height = 0;
    int                        components = 0; // as stb_image counts them
    std::vector<unsigned char> grey;           // width * height, filled when components == 1
  };

  // Each returns an empty string on success (components set, and the
  // samples when there is one component), else the reason it failed.
  std::string decode_png(const std::vector<unsigned char>& file, Decoded& out);
  std::string decode_jpeg(const std::vector<unsigned char>& file, Decoded& out);
  std::string decode_pnm(const std::vector<unsigned char>& file, Decoded& out);
  std::string decode_tga(const std::vector<unsigned char>& file, Decoded& out);

  // Picks the format from the file's first bytes in stb_image's order
  // (stbi__load_main: PNG, BMP, GIF, PSD, PIC, JPEG, PNM, HDR, then TGA,
  // whose header test is the weakest) and decodes it.
  std::string decode_image(const std::vector<unsigned char>& file, Decoded& out);
} // namespace dtio
