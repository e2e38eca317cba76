// Front end shared by the command-line tools (dt_distance_transform,
// dt_empiric_check): argument checks, input loading and error reporting.
//
// The usage lines, the order of the checks and their messages are the
// reference tools' interface (bin/distance_transform.cpp:27-61,
// bin/empiric_check.cpp:13-42), kept so that scripts written against them
// run unchanged; everything else here is this build's own.
#pragma once

#include <dt/image2d.hpp>
#include <dt/imread.hpp>

#include <charconv>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <functional>
#include <iostream>
#include <string>

namespace dtcli
{
  // A user error: the message goes to stderr as is and the tool exits 1.
  struct UserError
  {
    std::string message;
  };

  // Shortest text that reads back as the same float (what the reference's
  // std::format("{}", x) prints).
  inline std::string shortest(float v)
  {
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v);
    return std::string(buf, r.ptr);
  }

  // lambda as atof reads it, in [0, 1].
  inline float lambda_arg(const char* text)
  {
    const float lambda = (float)std::atof(text);
    if (lambda < 0 || lambda > 1)
      throw UserError{"Lambda argument must be in the range [0.0 - 1.0] (Got " + shortest(lambda) + ")"};
    return lambda;
  }

  // A uint8 mask file with the image's size (nonzero = seed).
  inline dt::image2d<std::uint8_t> mask_file(const char* path, int width, int height)
  {
    dt::image2d<std::uint8_t> m = dt::imread<std::uint8_t>(path);
    if (m.width() != width || m.height() != height)
      throw UserError{"Invalid mask shape (Got (w: " + std::to_string(m.width()) + " h: " + std::to_string(m.height()) +
                      "), expected (w: " + std::to_string(width) + " h: " + std::to_string(height) + "))"};
    return m;
  }

  // The default mask: a single seed at (width / 2, height / 2).
  inline dt::image2d<std::uint8_t> centre_seed(int width, int height)
  {
    dt::image2d<std::uint8_t> m(width, height);
    for (int y = 0; y < height; ++y)
    {
      std::uint8_t* row = &m(0, y);
      for (int x = 0; x < width; ++x)
        row[x] = 0;
    }
    if (width > 0 && height > 0)
      m(width / 2, height / 2) = 1;
    return m;
  }

  // Runs a tool: with fewer than `min_args` arguments prints
  // "Usage: <argv[0]> <usage>" and returns 1; else runs `body`, which
  // returns the exit status.  A UserError prints its message; any other
  // exception (unreadable file, device error) prints what() and a newline;
  // both exit 1.
  inline int run(int argc, char** argv, int min_args, const char* usage, const std::function<int()>& body)
  {
    if (argc < min_args)
    {
      std::cerr << "Usage: " << argv[0] << " " << usage << "\n";
      return 1;
    }
    try
    {
      return body();
    }
    catch (const UserError& e)
    {
      std::cerr << e.message;
    }
    catch (const std::exception& e)
    {
      std::cerr << e.what() << "\n";
    }
    return 1;
  }
} // namespace dtcli
