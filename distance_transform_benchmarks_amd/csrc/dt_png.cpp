// PNG reading and writing for the C++ API (dt::imread / dt::imsave) and the
// C ABI of include/dt_io.h.  The reference reads and writes images through
// the vendored stb_image / stb_image_write (dt/src/imread.cpp:9-29,
// dt/src/imsave.cpp:9-30); this is an independent codec on zlib that keeps
// the reference's observable behaviour:
//   - imread<uint8_t> accepts an image with exactly one 8-bit component, as
//     stb reports components.  PNG: grey of any bit depth (16-bit samples
//     keep their high byte, 1/2/4-bit samples are scaled to 0..255),
//     interlaced or not; grey with a tRNS chunk counts as two components;
//     RGB, RGBA, grey+alpha and palette images are rejected with the
//     reference's message.  JPEG, PGM and TGA are decoded by dt_jpeg.cpp and
//     dt_image_formats.cpp, which also recognise the formats that only ever
//     have colour.  Unreadable files raise "Error while loading <file>".
//   - imsave writes 8-bit grey or RGB, refuses a name that does not end in
//     ".png", and reads rows at the view's pitch.
// Chunk CRCs are not checked on input, as stb does not check them; they are
// written on output.
#include <dt/imread.hpp>
#include <dt/imsave.hpp>
#include <dt_io.h>

#include "dt_image_decode.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
  const unsigned char kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};

  uint32_t be32(const unsigned char* p)
  {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
  }

  void put32(std::vector<unsigned char>& v, uint32_t x)
  {
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
  }

  struct Header
  {
    uint32_t width = 0, height = 0;
    int      depth = 0, color = 0, interlace = 0;
  };

  int channels_of(int color)
  {
    switch (color)
    {
    case 0: return 1; // grey
    case 2: return 3; // RGB
    case 3: return 1; // palette index
    case 4: return 2; // grey + alpha
    case 6: return 4; // RGBA
    default: return 0;
    }
  }

  bool valid_depth(int color, int depth)
  {
    switch (color)
    {
    case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
    case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
    case 2:
    case 4:
    case 6: return depth == 8 || depth == 16;
    default: return false;
    }
  }

  unsigned char paeth(int a, int b, int c)
  {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc)
      return (unsigned char)a;
    return (unsigned char)(pb <= pc ? b : c);
  }

  // Undo the scanline filters of one (sub-)image in place.  raw holds rows of
  // 1 + rowbytes bytes; returns false on an unknown filter type.
  bool unfilter(unsigned char* raw, size_t rows, size_t rowbytes, size_t bpp)
  {
    const unsigned char* prev = nullptr;
    for (size_t y = 0; y < rows; ++y)
    {
      unsigned char* line = raw + y * (rowbytes + 1);
      const int      f    = line[0];
      unsigned char* cur  = line + 1;
      for (size_t i = 0; i < rowbytes; ++i)
      {
        const int a = i >= bpp ? cur[i - bpp] : 0;
        const int b = prev ? prev[i] : 0;
        const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
        switch (f)
        {
        case 0: break;
        case 1: cur[i] = (unsigned char)(cur[i] + a); break;
        case 2: cur[i] = (unsigned char)(cur[i] + b); break;
        case 3: cur[i] = (unsigned char)(cur[i] + ((a + b) >> 1)); break;
        case 4: cur[i] = (unsigned char)(cur[i] + paeth(a, b, c)); break;
        default: return false;
        }
      }
      prev = cur;
    }
    return true;
  }

  // Grey sample k of an unfiltered row, as 8 bits: 16-bit samples keep the
  // high byte, 1/2/4-bit samples are scaled by 255 / (2^depth - 1).
  unsigned char grey_sample(const unsigned char* row, size_t k, int depth)
  {
    switch (depth)
    {
    case 16: return row[2 * k];
    case 8: return row[k];
    default:
    {
      const int per  = 8 / depth;
      const int byte = row[k / per];
      const int sh   = 8 - depth * (int)(k % per + 1);
      const int v    = (byte >> sh) & ((1 << depth) - 1);
      return (unsigned char)(v * (255 / ((1 << depth) - 1)));
    }
    }
  }

  using dtio::Decoded;
} // namespace

// Decodes the file's header and, for a one-component image, its samples.
// Returns an empty string on success, else the reason.
std::string dtio::decode_png(const std::vector<unsigned char>& f, Decoded& out)
{
  if (f.size() < 8 || std::memcmp(f.data(), kSignature, 8) != 0)
    return "not a PNG file";
  size_t                     pos = 8;
  Header                     h;
  bool                       have_header = false, have_trns = false, have_plte = false;
  std::vector<unsigned char> idat;
  while (pos + 12 <= f.size())
  {
    const uint32_t len = be32(&f[pos]);
    if (len > f.size() - pos - 12)
      return "truncated chunk";
    const char*          type = reinterpret_cast<const char*>(&f[pos + 4]);
    const unsigned char* data = &f[pos + 8];
    if (std::memcmp(type, "IHDR", 4) == 0)
    {
      if (len != 13)
        return "bad IHDR";
      h.width     = be32(data);
      h.height    = be32(data + 4);
      h.depth     = data[8];
      h.color     = data[9];
      h.interlace = data[12];
      if (data[10] != 0 || data[11] != 0 || h.interlace > 1 || !valid_depth(h.color, h.depth))
        return "unsupported IHDR";
      if (h.width == 0 || h.height == 0 || h.width > (1u << 24) || h.height > (1u << 24) ||
          (unsigned long long)h.width * h.height > (1ull << 31))
        return "bad image size";
      have_header = true;
    }
    else if (std::memcmp(type, "PLTE", 4) == 0)
      have_plte = true;
    else if (std::memcmp(type, "tRNS", 4) == 0)
      have_trns = true;
    else if (std::memcmp(type, "IDAT", 4) == 0)
      idat.insert(idat.end(), data, data + len);
    else if (std::memcmp(type, "IEND", 4) == 0)
      break;
    pos += 12 + (size_t)len;
  }
  if (!have_header)
    return "missing IHDR";
  if (h.color == 3 && !have_plte)
    return "missing PLTE";
  out.width  = (int)h.width;
  out.height = (int)h.height;
  // Components as stb_image reports them: palette images expand to RGB(A),
  // a tRNS chunk adds an alpha component to grey and RGB images.
  if (h.color == 3)
    out.components = have_trns ? 4 : 3;
  else
    out.components = channels_of(h.color) + ((have_trns && (h.color == 0 || h.color == 2)) ? 1 : 0);
  if (out.components != 1 || out.header_only)
    return "";

  // Grey without tRNS: inflate, unfilter and convert to 8-bit samples.
  const size_t W = h.width, H = h.height;
  struct Pass
  {
    size_t x0, y0, dx, dy;
  };
  const Pass   adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                           {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  const Pass   whole    = {0, 0, 1, 1};
  const int    npass    = h.interlace ? 7 : 1;
  const size_t bpp      = std::max<size_t>(1, (size_t)h.depth / 8);
  size_t       need     = 0;
  for (int k = 0; k < npass; ++k)
  {
    const Pass&  p  = h.interlace ? adam7[k] : whole;
    const size_t pw = W > p.x0 ? (W - p.x0 + p.dx - 1) / p.dx : 0;
    const size_t ph = H > p.y0 ? (H - p.y0 + p.dy - 1) / p.dy : 0;
    if (pw && ph)
      need += ph * (1 + (pw * (size_t)h.depth + 7) / 8);
  }
  std::vector<unsigned char> raw(need);
  z_stream                   zs{};
  if (inflateInit(&zs) != Z_OK)
    return "zlib init failed";
  zs.next_in   = idat.data();
  zs.avail_in  = (uInt)idat.size();
  zs.next_out  = raw.data();
  zs.avail_out = (uInt)raw.size();
  const int zr = inflate(&zs, Z_FINISH);
  const size_t got = raw.size() - zs.avail_out;
  inflateEnd(&zs);
  if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR && zr != Z_OK) || got < need)
    return "corrupt image data";

  out.grey.assign(W * H, 0);
  size_t off = 0;
  for (int k = 0; k < npass; ++k)
  {
    const Pass&  p  = h.interlace ? adam7[k] : whole;
    const size_t pw = W > p.x0 ? (W - p.x0 + p.dx - 1) / p.dx : 0;
    const size_t ph = H > p.y0 ? (H - p.y0 + p.dy - 1) / p.dy : 0;
    if (!pw || !ph)
      continue;
    const size_t rowbytes = (pw * (size_t)h.depth + 7) / 8;
    if (!unfilter(raw.data() + off, ph, rowbytes, bpp))
      return "bad filter type";
    for (size_t y = 0; y < ph; ++y)
    {
      const unsigned char* row = raw.data() + off + y * (rowbytes + 1) + 1;
      for (size_t x = 0; x < pw; ++x)
        out.grey[(p.y0 + y * p.dy) * W + p.x0 + x * p.dx] = grey_sample(row, x, h.depth);
    }
    off += ph * (rowbytes + 1);
  }
  return "";
}

namespace
{
  bool read_file(const char* path, std::vector<unsigned char>& bytes)
  {
    std::ifstream in(path, std::ios::binary);
    if (!in)
      return false;
    bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return true;
  }

  void chunk(std::vector<unsigned char>& out, const char* type, const unsigned char* data, size_t len)
  {
    put32(out, (uint32_t)len);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + len);
    put32(out, (uint32_t)crc32(0L, out.data() + start, (uInt)(len + 4)));
  }

  // Encodes 8-bit grey (channels 1) or RGB (channels 3) rows read at pitch
  // bytes.  Each row takes the filter with the smallest sum of absolute
  // filtered bytes (as signed values), the usual PNG encoder heuristic.
  std::string encode(const unsigned char* buf, int width, int height, long long pitch, int channels,
                     std::vector<unsigned char>& png)
  {
    const size_t               bpp = (size_t)channels, rowbytes = (size_t)width * bpp;
    std::vector<unsigned char> filtered((rowbytes + 1) * (size_t)height), trial(rowbytes);
    for (int y = 0; y < height; ++y)
    {
      const unsigned char* cur  = buf + (long long)y * pitch;
      const unsigned char* prev = y > 0 ? buf + (long long)(y - 1) * pitch : nullptr;
      unsigned char*       dst  = filtered.data() + (size_t)y * (rowbytes + 1);
      long long            best = -1;
      for (int f = 0; f < 5; ++f)
      {
        long long sum = 0;
        for (size_t i = 0; i < rowbytes; ++i)
        {
          const int a = i >= bpp ? cur[i - bpp] : 0;
          const int b = prev ? prev[i] : 0;
          const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
          int       pred = 0;
          switch (f)
          {
          case 1: pred = a; break;
          case 2: pred = b; break;
          case 3: pred = (a + b) >> 1; break;
          case 4: pred = paeth(a, b, c); break;
          default: break;
          }
          trial[i] = (unsigned char)(cur[i] - pred);
          sum += std::abs((int)(signed char)trial[i]);
        }
        if (best < 0 || sum < best)
        {
          best   = sum;
          dst[0] = (unsigned char)f;
          std::memcpy(dst + 1, trial.data(), rowbytes);
        }
      }
    }
    uLongf                     zlen = compressBound((uLong)filtered.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, filtered.data(), (uLong)filtered.size(), 8) != Z_OK)
      return "zlib compression failed";
    png.assign(kSignature, kSignature + 8);
    unsigned char ihdr[13];
    const uint32_t w = (uint32_t)width, h = (uint32_t)height;
    const unsigned char hdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                                   (unsigned char)w,         (unsigned char)(h >> 24), (unsigned char)(h >> 16),
                                   (unsigned char)(h >> 8),  (unsigned char)h,         8,
                                   (unsigned char)(channels == 3 ? 2 : 0), 0, 0, 0};
    std::memcpy(ihdr, hdr, 13);
    chunk(png, "IHDR", ihdr, 13);
    chunk(png, "IDAT", z.data(), zlen);
    chunk(png, "IEND", nullptr, 0);
    return "";
  }

  bool ends_with_png(const char* name)
  {
    const size_t n = std::strlen(name);
    return n >= 4 && std::strcmp(name + n - 4, ".png") == 0;
  }

  void save(const char* filename, const unsigned char* buf, int width, int height, long long pitch, int channels)
  {
    if (!ends_with_png(filename))
      throw std::invalid_argument("Only PNG output is supported");
    std::vector<unsigned char> png;
    if (width <= 0 || height <= 0 || !encode(buf, width, height, pitch, channels, png).empty())
      throw std::runtime_error("Failed to save PNG");
    std::ofstream out(filename, std::ios::binary);
    out.write(reinterpret_cast<const char*>(png.data()), (std::streamsize)png.size());
    if (!out)
      throw std::runtime_error("Failed to save PNG");
  }

  thread_local std::string g_io_err;
} // namespace

namespace dt
{
  // dt/src/imread.cpp:9-29 of the reference.
  template <>
  image2d<std::uint8_t> imread(const char* filename)
  {
    std::vector<unsigned char> bytes;
    Decoded                    d;
    bool                       ok = false;
    try
    {
      ok = read_file(filename, bytes) && dtio::decode_image(bytes, d).empty();
    }
    catch (const std::bad_alloc&)
    {
      ok = false;
    }
    if (!ok)
      throw std::runtime_error(std::string("Error while loading ") + filename);
    if (d.components != 1)
      throw std::invalid_argument("Input image must have 1 component (grayscale) of 8 bits");
    image2d<std::uint8_t> out(d.width, d.height);
    for (int y = 0; y < d.height; ++y)
      for (int x = 0; x < d.width; ++x)
        out(x, y) = d.grey[(size_t)y * d.width + x];
    return out;
  }

  namespace impl
  {
    // dt/src/imsave.cpp:9-30 of the reference.
    template <>
    void imsave<std::uint8_t>(const char* filename, const std::uint8_t* buffer, int width, int height, int pitch)
    {
      save(filename, buffer, width, height, pitch, 1);
    }

    template <>
    void imsave<rgb8>(const char* filename, const rgb8* buffer, int width, int height, int pitch)
    {
      save(filename, reinterpret_cast<const unsigned char*>(buffer), width, height, pitch, 3);
    }
  } // namespace impl
} // namespace dt

extern "C"
{
  const char* dt_io_last_error(void) { return g_io_err.c_str(); }

  int dt_png_info(const char* path, int* width, int* height, int* components)
  {
    std::vector<unsigned char> bytes;
    Decoded                    d;
    std::string                why;
    try
    {
      if (!path || !read_file(path, bytes))
        why = "cannot read the file";
      else
      {
        d.header_only = true;
        why           = dtio::decode_image(bytes, d);
      }
    }
    catch (const std::exception& e)
    {
      why = e.what();
    }
    if (!why.empty())
    {
      g_io_err = std::string("Error while loading ") + (path ? path : "(null)") + ": " + why;
      return DT_IO_ERR_READ;
    }
    if (width)
      *width = d.width;
    if (height)
      *height = d.height;
    if (components)
      *components = d.components;
    return DT_IO_OK;
  }

  int dt_png_read_u8(const char* path, unsigned char* out, int width, int height, long long pitch)
  {
    try
    {
      const auto img = dt::imread<std::uint8_t>(path);
      if (img.width() != width || img.height() != height || !out || pitch < width)
      {
        g_io_err = "dt_png_read_u8: buffer does not match the image size";
        return DT_IO_ERR_ARGUMENT;
      }
      for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
          out[(long long)y * pitch + x] = img(x, y);
      return DT_IO_OK;
    }
    catch (const std::invalid_argument& e)
    {
      g_io_err = e.what();
      return DT_IO_ERR_FORMAT;
    }
    catch (const std::exception& e)
    {
      g_io_err = e.what();
      return DT_IO_ERR_READ;
    }
  }

  int dt_png_write(const char* path, const unsigned char* buf, int width, int height, long long pitch, int channels)
  {
    if (!path || !buf || (channels != 1 && channels != 3) || pitch < (long long)width * channels)
    {
      g_io_err = "dt_png_write: bad arguments";
      return DT_IO_ERR_ARGUMENT;
    }
    try
    {
      save(path, buf, width, height, pitch, channels);
      return DT_IO_OK;
    }
    catch (const std::invalid_argument& e)
    {
      g_io_err = e.what();
      return DT_IO_ERR_FORMAT;
    }
    catch (const std::exception& e)
    {
      g_io_err = e.what();
      return DT_IO_ERR_WRITE;
    }
  }
}
