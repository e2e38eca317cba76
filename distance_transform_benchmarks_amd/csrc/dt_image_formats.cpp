// Format selection for dt::imread, and the small formats: binary PGM/PPM
// and TGA.  Each follows what stbi_load(filename, &w, &h, &n, 0) returns for
// the format in the reference's stb_image 2.30 (dt/src/stb_image.h), the
// decoder behind dt/src/imread.cpp:12, so that the same files load to the
// same pixels or fail the same way:
//   - PGM (P5): 8-bit samples as stored.  16-bit samples (maxval > 255):
//     that decoder reads the big-endian samples into host-order 16-bit words
//     and keeps their high byte, which on the little-endian hosts it runs on
//     is each sample's second (least significant) byte; that byte is kept
//     here.  A header is "P5", whitespace and '#' comments, width, height,
//     maxval, and exactly one byte before the samples.  PPM (P6) has three
//     components.
//   - TGA: grey (image type 3 or 11) and 8-bit "true colour" (type 2 or 10)
//     images have one component; colour-mapped images (types 1 and 9) with
//     8-bit palette entries have one component, read through the palette
//     (an index past its end reads entry 0).  Rows are stored bottom-up
//     unless bit 5 of the descriptor is set.  Other depths have 2, 3 or 4
//     components.
//   - BMP, GIF, PSD, Softimage PIC and Radiance HDR files always decode to
//     three or four components there; they are recognised by the same
//     header tests and reported with that count (and their size where the
//     header holds it) without decoding, since imread rejects them either
//     way.  A damaged file of one of these formats therefore raises the
//     component error here where the reference would raise the loading
//     error.
// Truncated pixel data: that decoder returns the bytes it could read and
// leaves the rest uninitialised (TGA) or fails (PGM); here missing TGA bytes
// read as zeros and a short PGM fails.
#include "dt_image_decode.hpp"

#include <algorithm>
#include <climits>
#include <cstring>

namespace dtio
{
  namespace
  {
    constexpr int kMaxDim = 1 << 24; // the largest side the reference's decoder accepts

    uint32_t le16(const std::vector<unsigned char>& f, size_t i)
    {
      return i + 1 < f.size() ? (uint32_t)f[i] | (uint32_t)f[i + 1] << 8 : 0;
    }
    uint32_t le32(const std::vector<unsigned char>& f, size_t i) { return le16(f, i) | le16(f, i + 2) << 16; }
    uint32_t be16(const std::vector<unsigned char>& f, size_t i)
    {
      return i + 1 < f.size() ? (uint32_t)f[i] << 8 | (uint32_t)f[i + 1] : 0;
    }
    uint32_t be32(const std::vector<unsigned char>& f, size_t i) { return be16(f, i) << 16 | be16(f, i + 2); }
    bool     starts(const std::vector<unsigned char>& f, size_t at, const char* s)
    {
      const size_t n = std::strlen(s);
      return f.size() >= at + n && std::memcmp(f.data() + at, s, n) == 0;
    }

    // --- header tests, in the order the reference's decoder tries them ---

    bool is_png(const std::vector<unsigned char>& f)
    {
      static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
      return f.size() >= 8 && std::memcmp(f.data(), sig, 8) == 0;
    }
    bool is_bmp(const std::vector<unsigned char>& f)
    {
      if (!starts(f, 0, "BM"))
        return false;
      const uint32_t hsz = le32(f, 14);
      return hsz == 12 || hsz == 40 || hsz == 56 || hsz == 108 || hsz == 124;
    }
    bool is_gif(const std::vector<unsigned char>& f) { return starts(f, 0, "GIF87a") || starts(f, 0, "GIF89a"); }
    bool is_psd(const std::vector<unsigned char>& f) { return starts(f, 0, "8BPS"); }
    bool is_pic(const std::vector<unsigned char>& f) { return starts(f, 0, "\x53\x80\xF6\x34") && starts(f, 88, "PICT"); }
    bool is_jpeg(const std::vector<unsigned char>& f)
    {
      // A marker: 0xFF, any fill bytes 0xFF, then SOI (0xD8).
      size_t i = 0;
      if (f.empty() || f[0] != 0xFF)
        return false;
      while (i < f.size() && f[i] == 0xFF)
        ++i;
      return i < f.size() && f[i] == 0xD8;
    }
    bool is_pnm(const std::vector<unsigned char>& f) { return f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6'); }
    bool is_hdr(const std::vector<unsigned char>& f) { return starts(f, 0, "#?RADIANCE\n") || starts(f, 0, "#?RGBE\n"); }
    bool is_tga(const std::vector<unsigned char>& f)
    {
      if (f.size() < 17)
        return false;
      const int cmap = f[1], type = f[2];
      if (cmap > 1)
        return false;
      if (cmap == 1)
      {
        if (type != 1 && type != 9)
          return false;
        const int pal_bits = f[7];
        if (pal_bits != 8 && pal_bits != 15 && pal_bits != 16 && pal_bits != 24 && pal_bits != 32)
          return false;
      }
      else if (type != 2 && type != 3 && type != 10 && type != 11)
        return false;
      if (le16(f, 12) < 1 || le16(f, 14) < 1)
        return false;
      const int bpp = f[16];
      if (cmap == 1 && bpp != 8 && bpp != 16)
        return false;
      return bpp == 8 || bpp == 15 || bpp == 16 || bpp == 24 || bpp == 32;
    }

    // Formats that never decode to one component: count and size only.
    void colour_only(const std::vector<unsigned char>& f, char kind, Decoded& out)
    {
      switch (kind)
      {
      case 'B': // BMP: 32-bit pixels carry alpha
        out.width      = (int)le32(f, 18);
        out.height     = std::abs((int)le32(f, 22));
        out.components = le16(f, 28) == 32 ? 4 : 3;
        break;
      case 'G': // GIF: always RGBA
        out.width      = (int)le16(f, 6);
        out.height     = (int)le16(f, 8);
        out.components = 4;
        break;
      case 'P': // PSD: always RGBA
        out.height     = (int)be32(f, 14);
        out.width      = (int)be32(f, 18);
        out.components = 4;
        break;
      case 'S': // Softimage PIC: RGB or RGBA by its channel packets
        out.width      = (int)be16(f, 92);
        out.height     = (int)be16(f, 94);
        out.components = 4;
        break;
      default: // Radiance HDR: RGB; the size is in a text line not parsed here
        out.components = 3;
        break;
      }
    }
  } // namespace

  std::string decode_pnm(const std::vector<unsigned char>& f, Decoded& out)
  {
    if (!is_pnm(f))
      return "not a binary PGM/PPM file";
    size_t pos  = 2;
    auto   eof  = [&] { return pos >= f.size(); };
    auto   next = [&] { return (char)(pos < f.size() ? f[pos++] : 0); };
    auto   sp   = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; };
    char   c    = next();
    auto   skip_space = [&] {
      for (;;)
      {
        while (!eof() && sp(c))
          c = next();
        if (eof() || c != '#')
          break;
        while (!eof() && c != '\n' && c != '\r')
          c = next();
      }
    };
    bool overflow = false;
    auto integer  = [&] {
      int v = 0;
      while (!eof() && c >= '0' && c <= '9')
      {
        v = v * 10 + (c - '0');
        c = next();
        if (v > 214748364 || (v == 214748364 && c > '7'))
        {
          overflow = true;
          return 0;
        }
      }
      return v;
    };
    skip_space();
    const int w = integer();
    if (!w || overflow)
      return "bad PNM width";
    skip_space();
    const int h = integer();
    if (!h || overflow)
      return "bad PNM height";
    skip_space();
    // An overflowing maxval parses as 0 (8-bit samples), as in the reference's decoder.
    const int maxv = integer();
    if (maxv > 65535)
      return "bad PNM maximum value";
    if (w > kMaxDim || h > kMaxDim)
      return "image too large";
    const int       n     = f[1] == '6' ? 3 : 1;
    const int       bytes = maxv > 255 ? 2 : 1;
    const long long need  = (long long)n * w * h * bytes;
    if (need > INT_MAX)
      return "image too large";
    if ((long long)(f.size() - std::min(pos, f.size())) < need)
      return "PNM file truncated";
    out.width      = w;
    out.height     = h;
    out.components = n;
    if (n != 1 || out.header_only)
      return "";
    out.grey.resize((size_t)w * h);
    for (size_t i = 0; i < out.grey.size(); ++i)
      out.grey[i] = f[pos + i * bytes + (bytes - 1)];
    return "";
  }

  std::string decode_tga(const std::vector<unsigned char>& f, Decoded& out)
  {
    if (!is_tga(f))
      return "not a TGA file";
    const int  id_len = f[0], indexed = f[1], type = f[2];
    const int  pal_start = (int)le16(f, 3), pal_len = (int)le16(f, 5), pal_bits = f[7];
    const int  w = (int)le16(f, 12), h = (int)le16(f, 14), bpp = f[16];
    const int  desc     = f.size() > 17 ? f[17] : 0;
    const bool rle      = type >= 8;
    const int  base     = rle ? type - 8 : type;
    const bool inverted = ((desc >> 5) & 1) == 0; // bottom-up rows
    // Components by the reference's rules: the palette entry size for
    // colour-mapped images, else the pixel size (16-bit grey has alpha).
    const int depth = indexed ? pal_bits : bpp;
    int       comp  = 0;
    switch (depth)
    {
    case 8: comp = 1; break;
    case 16: comp = (!indexed && base == 3) ? 2 : 3; break;
    case 15: comp = 3; break;
    case 24: comp = 3; break;
    case 32: comp = 4; break;
    default: return "unsupported TGA pixel format";
    }
    out.width      = w;
    out.height     = h;
    out.components = comp;
    if (comp != 1 || out.header_only)
      return "";
    if ((long long)w * h > INT_MAX)
      return "image too large";

    size_t pos  = 18 + (size_t)id_len;
    auto   get8 = [&]() -> int { return pos < f.size() ? f[pos++] : 0; };
    out.grey.assign((size_t)w * h, 0);
    if (!indexed && !rle)
    {
      for (int i = 0; i < h; ++i)
        for (int x = 0; x < w; ++x)
          out.grey[(size_t)i * w + x] = (unsigned char)get8();
    }
    else
    {
      std::vector<unsigned char> pal;
      if (indexed)
      {
        if (pal_len == 0)
          return "TGA palette is empty";
        pos = std::min(pos + (size_t)pal_start, f.size()); // the reference's decoder skips this many bytes
        if (f.size() - pos < (size_t)pal_len)
          return "TGA palette truncated";
        pal.assign(f.begin() + (long)pos, f.begin() + (long)pos + pal_len);
        pos += (size_t)pal_len;
      }
      int           run = 0;
      bool          repeat = false, fresh = true;
      unsigned char v = 0;
      for (size_t i = 0; i < out.grey.size(); ++i)
      {
        if (rle)
        {
          if (run == 0)
          {
            const int cmd = get8();
            run           = 1 + (cmd & 127);
            repeat        = (cmd >> 7) != 0;
            fresh         = true;
          }
          else if (!repeat)
            fresh = true;
        }
        else
          fresh = true;
        if (fresh)
        {
          if (indexed)
          {
            int idx = get8();
            if (bpp == 16)
              idx |= get8() << 8;
            v = pal[idx < pal_len ? idx : 0];
          }
          else
            v = (unsigned char)get8();
          fresh = false;
        }
        out.grey[i] = v;
        --run;
      }
    }
    if (inverted)
      for (int y = 0; y < h / 2; ++y)
        std::swap_ranges(out.grey.begin() + (long)y * w, out.grey.begin() + (long)(y + 1) * w,
                         out.grey.begin() + (long)(h - 1 - y) * w);
    return "";
  }

  std::string decode_image(const std::vector<unsigned char>& f, Decoded& out)
  {
    if (is_png(f))
      return decode_png(f, out);
    if (is_bmp(f))
      return colour_only(f, 'B', out), "";
    if (is_gif(f))
      return colour_only(f, 'G', out), "";
    if (is_psd(f))
      return colour_only(f, 'P', out), "";
    if (is_pic(f))
      return colour_only(f, 'S', out), "";
    if (is_jpeg(f))
      return decode_jpeg(f, out);
    if (is_pnm(f))
      return decode_pnm(f, out);
    if (is_hdr(f))
      return colour_only(f, 'H', out), "";
    if (is_tga(f))
      return decode_tga(f, out);
    return "unknown image format";
  }
} // namespace dtio
