// One-component JPEG decoding for dt::imread: baseline and extended
// sequential (SOF0/SOF1) and progressive (SOF2) Huffman streams of 8-bit
// samples, decoded per ITU-T T.81.  A colour JPEG (three or four
// components) is reported as three components without decoding it, which is
// what the reference's decoder reports (dt/src/stb_image.h:4020) before
// imread rejects it.
//
// Pixels equal those of the decoder the reference runs (stb_image 2.30 built
// for x86-64, dt/src/stb_image.h): entropy decoding is fixed by the standard,
// and the inverse DCT below reproduces that decoder's integer arithmetic (the
// LL&M factorisation with 12-bit constants, its rounding, and the 16-bit
// saturation between the column and row passes of its SSE2 code,
// dt/src/stb_image.h:2426-2521 and :2530-2640).  Stream edge cases follow it
// too: bits past a marker or the end of the file read as zeros, a marker
// other than a restart where a restart is due ends the scan, and an unknown
// marker after the frame header ends the image without an error.  Where that
// decoder returns uninitialised memory (blocks a truncated scan never
// reached, a progressive file without its EOI marker), this one returns the
// decoded coefficients' pixels, zeros for blocks never coded.
#include "dt_image_decode.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace dtio
{
  namespace
  {
    constexpr int kNone = -1; // no marker

    // Natural (row-major) index of the k-th coefficient of the zig-zag scan
    // (T.81 figure A.6).  Entries 64..79 repeat 63, so a zero run that
    // overshoots the band of a corrupt stream stays inside the block.
    struct ZigZag
    {
      uint8_t natural[80];
      ZigZag()
      {
        int k = 0;
        for (int d = 0; d < 15; ++d) // anti-diagonal row + col = d
        {
          const int lo = std::max(0, d - 7), hi = std::min(d, 7);
          for (int i = 0; i <= hi - lo; ++i)
          {
            const int row = (d & 1) ? lo + i : hi - i; // odd diagonals run down-left, even ones up-right
            natural[k++]  = (uint8_t)(row * 8 + d - row);
          }
        }
        for (; k < 80; ++k)
          natural[k] = 63;
      }
    };
    const ZigZag kZig;

    // The 8-point inverse DCT as an exact integer matrix.  The reference's
    // decoder evaluates the Loeffler-Ligtenberg-Moschytz flowgraph with its
    // rotation constants in 12-bit fixed point; every node of the flowgraph
    // is an integer sum or a product by a constant, so before the rounding
    // shift each output is an exact integer linear combination of the
    // inputs.  The combination is found once by running the flowgraph on
    // unit impulses, and the transform is then applied as that matrix.
    struct IdctMatrix
    {
      int32_t m[8][8]; // out[n] = sum_k m[n][k] * in[k], scaled by 2^12 * sqrt(8) / 2 per pass

      // A constant of the flowgraph in 12-bit fixed point, rounded the way
      // the reference's decoder rounds it: (int)(c * 4096 + 0.5), which for
      // the negative constants truncates toward zero (-1.8478 -> -7567).
      static int64_t fix12(double c) { return (int64_t)(int)(c * 4096.0 + 0.5); }

      static void flowgraph(const int64_t s[8], int64_t o[8])
      {
        const double r2 = std::sqrt(2.0);
        double       cs[8];
        for (int k = 0; k < 8; ++k)
          cs[k] = std::cos(k * M_PI / 16.0);
        // Even half: inputs 0, 2, 4, 6; one rotation by 3*pi/8.
        const int64_t rot = (s[2] + s[6]) * fix12(r2 * cs[6]);
        const int64_t e2  = rot + s[6] * fix12(-r2 * (cs[2] + cs[6]));
        const int64_t e3  = rot + s[2] * fix12(r2 * (cs[2] - cs[6]));
        const int64_t sum = (s[0] + s[4]) * 4096, dif = (s[0] - s[4]) * 4096;
        const int64_t a0 = sum + e3, a3 = sum - e3, a1 = dif + e2, a2 = dif - e2;
        // Odd half: inputs 1, 3, 5, 7; a shared product and four pair products.
        const int64_t p71 = s[7] + s[1], p53 = s[5] + s[3], p73 = s[7] + s[3], p51 = s[5] + s[1];
        const int64_t z   = (p73 + p51) * fix12(r2 * cs[3]);
        const int64_t q71 = z + p71 * fix12(r2 * (cs[7] - cs[3]));
        const int64_t q53 = z + p53 * fix12(-r2 * (cs[1] + cs[3]));
        const int64_t q73 = p73 * fix12(-r2 * (cs[3] + cs[5]));
        const int64_t q51 = p51 * fix12(r2 * (cs[5] - cs[3]));
        const int64_t b0  = s[7] * fix12(r2 * (-cs[1] + cs[3] + cs[5] - cs[7])) + q71 + q73;
        const int64_t b1  = s[5] * fix12(r2 * (cs[1] + cs[3] - cs[5] + cs[7])) + q53 + q51;
        const int64_t b2  = s[3] * fix12(r2 * (cs[1] + cs[3] + cs[5] - cs[7])) + q53 + q73;
        const int64_t b3  = s[1] * fix12(r2 * (cs[1] + cs[3] - cs[5] - cs[7])) + q71 + q51;
        o[0] = a0 + b3;
        o[7] = a0 - b3;
        o[1] = a1 + b2;
        o[6] = a1 - b2;
        o[2] = a2 + b1;
        o[5] = a2 - b1;
        o[3] = a3 + b0;
        o[4] = a3 - b0;
      }

      IdctMatrix()
      {
        for (int k = 0; k < 8; ++k)
        {
          int64_t in[8] = {0, 0, 0, 0, 0, 0, 0, 0}, out[8];
          in[k] = 1;
          flowgraph(in, out);
          for (int n = 0; n < 8; ++n)
            m[n][k] = (int32_t)out[n];
        }
      }
    };
    const IdctMatrix kIdct;

    // Inverse DCT of one dequantised block into 8 x 8 pixels.  Sums run in
    // unsigned 32-bit arithmetic, i.e. modulo 2^32 as the reference's
    // decoder's 32-bit lanes do; the column pass keeps two extra bits and
    // saturates to 16 bits, the row pass removes 2^17 with rounding, adds
    // the 128 level shift and clamps to 0..255.
    void idct_block(const int16_t* coef, unsigned char* out, size_t stride)
    {
      int32_t v[64];
      for (int c = 0; c < 8; ++c)
      {
        bool dc_only = true;
        for (int k = 1; k < 8; ++k)
          dc_only = dc_only && coef[k * 8 + c] == 0;
        if (dc_only) // every output is (4096 * dc + 512) >> 10 = 4 * dc
        {
          const int32_t d = std::clamp(4 * (int32_t)coef[c], -32768, 32767);
          for (int n = 0; n < 8; ++n)
            v[n * 8 + c] = d;
          continue;
        }
        for (int n = 0; n < 8; ++n)
        {
          uint32_t acc = 512u;
          for (int k = 0; k < 8; ++k)
            acc += (uint32_t)kIdct.m[n][k] * (uint32_t)(int32_t)coef[k * 8 + c];
          v[n * 8 + c] = std::clamp((int32_t)acc >> 10, -32768, 32767);
        }
      }
      for (int r = 0; r < 8; ++r)
      {
        const int32_t* in   = v + r * 8;
        bool           flat = true;
        for (int k = 1; k < 8; ++k)
          flat = flat && in[k] == 0;
        for (int n = 0; n < 8; ++n)
        {
          uint32_t acc = 65536u + (128u << 17);
          if (flat) // every kIdct.m[n][0] is 4096
            acc += 4096u * (uint32_t)in[0];
          else
            for (int k = 0; k < 8; ++k)
              acc += (uint32_t)kIdct.m[n][k] * (uint32_t)in[k];
          out[r * stride + n] = (unsigned char)std::clamp((int32_t)acc >> 17, 0, 255);
        }
      }
    }

    // Entropy-coded bits, most significant first, in a 64-bit window.
    // 0xFF 0x00 is a data byte 0xFF (fill bytes 0xFF before it are skipped);
    // 0xFF followed by anything else is a marker, which stops the reading:
    // from there on, and past the end of the file, the stream reads as
    // zero bits.
    class Bits
    {
    public:
      Bits(const std::vector<unsigned char>& f, size_t pos) : f_(f), pos_(pos) {}

      void fill()
      {
        while (n_ <= 56)
        {
          unsigned b = 0;
          if (marker_ == kNone && pos_ < f_.size())
          {
            b = f_[pos_++];
            if (b == 0xFF)
            {
              size_t q = pos_;
              while (q < f_.size() && f_[q] == 0xFF)
                ++q;
              if (q < f_.size() && f_[q] != 0)
              {
                marker_ = f_[q];
                b       = 0;
              }
              pos_ = std::min(q + 1, f_.size());
            }
          }
          win_ |= (uint64_t)b << (56 - n_);
          n_ += 8;
        }
      }
      unsigned peek(int k)
      {
        if (n_ < k)
          fill();
        return (unsigned)(win_ >> (64 - k));
      }
      void skip(int k)
      {
        win_ <<= k;
        n_ -= k;
      }
      unsigned get(int k)
      {
        if (k == 0)
          return 0;
        const unsigned v = peek(k);
        skip(k);
        return v;
      }
      // RECEIVE then EXTEND (T.81 F.2.2.1): k magnitude bits as a signed value.
      int extend(int k)
      {
        const int v = (int)get(k);
        return v < (1 << (k - 1)) ? v - (1 << k) + 1 : v;
      }
      // Drops the window and any marker seen, after a restart marker.
      void restart()
      {
        win_    = 0;
        n_      = 0;
        marker_ = kNone;
      }
      int    marker() const { return marker_; }
      size_t pos() const { return pos_; }

    private:
      const std::vector<unsigned char>& f_;
      size_t                            pos_;
      uint64_t                          win_    = 0;
      int                               n_      = 0;
      int                               marker_ = kNone;
    };

    // Canonical Huffman table (T.81 annex C, decoding per F.2.2.3), with a
    // direct lookup for codes of at most kFastBits bits.
    constexpr int kFastBits = 9;
    struct Huffman
    {
      bool     defined = false;
      uint8_t  value[256];
      int32_t  maxcode[17]; // largest code of each length, -1 if that length has none
      int32_t  mincode[17];
      int      first[17];   // index in value[] of the smallest code of each length
      uint16_t fast[1 << kFastBits]; // (length << 8 | index) of the code a prefix starts with, 0 if longer

      bool build(const int count[16])
      {
        std::memset(fast, 0, sizeof(fast));
        std::memset(value, 0, sizeof(value)); // symbols past the table's count read as 0
        uint32_t code = 0;
        int      k    = 0;
        for (int len = 1; len <= 16; ++len)
        {
          first[len]   = k;
          mincode[len] = (int32_t)code;
          for (int i = 0; i < count[len - 1]; ++i, ++code, ++k)
          {
            if (code >= (1u << len))
              return false; // more codes than the length can hold
            if (len <= kFastBits)
              for (uint32_t j = 0; j < (1u << (kFastBits - len)); ++j)
                fast[(code << (kFastBits - len)) | j] = (uint16_t)(len << 8 | k);
          }
          maxcode[len] = count[len - 1] ? (int32_t)code - 1 : -1;
          code <<= 1;
        }
        defined = true;
        return true;
      }

      // Returns the decoded symbol, or -1 for a bit pattern that is no code.
      int decode(Bits& b) const
      {
        const unsigned f = fast[b.peek(kFastBits)];
        if (f)
        {
          b.skip(f >> 8);
          return value[f & 0xFF];
        }
        for (int len = kFastBits + 1; len <= 16; ++len)
        {
          const int32_t c = (int32_t)b.peek(len);
          if (maxcode[len] >= 0 && c <= maxcode[len])
          {
            b.skip(len);
            return value[first[len] + c - mincode[len]];
          }
        }
        return -1;
      }
    };

    bool fits_short(long long v) { return v >= SHRT_MIN && v <= SHRT_MAX; }

    struct Jpeg
    {
      const std::vector<unsigned char>& f;
      size_t                            pos = 0;
      int                               marker = kNone; // a marker read ahead by a scan

      uint16_t quant[4][64] = {}; // natural order
      Huffman  dc[4], ac[4];
      int      restart_interval = 0;

      bool progressive = false;
      int  width = 0, height = 0, ncomp = 0;
      int  comp_id = 0, comp_tq = 0;

      // Scan state.
      int bw = 0, bh = 0; // blocks across and down
      int td = 0, ta = 0, ss = 0, se = 63, ah = 0, al = 0;
      int pred = 0, eobrun = 0;

      std::vector<unsigned char> pix;  // bw*8 x bh*8 (sequential)
      std::vector<int16_t>       coef; // bw*bh blocks of 64 (progressive)

      explicit Jpeg(const std::vector<unsigned char>& file) : f(file) {}

      // Header bytes; past the end of the file they read as zeros.
      int  get8() { return pos < f.size() ? f[pos++] : 0; }
      int  get16() { return get8() << 8 | get8(); }
      void skip(long long n) { pos = (size_t)std::min<long long>((long long)f.size(), (long long)pos + std::max(0ll, n)); }
      bool at_end() const { return pos >= f.size(); }

      int get_marker()
      {
        if (marker != kNone)
        {
          const int m = marker;
          marker      = kNone;
          return m;
        }
        int x = get8();
        if (x != 0xFF)
          return kNone;
        while (x == 0xFF)
          x = get8();
        return x;
      }

      // DRI, DQT, DHT, APPn and COM segments; anything else is an error.
      bool process_marker(int m)
      {
        if (m == 0xDD)
        {
          if (get16() != 4)
            return false;
          restart_interval = get16();
          return true;
        }
        if (m == 0xDB)
        {
          int len = get16() - 2;
          while (len > 0)
          {
            const int pq = get8(), prec = pq >> 4, t = pq & 15;
            if (prec > 1 || t > 3)
              return false;
            for (int i = 0; i < 64; ++i)
              quant[t][kZig.natural[i]] = (uint16_t)(prec ? get16() : get8());
            len -= prec ? 129 : 65;
          }
          return len == 0;
        }
        if (m == 0xC4)
        {
          int len = get16() - 2;
          while (len > 0)
          {
            const int tcth = get8(), tc = tcth >> 4, th = tcth & 15;
            if (tc > 1 || th > 3)
              return false;
            int count[16], n = 0;
            for (int i = 0; i < 16; ++i)
              n += count[i] = get8();
            if (n > 256)
              return false;
            Huffman& h = tc ? ac[th] : dc[th];
            if (!h.build(count))
              return false;
            for (int i = 0; i < n; ++i)
              h.value[i] = (uint8_t)get8();
            len -= 17 + n;
          }
          return len == 0;
        }
        if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE)
        {
          const int len = get16();
          if (len < 2)
            return false;
          skip(len - 2);
          return true;
        }
        return false;
      }

      std::string frame_header()
      {
        const int lf = get16();
        if (lf < 11)
          return "bad SOF length";
        if (get8() != 8)
          return "only 8-bit samples are supported";
        height = get16();
        width  = get16();
        if (!height || !width)
          return "bad image size";
        ncomp = get8();
        if (ncomp != 1 && ncomp != 3 && ncomp != 4)
          return "bad component count";
        if (lf != 8 + 3 * ncomp)
          return "bad SOF length";
        int h[4], v[4], hmax = 1, vmax = 1;
        for (int i = 0; i < ncomp; ++i)
        {
          const int id = get8(), hv = get8(), tq = get8();
          h[i] = hv >> 4;
          v[i] = hv & 15;
          if (h[i] < 1 || h[i] > 4 || v[i] < 1 || v[i] > 4 || tq > 3)
            return "bad component parameters";
          if (i == 0)
          {
            comp_id = id;
            comp_tq = tq;
          }
          hmax = std::max(hmax, h[i]);
          vmax = std::max(vmax, v[i]);
        }
        if ((long long)width * height * ncomp > INT_MAX)
          return "image too large";
        for (int i = 0; i < ncomp; ++i)
          if (hmax % h[i] || vmax % v[i])
            return "bad sampling factors";
        return "";
      }

      std::string scan_header()
      {
        const int ls = get16(), ns = get8();
        if (ns < 1 || ns > ncomp)
          return "bad SOS component count";
        if (ls != 6 + 2 * ns)
          return "bad SOS length";
        for (int i = 0; i < ns; ++i)
        {
          const int id = get8(), t = get8();
          if (id != comp_id) // a one-component image: the scan must name it
            return "unknown component in SOS";
          td = t >> 4;
          ta = t & 15;
          if (td > 3 || ta > 3)
            return "bad Huffman table selector";
        }
        ss = get8();
        se = get8();
        const int a = get8();
        ah = a >> 4;
        al = a & 15;
        if (progressive)
        {
          if (ss > 63 || se > 63 || ss > se || ah > 13 || al > 13)
            return "bad progressive scan parameters";
        }
        else
        {
          if (ss != 0 || ah != 0 || al != 0)
            return "bad sequential scan parameters";
          se = 63;
        }
        return "";
      }

      // One block of a sequential scan, dequantised and transformed.
      bool sequential_block(Bits& b, int bx, int by)
      {
        const Huffman& hd = dc[td];
        const Huffman& ha = ac[ta];
        if (!hd.defined || !ha.defined)
          return false;
        const uint16_t* q = quant[comp_tq];
        int16_t         blk[64];
        std::memset(blk, 0, sizeof(blk));
        const int t = hd.decode(b);
        if (t < 0 || t > 15)
          return false;
        const long long dcv = (long long)pred + (t ? b.extend(t) : 0);
        if (dcv < INT_MIN || dcv > INT_MAX)
          return false;
        pred = (int)dcv;
        if (!fits_short((long long)pred * q[0]))
          return false;
        blk[0] = (int16_t)(pred * q[0]);
        for (int k = 1; k < 64;)
        {
          const int rs = ha.decode(b);
          if (rs < 0)
            return false;
          const int r = rs >> 4, s = rs & 15;
          if (s == 0)
          {
            if (rs != 0xF0)
              break; // end of block
            k += 16;
            continue;
          }
          k += r;
          const int z = kZig.natural[k++];
          blk[z]      = (int16_t)(b.extend(s) * q[z]);
        }
        const size_t stride = (size_t)bw * 8;
        idct_block(blk, &pix[(size_t)by * 8 * stride + (size_t)bx * 8], stride);
        return true;
      }

      // Refinement of one already non-zero coefficient by one bit (G.1.2.3).
      static void refine(Bits& b, int16_t& c, int bit)
      {
        if (b.get(1) && (c & bit) == 0)
          c = (int16_t)(c > 0 ? c + bit : c - bit);
      }

      // One block of a progressive scan, into the coefficient store.
      bool progressive_block(Bits& b, int16_t* blk)
      {
        if (ss == 0) // DC scans (G.1.2.1)
        {
          if (se != 0)
            return false;
          if (ah == 0)
          {
            const Huffman& hd = dc[td];
            if (!hd.defined)
              return false;
            std::memset(blk, 0, 64 * sizeof(int16_t));
            const int t = hd.decode(b);
            if (t < 0 || t > 15)
              return false;
            const long long dcv = (long long)pred + (t ? b.extend(t) : 0);
            if (dcv < INT_MIN || dcv > INT_MAX)
              return false;
            pred = (int)dcv;
            if (!fits_short((long long)pred * (1 << al)))
              return false;
            blk[0] = (int16_t)(pred * (1 << al));
          }
          else if (b.get(1))
            blk[0] = (int16_t)(blk[0] + (1 << al));
          return true;
        }
        const Huffman& ha = ac[ta];
        if (!ha.defined)
          return false;
        if (ah == 0) // AC first scan (G.1.2.2)
        {
          if (eobrun > 0)
          {
            --eobrun;
            return true;
          }
          for (int k = ss; k <= se;)
          {
            const int rs = ha.decode(b);
            if (rs < 0)
              return false;
            const int r = rs >> 4, s = rs & 15;
            if (s == 0)
            {
              if (r < 15)
              {
                eobrun = (1 << r) - 1 + (int)b.get(r);
                break;
              }
              k += 16;
              continue;
            }
            k += r;
            blk[kZig.natural[k++]] = (int16_t)(b.extend(s) * (1 << al));
          }
          return true;
        }
        // AC refinement scan (G.1.2.3): a new coefficient is +-1 << al, and
        // every coefficient passed that is already non-zero takes one
        // correction bit.
        const int bit = 1 << al;
        if (eobrun > 0)
        {
          --eobrun;
          for (int k = ss; k <= se; ++k)
          {
            int16_t& c = blk[kZig.natural[k]];
            if (c)
              refine(b, c, bit);
          }
          return true;
        }
        for (int k = ss; k <= se;)
        {
          const int rs = ha.decode(b);
          if (rs < 0)
            return false;
          int r = rs >> 4, s = rs & 15, value = 0;
          if (s == 0)
          {
            if (r < 15)
            {
              eobrun = (1 << r) - 1 + (int)b.get(r);
              r      = 64; // refine the rest of the band, then end the block
            }
            // r == 15: pass 16 zero coefficients (the 16th takes value 0)
          }
          else
          {
            if (s != 1)
              return false;
            value = b.get(1) ? bit : -bit;
          }
          while (k <= se)
          {
            int16_t& c = blk[kZig.natural[k++]];
            if (c)
              refine(b, c, bit);
            else if (r == 0)
            {
              c = (int16_t)value;
              break;
            }
            else
              --r;
          }
        }
        return true;
      }

      // Entropy-coded segment of one scan; returns false on a coding error.
      bool scan()
      {
        Bits      b(f, pos);
        int       todo = restart_interval ? restart_interval : INT_MAX;
        const int n    = bw * bh;
        pred           = 0;
        eobrun         = 0;
        for (int i = 0; i < n; ++i)
        {
          const int bx = i % bw, by = i / bw;
          const bool ok = progressive ? progressive_block(b, &coef[(size_t)i * 64]) : sequential_block(b, bx, by);
          if (!ok)
            return false;
          if (--todo <= 0)
          {
            b.fill();
            if (b.marker() < 0xD0 || b.marker() > 0xD7)
              break; // a restart is due and none follows: the scan ends here
            b.restart();
            pred   = 0;
            eobrun = 0;
            todo   = restart_interval ? restart_interval : INT_MAX;
          }
        }
        pos    = b.pos();
        marker = b.marker();
        return true;
      }

      // After a scan that ended without meeting a marker: skip bytes up to
      // the next marker (0xFF followed by neither 0x00 nor 0xFF).
      int skip_to_marker()
      {
        while (!at_end())
        {
          int x = get8();
          while (x == 0xFF)
          {
            if (at_end())
              return kNone;
            x = get8();
            if (x != 0x00 && x != 0xFF)
              return x;
          }
        }
        return kNone;
      }

      std::string decode(Decoded& out)
      {
        if (get_marker() != 0xD8)
          return "not a JPEG file";
        int m = get_marker();
        while (m != 0xC0 && m != 0xC1 && m != 0xC2)
        {
          if (!process_marker(m))
            return "bad or unsupported marker before the frame header";
          m = get_marker();
          while (m == kNone)
          {
            if (at_end())
              return "no frame header";
            m = get_marker();
          }
        }
        progressive = m == 0xC2;
        std::string why = frame_header();
        if (!why.empty())
          return why;
        out.width      = width;
        out.height     = height;
        out.components = ncomp >= 3 ? 3 : 1;
        if (out.components != 1 || out.header_only)
          return "";

        bw = (width + 7) >> 3;
        bh = (height + 7) >> 3;
        if (progressive)
          coef.assign((size_t)bw * bh * 64, 0);
        else
          pix.assign((size_t)bw * 8 * bh * 8, 0);

        m = get_marker();
        while (m != 0xD9)
        {
          if (m == 0xDA)
          {
            why = scan_header();
            if (!why.empty())
              return why;
            if (!scan())
              return "corrupt entropy-coded data";
            if (marker == kNone)
              marker = skip_to_marker();
            m = get_marker();
            if (m >= 0xD0 && m <= 0xD7)
              m = get_marker();
          }
          else if (m == 0xDC) // DNL: must repeat the frame height
          {
            const int ld = get16(), nl = get16();
            if (ld != 4 || nl != height)
              return "bad DNL segment";
            m = get_marker();
          }
          else
          {
            if (!process_marker(m))
              break; // an unknown marker (or none) ends the image
            m = get_marker();
          }
        }

        if (progressive)
        {
          pix.assign((size_t)bw * 8 * bh * 8, 0);
          const uint16_t* q      = quant[comp_tq];
          const size_t    stride = (size_t)bw * 8;
          for (int i = 0; i < bw * bh; ++i)
          {
            int16_t* blk = &coef[(size_t)i * 64];
            for (int k = 0; k < 64; ++k)
              blk[k] = (int16_t)(blk[k] * q[k]);
            idct_block(blk, &pix[(size_t)(i / bw) * 8 * stride + (size_t)(i % bw) * 8], stride);
          }
        }
        out.grey.resize((size_t)width * height);
        const size_t stride = (size_t)bw * 8;
        for (int y = 0; y < height; ++y)
          std::memcpy(&out.grey[(size_t)y * width], &pix[(size_t)y * stride], (size_t)width);
        return "";
      }
    };
  } // namespace

  std::string decode_jpeg(const std::vector<unsigned char>& file, Decoded& out)
  {
    Jpeg j(file);
    return j.decode(out);
  }
} // namespace dtio
