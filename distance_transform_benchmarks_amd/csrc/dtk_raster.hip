// Raster GGDT passes with the exact reference schedule (gfx950).
//
// Reference: dt/src/generalised_distance_transform.cu:17-99 (passes) and
// :101-134 (driver).  A row pass visits x = 1..W-1 (forward) or W-2..0
// (backward); every pixel of column x reads column x-1 as it is after step
// x-1 of the same pass.  The column passes do the same along y.  The CPU
// oracle runs exactly this schedule, so after N iterations GPU and CPU
// agree bit-for-bit (and at convergence for any N).
//
// MI355X mapping.  A pass is split into strips of R = 64 - 2G lines (rows
// for a row pass, columns for a column pass; G = 20, R = 24 for row passes
// with image costs, G = 16, R = 32 otherwise).  One wave holds 64 lanes = its R lines plus G
// ghost lines on each side, and advances in chunks of C = G steps: at step
// j of a chunk, lane l is exact when lanes l-1..l+1 were exact at step j-1,
// so after C steps the R owned lanes are exact even though the outermost
// ghost lanes had no neighbours.  A chunk needs, for
// its ghost lanes, the neighbour strips' final values of the step before it
// (their other values give the same result whenever they are read).  So at
// the end of every chunk a strip publishes the final values of its first
// and last G owned lines to a mailbox, as 64-bit (stamp << 32 | value)
// words, and the neighbours poll those words: one memory round trip per
// chunk, no flag and no fence.  The mailbox is double-buffered by chunk
// parity (a strip is never more than one chunk ahead of its neighbours) and
// stamped with (pass << 16 | chunk + 1).  Strips form a wavefront pipeline;
// all strips of a pass are resident at once (checked by the host).
//
// Edge costs are formed on the fly from the image values of the current and
// previous step (DPP for the neighbouring lines), with the same arithmetic
// as the tile solver: cur = (D(q) + e) + g, one add when lambda is 0 or 1.
#include <hip/hip_runtime.h>

#include "dtk_launch.hpp"
#include "dtk_tile.hpp"

namespace dtk
{
  namespace
  {
    // Mailbox polls kept in flight while a chunk waits for its neighbours:
    // 3 spaced by s_sleep 4 with image costs, 2 spaced by s_sleep 2 at
    // lambda 0.  A poll that arrives before the neighbour's publish then
    // costs one spacing instead of a round trip (on the synthetic corpus
    // 3.5% faster at lambda 0.5, 2.4% at lambda 1 and 2.6% at lambda 0; with
    // image costs 2 or 4 polls gain less and 6 lose, at lambda 0 4 polls
    // lose 16% and a spacing of 8 11%).  Issuing the first poll before the
    // chunk's loads was slower at every lambda.
    __host__ __device__ constexpr int raster_poll_train(int cm) { return cm != 0 ? 3 : 2; }
    __host__ __device__ constexpr int raster_poll_nap(int cm) { return cm != 0 ? 4 : 2; }
    // Ghost lines per side (== steps per chunk) of a pass: 20 for row passes
    // with image costs, 16 otherwise.  A chunk costs a fixed handshake plus a
    // part that grows with its steps.  Row passes with image costs gain from
    // fewer chunks (7-10% per pass on the synthetic corpus); column passes
    // load one row segment per strip and step, and the narrower strips of
    // G = 20 (24 owned lines instead of 32) load a third more of them (2-3%
    // slower per pass), and at lambda 0 G = 20 is 11% slower.  20 is the
    // widest multiple of four (for the row vectors) whose ghost lines still
    // lie in the two neighbour strips.
    __host__ __device__ constexpr int raster_ghost(int cm, bool rows) { return cm != 0 && rows ? 20 : 16; }
    __host__ __device__ constexpr int raster_owned(int cm, bool rows) { return 64 - 2 * raster_ghost(cm, rows); }
    // A strip waits only for its two neighbours, so its ghost lines must lie
    // within them.
    static_assert(raster_ghost(0, true) <= raster_owned(0, true) && raster_ghost(2, true) <= raster_owned(2, true) &&
                    raster_ghost(2, false) <= raster_owned(2, false),
                  "ghost lines must not reach beyond the neighbour strips");
  }

  __global__ void k_raster_init(ImgDesc d, KParams p, uint32_t* err)
  {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= d.W || y >= d.H)
      return;
    float d0;
    if (p.mask_kind == MASK_U8)
      d0 = p.v * (float)(d.mask[(long long)y * d.mask_pitch + x] == 0);
    else
    {
      const float s = reinterpret_cast<const float*>(d.mask + (long long)y * d.mask_pitch)[x];
      if (!(s >= 0.f && s <= 1.f))
        __hip_atomic_fetch_or(err, (uint32_t)ERR_MASK_RANGE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      d0 = p.v * (1.f - s);
    }
    reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d.D) + (long long)y * d.D_pitch)[x] = d0;
  }

  template <int KIND>
  __device__ __forceinline__ uint32_t img_at(const ImgDesc& d, int x, int y)
  {
    return load_pix<KIND>(d.img + (long long)y * d.img_pitch, x);
  }

  __device__ __forceinline__ float& d_at(const ImgDesc& d, int x, int y)
  {
    return reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d.D) + (long long)y * d.D_pitch)[x];
  }

  // Mailbox word of strip s: side 0 = its first G owned lines, side 1 = its
  // last G; par = chunk parity; k = line within the side.
  template <int G>
  __device__ __forceinline__ size_t mail_index(int strip, int side, int par, int k)
  {
    return ((size_t)(strip * 2 + side) * 2 + par) * G + k;
  }

  // Image value of one pixel as the raster passes hold it: uint8 as a float
  // (differences of small integers are exact in float, so |a - b| is the
  // same value as the integer difference), float as float, RGB8 packed.
  template <int KIND>
  struct RasterPix
  {
    using T = float;
  };
  template <>
  struct RasterPix<IMG_RGB8>
  {
    using T = uint32_t;
  };

  template <int KIND>
  __device__ __forceinline__ typename RasterPix<KIND>::T raster_img(const ImgDesc& d, int x, int y)
  {
    const unsigned char* r = d.img + (long long)y * d.img_pitch;
    if constexpr (KIND == IMG_U8)
      return (float)gload<unsigned char>(r, x);
    else if constexpr (KIND == IMG_F32)
      return gload<float>(r, 4ll * x);
    else
      return load_pix<KIND>(r, x);
  }

  // Wave shifts of image values, copied through an opaque move so the DPP
  // is never folded into the subtract that uses it (see opaque_u).
  template <class T>
  __device__ __forceinline__ T shr_img(T v)
  {
    if constexpr (sizeof(T) == 4 && !__is_same(T, float))
      return shr_u(v);
    else
      return __uint_as_float(opaque_u(__float_as_uint(shr_z(v))));
  }
  template <class T>
  __device__ __forceinline__ T shl_img(T v)
  {
    if constexpr (sizeof(T) == 4 && !__is_same(T, float))
      return shl_u(v);
    else
      return __uint_as_float(opaque_u(__float_as_uint(shl_z(v))));
  }

  // Cost term of an edge: CM 2: l_grad * |I(p) - I(q)|; CM 1 (l_grad = 1):
  // |I(p) - I(q)|.  A NaN (float images only) becomes +inf, which never wins
  // a strict compare, as a NaN would not.
  template <int KIND, int CM, class T>
  __device__ __forceinline__ float raster_cost(T a, T b, float l_grad)
  {
    float g;
    if constexpr (KIND == IMG_RGB8)
      g = pix_dist<IMG_RGB8>(a, b);
    else
      g = fabsf(a - b);
    if constexpr (CM == 2)
      g = l_grad * g;
    if constexpr (KIND == IMG_F32)
      g = (g == g) ? g : DTK_INF;
    return g;
  }

  // N consecutive pixels of one row (N = G, a multiple of four), for a
  // whole chunk of a row pass (a lane is a row there, so every scalar access
  // of a step touches its own cache line).  Float rows are only 4-byte
  // aligned, and 16-byte loads and stores need only dword alignment.
  typedef float f4d __attribute__((ext_vector_type(4), aligned(4)));
  typedef const __attribute__((address_space(1))) f4d* gf4d_r;
  typedef __attribute__((address_space(1))) f4d*       gf4d_w;

  template <int N>
  __device__ __forceinline__ void load_f32_row(const unsigned char* base, long long pitch, int y, int x0, float (&v)[N])
  {
    static_assert(N % 4 == 0, "row vectors of whole dwordx4");
    const gf4d_r q = reinterpret_cast<gf4d_r>((gbytes)base + (long long)y * pitch + 4ll * x0);
#pragma unroll
    for (int k = 0; k < N / 4; ++k)
    {
      const f4d t = q[k];
      v[4 * k]     = t.x;
      v[4 * k + 1] = t.y;
      v[4 * k + 2] = t.z;
      v[4 * k + 3] = t.w;
    }
  }

  template <int N>
  __device__ __forceinline__ void store_f32_row(void* base, long long pitch, int y, int x0, const float (&v)[N])
  {
    static_assert(N % 4 == 0, "row vectors of whole dwordx4");
    const gf4d_w q = reinterpret_cast<gf4d_w>((gbytes_w)base + (long long)y * pitch + 4ll * x0);
#pragma unroll
    for (int k = 0; k < N / 4; ++k)
      q[k] = f4d{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
  }

  // Bytes x0 .. x0+N-1 of a uint8 row as floats: the N/4 or N/4 + 1 aligned
  // dwords that hold them, realigned with v_alignbyte.  The last dword is
  // read only when x0 is not 4-aligned, and then it holds byte x0+N-1, so no
  // dword without a pixel of the row is read.
  template <int N>
  __device__ __forceinline__ void load_u8_row(const ImgDesc& d, int y, int x0, float (&v)[N])
  {
    static_assert(N % 4 == 0, "whole dwords");
    constexpr int   M  = N / 4;
    const uintptr_t a  = reinterpret_cast<uintptr_t>(d.img + (long long)y * d.img_pitch + x0);
    const uint32_t  sh = (uint32_t)(a & 3u);
    const auto*     q  = reinterpret_cast<const __attribute__((address_space(1))) uint32_t*>(a - sh);
    uint32_t        w[M + 1];
#pragma unroll
    for (int k = 0; k < M; ++k)
      w[k] = q[k];
    w[M] = 0u;
    if (sh != 0u)
      w[M] = q[M];
#pragma unroll
    for (int k = 0; k < M; ++k)
    {
      const uint32_t b = __builtin_amdgcn_alignbyte(w[k + 1], w[k], sh);
      v[4 * k]         = (float)(b & 255u);
      v[4 * k + 1]     = (float)((b >> 8) & 255u);
      v[4 * k + 2]     = (float)((b >> 16) & 255u);
      v[4 * k + 3]     = (float)(b >> 24);
    }
  }

  // ROWS: row pass (steps along x, lanes are rows); else column pass.
  // CM: 0 = lambda 0 (chamfer costs e_axis / e_diag, the image is not read),
  // 1 = lambda 1 (cur = D(q) + g), 2 = otherwise (cur = (D(q) + e) + g).
  //
  // A chunk runs in three phases.  Its image values, its distances and the
  // first mailbox poll are issued first; the edge costs of all its steps
  // (which depend on the image only) are formed while the distances and the
  // poll are in flight; the step chain then needs two adds per candidate and
  // two minima per step, without branches.  Lanes outside the image, and
  // the steps of the pass's last chunk past its end, get +inf costs (+inf
  // chamfer steps at lambda 0), so they leave D at +inf or unchanged.  Ghost
  // lanes take no further masks: the outermost lanes (0 and 63) read a zero
  // from the wave shift, but a wrong value on a ghost lane moves in by one
  // lane per step and a chunk has G steps, so owned lanes never see it.
  template <int KIND, bool ROWS, bool FWD, bool C8, int CM>
  __global__ void __launch_bounds__(64) k_raster_pass(ImgDesc d, KParams p, uint32_t* changed,
                                                      unsigned long long* mail, uint32_t epoch)
  {
    using T                 = typename RasterPix<KIND>::T;
    constexpr int kGhost    = raster_ghost(CM, ROWS);
    constexpr int kStripR   = raster_owned(CM, ROWS);
    const int lane     = threadIdx.x;
    const int strip    = blockIdx.x;
    const int nlines   = ROWS ? d.H : d.W; // lines (lanes) dimension
    const int nsteps   = ROWS ? d.W : d.H; // step dimension
    const int line     = strip * kStripR - kGhost + lane;
    const bool lin     = line >= 0 && line < nlines;
    const bool own     = lane >= kGhost && lane < kGhost + kStripR && lin;
    const int  nstrips = gridDim.x;

    const int len   = nsteps - 1; // steps of the pass
    const int start = FWD ? 1 : nsteps - 2;
    const int dir   = FWD ? 1 : -1;
    if (len <= 0)
      return;
    const int nchunks = (len + kGhost - 1) / kGhost;

    const float ea = p.e_axis;
    // e_diag in a vector register, so the shifted distance and e_diag can
    // form one DPP add (DPP takes no scalar operand).
    float ed;
    asm volatile("v_mov_b32 %0, %1" : "=v"(ed) : "s"(p.e_diag));
    const uint64_t own_bits = __ballot(own);
    uint64_t       ch       = 0;

    // Values of the previous step (the step before the chunk's first).
    float Dprev = DTK_INF;
    T     Iprev = T(0);

    // The chunk's own steps are loaded before waiting: no other strip
    // stores an owned line's pixels, and a ghost line read before or after
    // its owner stored this chunk gives the same result (the owner's value
    // is min(old value, the relaxations recomputed here)).  (Loading them a
    // chunk ahead is slower: loads complete in order, so the mailbox poll
    // would wait for the whole next chunk's loads.)
    const int cl = min(max(line, 0), nlines - 1);
    for (int c = 0; c < nchunks; ++c)
    {
      const int k0 = c * kGhost;
      const int kn = min(kGhost, len - k0);

      // Mailbox word of a ghost lane: lanes 0..G-1 are the last G owned
      // lines of strip - 1 (its side 1), lanes R+G..63 the first G owned
      // lines of strip + 1 (its side 0).
      const bool         lower = lane < kGhost && strip > 0;
      const bool         upper = lane >= kGhost + kStripR && strip + 1 < nstrips;
      const bool         need  = c > 0 && (lower || upper) && lin;
      const size_t       at    = lower ? mail_index<kGhost>(strip - 1, 1, c & 1, lane)
                                       : mail_index<kGhost>(strip + 1, 0, c & 1, max(lane - kGhost - kStripR, 0));
      const uint32_t     want  = epoch | (uint32_t)c;
      unsigned long long w     = 0;

      T     Iv[kGhost];
      float Dv[kGhost];
      // A whole chunk of a row pass covers pixels xlo .. xlo+G-1 of each
      // lane's row, in step order for a forward pass and reversed for a
      // backward one; it is loaded (and stored) as row vectors.
      const bool rowvec = ROWS && kn == kGhost;
      const int  xlo    = FWD ? start + k0 : start - k0 - (kGhost - 1);
      if constexpr (CM != 0)
      {
        if (ROWS && KIND != IMG_RGB8 && rowvec)
        {
          float t[kGhost];
          if constexpr (KIND == IMG_U8)
            load_u8_row<kGhost>(d, cl, xlo, t);
          else
            load_f32_row<kGhost>(d.img, d.img_pitch, cl, xlo, t);
#pragma unroll
          for (int j = 0; j < kGhost; ++j)
            Iv[j] = (T)t[FWD ? j : kGhost - 1 - j];
        }
        else
        {
#pragma unroll
          for (int j = 0; j < kGhost; ++j)
          {
            const int s = start + dir * (k0 + min(j, kn - 1));
            Iv[j]       = raster_img<KIND>(d, ROWS ? s : cl, ROWS ? cl : s);
          }
        }
      }
      if (rowvec)
      {
        float t[kGhost];
        load_f32_row<kGhost>(reinterpret_cast<const unsigned char*>(d.D), d.D_pitch, cl, xlo, t);
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
          Dv[j] = lin ? t[FWD ? j : kGhost - 1 - j] : DTK_INF;
      }
      else
      {
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
        {
          const int   s  = start + dir * (k0 + min(j, kn - 1));
          const int   x = ROWS ? s : cl, y = ROWS ? cl : s;
          const float dv = gload<float>(d.D, (long long)y * d.D_pitch + 4ll * x);
          Dv[j]          = (j < kn && lin) ? dv : DTK_INF;
        }
      }

      // Previous step: from memory in the first chunk.  Afterwards every
      // lane keeps its image value (the last step's), owned lanes keep their
      // distance, and ghost lanes take the neighbour's final distance of the
      // step before this chunk from its mailbox.
      if (c == 0)
      {
        const int s = start + dir * (k0 - 1);
        // Clamped unconditional loads, masked afterwards.
        const int   x = ROWS ? s : cl, y = ROWS ? cl : s;
        const float dv = gload<float>(d.D, (long long)y * d.D_pitch + 4ll * x);
        Dprev          = lin ? dv : DTK_INF;
        if constexpr (CM != 0)
          Iprev = raster_img<KIND>(d, x, y);
      }
      else if (need)
        w = __hip_atomic_load(&mail[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

      // Edge costs of the chunk: centre, left (line - 1) and right
      // (line + 1) neighbour of the previous step; at lambda 0 the chamfer
      // steps.  +inf on lanes outside the image and past the pass's end.
      float Gc[kGhost], Gl[kGhost], Gr[kGhost];
      {
        T ip = Iprev;
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
        {
          const bool v = lin && j < kn;
          if constexpr (CM == 0)
          {
            Gc[j] = v ? ea : DTK_INF;
            Gl[j] = Gr[j] = v ? p.e_diag : DTK_INF;
          }
          else
          {
            const T iv = Iv[j];
            Gc[j]      = v ? raster_cost<KIND, CM>(iv, ip, p.l_grad) : DTK_INF;
            if constexpr (C8)
            {
              // Shifted outside the selects: an opaque move cannot be
              // evaluated speculatively, so inside them it would branch.
              const T il = shr_img(ip), ir = shl_img(ip);
              Gl[j]      = v ? raster_cost<KIND, CM>(iv, il, p.l_grad) : DTK_INF;
              Gr[j]      = v ? raster_cost<KIND, CM>(iv, ir, p.l_grad) : DTK_INF;
            }
            ip = iv;
          }
        }
        if constexpr (CM != 0)
          Iprev = Iv[kGhost - 1];
      }

      if (c > 0)
      {
        bool  have = !need || (uint32_t)(w >> 32) == want;
        float got  = __uint_as_float((uint32_t)w);
        // A train of polls: one or two more loads of the word in flight, each
        // checked in turn (the oldest first, so each check waits for one load
        // only) and reissued.
        static_assert(raster_poll_train(CM) > 1, "the wait keeps at least one more poll in flight");
        if (__ballot(!have) != 0)
        {
          constexpr int      kTrain = raster_poll_train(CM) - 1;
          unsigned long long q[kTrain];
#pragma unroll
          for (int k = 0; k < kTrain; ++k)
          {
            __builtin_amdgcn_s_sleep(raster_poll_nap(CM));
            q[k] = have ? 0ull : __hip_atomic_load(&mail[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          for (uint32_t spin = 0; __ballot(!have) != 0; spin += kTrain)
          {
            if (spin >= kSpinLimit)
            {
              // changed[1] is the error word of a raster call.
              if (lane == 0)
                __hip_atomic_fetch_or(&changed[1], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
#pragma unroll
            for (int k = 0; k < kTrain; ++k)
            {
              if (!have && (uint32_t)(q[k] >> 32) == want)
              {
                have = true;
                got  = __uint_as_float((uint32_t)q[k]);
              }
              __builtin_amdgcn_s_sleep(raster_poll_nap(CM));
              q[k] = have ? 0ull : __hip_atomic_load(&mail[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
        if (need)
          Dprev = got;
      }

#pragma unroll
      for (int j = 0; j < kGhost; ++j)
      {
        // CM 0 and 1 add one term (the chamfer step, or g with e = 0).
        float nd = minf_nn(Dv[j], CM == 2 ? (Dprev + ea) + Gc[j] : Dprev + Gc[j]);
        if constexpr (C8)
        {
          const float dl = shr_z(Dprev), dr = shl_z(Dprev);
          nd = min3f(nd, CM == 2 ? (dl + ed) + Gl[j] : dl + Gl[j], CM == 2 ? (dr + ed) + Gr[j] : dr + Gr[j]);
        }
        ch |= __ballot(lt_nn(nd, Dv[j]));
        Dv[j] = nd;
        Dprev = nd;
      }

      // Store the owned lines of the chunk and publish it.
      if (own && rowvec)
      {
        float t[kGhost];
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
          t[FWD ? j : kGhost - 1 - j] = Dv[j];
        store_f32_row<kGhost>(d.D, d.D_pitch, line, xlo, t);
      }
      else if (own)
      {
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
        {
          if (j < kn)
          {
            const int s = start + dir * (k0 + j);
            const int x = ROWS ? s : line, y = ROWS ? line : s;
            // Neighbouring strips load these pixels as ghost lines, and the
            // value before or after this store gives them the same result,
            // so no ordering is needed; a device-scope store and a plain one
            // measured the same.
            __hip_atomic_store(&d_at(d, x, y), Dv[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
      // Publish the chunk's final values of the first and last G owned
      // lines (Dprev holds the last step's value; every chunk but the last
      // has G steps).
      // First G owned lanes: [G, 2G) -> side 0; last G: [R, R + G) -> side 1.
      // With 2G > R the two sets overlap (G = 20, R = 24: lanes 24..39), and
      // a lane in both publishes to both sides.
      const bool first = lane >= kGhost && lane < 2 * kGhost;
      const bool last  = lane >= kStripR && lane < kStripR + kGhost;
      if (c + 1 < nchunks)
      {
        const unsigned long long wv = ((unsigned long long)(epoch | (uint32_t)(c + 1)) << 32) | __float_as_uint(Dprev);
        if (first)
          __hip_atomic_store(&mail[mail_index<kGhost>(strip, 0, (c + 1) & 1, lane - kGhost)], wv, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        if (last)
          __hip_atomic_store(&mail[mail_index<kGhost>(strip, 1, (c + 1) & 1, lane - kStripR)], wv, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if ((ch & own_bits) != 0 && lane == 0)
      __hip_atomic_store(changed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // pass0: index of the iteration's first pass within the call (1-based,
  // pass0 + 3 < 65536); the passes stamp their mailbox words with it.
  template <int KIND, bool C8, int CM>
  static hipError_t raster_iter_t(const ImgDesc& d, const KParams& p, uint32_t* changed, unsigned long long* mail,
                                  uint32_t pass0, hipStream_t s)
  {
    constexpr int RR = raster_owned(CM, true), RC = raster_owned(CM, false);
    const int     rs = (d.H + RR - 1) / RR; // row-pass strips
    const int     cs = (d.W + RC - 1) / RC; // column-pass strips
    hipLaunchKernelGGL((k_raster_pass<KIND, true, true, C8, CM>), dim3(rs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 0u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, true, false, C8, CM>), dim3(rs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 1u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, false, true, C8, CM>), dim3(cs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 2u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, false, false, C8, CM>), dim3(cs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 3u) << 16);
    return hipGetLastError();
  }

  // Cost mode of a raster pass: 0 = lambda 0, 1 = lambda 1, 2 = otherwise.
  static int raster_cm(const KParams& p) { return p.zero_lambda ? 0 : p.l_grad == 1.f ? 1 : 2; }

  template <int KIND, bool C8>
  static hipError_t raster_iter_c(const ImgDesc& d, const KParams& p, uint32_t* changed, unsigned long long* mail,
                                  uint32_t pass0, hipStream_t s)
  {
    switch (raster_cm(p))
    {
    case 0:
      return raster_iter_t<KIND, C8, 0>(d, p, changed, mail, pass0, s);
    case 1:
      return raster_iter_t<KIND, C8, 1>(d, p, changed, mail, pass0, s);
    default:
      return raster_iter_t<KIND, C8, 2>(d, p, changed, mail, pass0, s);
    }
  }

  template <int KIND>
  static hipError_t raster_iter_k(const ImgDesc& d, const KParams& p, bool c8, uint32_t* changed,
                                  unsigned long long* mail, uint32_t pass0, hipStream_t s)
  {
    return c8 ? raster_iter_c<KIND, true>(d, p, changed, mail, pass0, s)
              : raster_iter_c<KIND, false>(d, p, changed, mail, pass0, s);
  }

  hipError_t launch_raster_init(const ImgDesc& d, const KParams& p, uint32_t* err, hipStream_t s)
  {
    if (d.W <= 0 || d.H <= 0)
      return hipSuccess;
    hipLaunchKernelGGL(k_raster_init, dim3((d.W + 255) / 256, d.H), dim3(256), 0, s, d, p, err);
    return hipGetLastError();
  }

  hipError_t launch_raster_iteration(const ImgDesc& d, const KParams& p, bool c8, uint32_t* changed,
                                     unsigned long long* mail, uint32_t pass0, hipStream_t s)
  {
    switch (p.img_kind)
    {
    case IMG_U8:
      return raster_iter_k<IMG_U8>(d, p, c8, changed, mail, pass0, s);
    case IMG_RGB8:
      return raster_iter_k<IMG_RGB8>(d, p, c8, changed, mail, pass0, s);
    default:
      return raster_iter_k<IMG_F32>(d, p, c8, changed, mail, pass0, s);
    }
  }

  template <int KIND, bool C8, int CM>
  static int raster_occ_t()
  {
    int n = 1 << 30, k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, true, true, C8, CM>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, true, false, C8, CM>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, false, true, C8, CM>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, false, false, C8, CM>, 64, 0) == hipSuccess)
      n = min(n, k);
    return n == (1 << 30) ? 0 : n;
  }

  template <int KIND, bool C8>
  static int raster_occ_c(int cm)
  {
    return cm == 0 ? raster_occ_t<KIND, C8, 0>() : cm == 1 ? raster_occ_t<KIND, C8, 1>() : raster_occ_t<KIND, C8, 2>();
  }

  template <int KIND>
  static int raster_occ_k(bool c8, int cm)
  {
    return c8 ? raster_occ_c<KIND, true>(cm) : raster_occ_c<KIND, false>(cm);
  }

  int raster_occupancy(const KParams& p, bool c8)
  {
    const int cm = raster_cm(p);
    switch (p.img_kind)
    {
    case IMG_U8:
      return raster_occ_k<IMG_U8>(c8, cm);
    case IMG_RGB8:
      return raster_occ_k<IMG_RGB8>(c8, cm);
    default:
      return raster_occ_k<IMG_F32>(c8, cm);
    }
  }

  size_t raster_strips(const ImgDesc& d, const KParams& p)
  {
    const int cm = raster_cm(p);
    const int RR = raster_owned(cm, true), RC = raster_owned(cm, false);
    const int rs = (d.H + RR - 1) / RR;
    const int cs = (d.W + RC - 1) / RC;
    return (size_t)(rs > cs ? rs : cs);
  }

  size_t raster_mail_words(const ImgDesc& d, const KParams& p)
  {
    const int    cm = raster_cm(p);
    const int    RR = raster_owned(cm, true), RC = raster_owned(cm, false);
    const size_t wr = (size_t)((d.H + RR - 1) / RR) * 2 * 2 * raster_ghost(cm, true);
    const size_t wc = (size_t)((d.W + RC - 1) / RC) * 2 * 2 * raster_ghost(cm, false);
    return wr > wc ? wr : wc;
  }
} // namespace dtk
