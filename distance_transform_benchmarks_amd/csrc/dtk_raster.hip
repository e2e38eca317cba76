// Raster GGDT passes with the exact reference schedule (gfx950).
//
// Reference: dt/src/generalised_distance_transform.cu:17-99 (passes) and
// :101-134 (driver).  A row pass visits x = 1..W-1 (forward) or W-2..0
// (backward); every pixel of column x reads column x-1 as it is after step
// x-1 of the same pass.  The column passes do the same along y.  The CPU
// oracle runs exactly this schedule, so after N iterations GPU and CPU
// agree bit-for-bit (and at convergence for any N).
//
// MI355X mapping.  A pass is split into strips of R = 32 lines (rows for a
// row pass, columns for a column pass).  One wave holds 64 lanes = its 32
// lines plus G = 16 ghost lines on each side, and advances in chunks of
// C = G steps: at step j of a chunk, lane l is exact when lanes l-1..l+1
// were exact at step j-1, so after C steps the 32 owned lanes are exact even
// though the outermost ghost lanes had no neighbours.  A chunk needs, for
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
#ifndef DTK_RASTER_GHOST
#define DTK_RASTER_GHOST 16
#endif
    constexpr int kGhost  = DTK_RASTER_GHOST; // ghost lines per side == steps per chunk
    constexpr int kStripR = 64 - 2 * kGhost;  // owned lines per strip
    // A strip waits only for its two neighbours, so its ghost lines must lie
    // within them.
    static_assert(kGhost <= kStripR, "ghost lines must not reach beyond the neighbour strips");
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
  __device__ __forceinline__ size_t mail_index(int strip, int side, int par, int k)
  {
    return ((size_t)(strip * 2 + side) * 2 + par) * kGhost + k;
  }

  // ROWS: row pass (steps along x, lanes are rows); else column pass.
  template <int KIND, bool ROWS, bool FWD, bool C8, bool TWO>
  __global__ void __launch_bounds__(64) k_raster_pass(ImgDesc d, KParams p, uint32_t* changed,
                                                      unsigned long long* mail, uint32_t epoch)
  {
    const int lane   = threadIdx.x;
    const int strip  = blockIdx.x;
    const int nlines = ROWS ? d.H : d.W; // lines (lanes) dimension
    const int nsteps = ROWS ? d.W : d.H; // step dimension
    const int line   = strip * kStripR - kGhost + lane;
    const bool lin   = line >= 0 && line < nlines;
    const bool own   = lane >= kGhost && lane < kGhost + kStripR && lin;
    const bool linL  = line - 1 >= 0 && line - 1 < nlines;
    const bool linR  = line + 1 >= 0 && line + 1 < nlines;
    const int  nstrips = gridDim.x;

    const int len   = nsteps - 1;         // steps of the pass
    const int start = FWD ? 1 : nsteps - 2;
    const int dir   = FWD ? 1 : -1;
    if (len <= 0)
      return;
    const int nchunks = (len + kGhost - 1) / kGhost;

    const float ea = p.e_axis, ed = p.e_diag;
    bool        any = false;

    // Values of the previous step (the step before the chunk's first).
    float    Dprev = DTK_INF;
    uint32_t Iprev = 0;

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

      float    Dv[kGhost];
      uint32_t Iv[kGhost];
#pragma unroll
      for (int j = 0; j < kGhost; ++j)
      {
        const int      s  = start + dir * (k0 + min(j, kn - 1));
        const int      x = ROWS ? s : cl, y = ROWS ? cl : s;
        const float    dv = gload<float>(d.D, (long long)y * d.D_pitch + 4ll * x);
        const uint32_t iv = img_at<KIND>(d, x, y);
        Dv[j]             = (j < kn && lin) ? dv : DTK_INF;
        Iv[j]             = (j < kn && lin) ? iv : 0u;
      }

      // Previous step: from memory in the first chunk.  Afterwards every
      // lane keeps its image value (the last step's), owned lanes keep their
      // distance, and ghost lanes take the neighbour's final distance of the
      // step before this chunk from its mailbox.
      if (c == 0)
      {
        const int s = start + dir * (k0 - 1);
        // Clamped unconditional loads, masked afterwards.
        const int      x = ROWS ? s : cl, y = ROWS ? cl : s;
        const float    dv = gload<float>(d.D, (long long)y * d.D_pitch + 4ll * x);
        const uint32_t iv = img_at<KIND>(d, x, y);
        Dprev             = lin ? dv : DTK_INF;
        Iprev             = lin ? iv : 0u;
      }
      else
      {
        // Lanes 0..G-1 are the last G owned lines of strip - 1 (its side 1),
        // lanes R+G..63 the first G owned lines of strip + 1 (its side 0).
        const bool lower = lane < kGhost && strip > 0;
        const bool upper = lane >= kGhost + kStripR && strip + 1 < nstrips;
        const bool need  = (lower || upper) && lin;
        const size_t   at = lower ? mail_index(strip - 1, 1, c & 1, lane)
                                  : mail_index(strip + 1, 0, c & 1, max(lane - kGhost - kStripR, 0));
        const uint32_t want = epoch | (uint32_t)c;
        bool           have = !need;
        float          got  = DTK_INF;
        for (uint32_t spin = 0;; ++spin)
        {
          if (!have)
          {
            const unsigned long long w = __hip_atomic_load(&mail[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(w >> 32) == want)
            {
              have = true;
              got  = __uint_as_float((uint32_t)w);
            }
          }
          if (__ballot(!have) == 0)
            break;
          if (spin >= kSpinLimit)
          {
            // changed[1] is the error word of a raster call.
            if (lane == 0)
              __hip_atomic_fetch_or(&changed[1], (uint32_t)ERR_WATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (need)
          Dprev = got;
      }
#pragma unroll
      for (int j = 0; j < kGhost; ++j)
      {
        if (j < kn)
        {
          const uint32_t ip = Iv[j];
          float          nd = Dv[j];
          {
            float g = p.l_grad * pix_dist<KIND>(ip, Iprev);
            g       = (g == g) ? g : DTK_INF;
            const float cost = p.zero_lambda ? ea : g;
            const float cur  = TWO ? (Dprev + ea) + cost : Dprev + cost;
            nd               = minf_nn(nd, cur);
          }
          if constexpr (C8)
          {
            const float    dl = shr_z(Dprev), dr = shl_z(Dprev);
            const uint32_t il = shr_u(Iprev), ir = shl_u(Iprev);
            float          gl = p.l_grad * pix_dist<KIND>(ip, il);
            float          gr = p.l_grad * pix_dist<KIND>(ip, ir);
            gl                = (gl == gl) ? gl : DTK_INF;
            gr                = (gr == gr) ? gr : DTK_INF;
            const float cl    = p.zero_lambda ? ed : gl;
            const float cr    = p.zero_lambda ? ed : gr;
            float       curl  = TWO ? (dl + ed) + cl : dl + cl;
            float       curr  = TWO ? (dr + ed) + cr : dr + cr;
            curl              = (linL && lane > 0) ? curl : DTK_INF;
            curr              = (linR && lane < 63) ? curr : DTK_INF;
            nd                = min3f(nd, curl, curr);
          }
          if (lin && lt_nn(nd, Dv[j]))
          {
            Dv[j] = nd;
            any |= own;
          }
          Dprev = Dv[j];
          Iprev = ip;
        }
      }

      // Store the owned lines of the chunk and publish it.
      if (own)
      {
#pragma unroll
        for (int j = 0; j < kGhost; ++j)
        {
          if (j < kn)
          {
            const int s = start + dir * (k0 + j);
            const int x = ROWS ? s : line, y = ROWS ? line : s;
            // Device-scope store: written through, so the release fence
            // below finds no dirty L2 lines to write back.
            __hip_atomic_store(&d_at(d, x, y), Dv[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
      // Publish the chunk's final values of the first and last G owned
      // lines (Dprev holds the last step's value).
      // First G owned lanes: [G, 2G) -> side 0; last G: [R, R + G) -> side 1.
      const bool first = lane >= kGhost && lane < 2 * kGhost;
      const bool last  = lane >= kStripR && lane < kStripR + kGhost;
      if (c + 1 < nchunks && (first || last))
      {
        const unsigned long long w = ((unsigned long long)(epoch | (uint32_t)(c + 1)) << 32) | __float_as_uint(Dprev);
        __hip_atomic_store(&mail[mail_index(strip, first ? 0 : 1, (c + 1) & 1, first ? lane - kGhost : lane - kStripR)],
                           w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (__ballot(any) != 0 && lane == 0)
      __hip_atomic_store(changed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // pass0: index of the iteration's first pass within the call (1-based,
  // pass0 + 3 < 65536); the passes stamp their mailbox words with it.
  template <int KIND, bool C8, bool TWO>
  static hipError_t raster_iter_t(const ImgDesc& d, const KParams& p, uint32_t* changed, unsigned long long* mail,
                                  uint32_t pass0, hipStream_t s)
  {
    const int rs = (d.H + kStripR - 1) / kStripR; // row-pass strips
    const int cs = (d.W + kStripR - 1) / kStripR; // column-pass strips
    hipLaunchKernelGGL((k_raster_pass<KIND, true, true, C8, TWO>), dim3(rs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 0u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, true, false, C8, TWO>), dim3(rs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 1u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, false, true, C8, TWO>), dim3(cs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 2u) << 16);
    hipLaunchKernelGGL((k_raster_pass<KIND, false, false, C8, TWO>), dim3(cs), dim3(64), 0, s, d, p, changed, mail,
                       (pass0 + 3u) << 16);
    return hipGetLastError();
  }

  template <int KIND>
  static hipError_t raster_iter_k(const ImgDesc& d, const KParams& p, bool c8, uint32_t* changed,
                                  unsigned long long* mail, uint32_t pass0, hipStream_t s)
  {
    const bool two = !(p.zero_lambda || p.l_grad == 1.f);
    if (c8)
      return two ? raster_iter_t<KIND, true, true>(d, p, changed, mail, pass0, s)
                 : raster_iter_t<KIND, true, false>(d, p, changed, mail, pass0, s);
    return two ? raster_iter_t<KIND, false, true>(d, p, changed, mail, pass0, s)
               : raster_iter_t<KIND, false, false>(d, p, changed, mail, pass0, s);
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

  template <int KIND, bool C8, bool TWO>
  static int raster_occ_t()
  {
    int n = 1 << 30, k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, true, true, C8, TWO>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, true, false, C8, TWO>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, false, true, C8, TWO>, 64, 0) == hipSuccess)
      n = min(n, k);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_raster_pass<KIND, false, false, C8, TWO>, 64, 0) == hipSuccess)
      n = min(n, k);
    return n == (1 << 30) ? 0 : n;
  }

  template <int KIND>
  static int raster_occ_k(bool c8, bool two)
  {
    if (c8)
      return two ? raster_occ_t<KIND, true, true>() : raster_occ_t<KIND, true, false>();
    return two ? raster_occ_t<KIND, false, true>() : raster_occ_t<KIND, false, false>();
  }

  int raster_occupancy(const KParams& p, bool c8)
  {
    const bool two = !(p.zero_lambda || p.l_grad == 1.f);
    switch (p.img_kind)
    {
    case IMG_U8:
      return raster_occ_k<IMG_U8>(c8, two);
    case IMG_RGB8:
      return raster_occ_k<IMG_RGB8>(c8, two);
    default:
      return raster_occ_k<IMG_F32>(c8, two);
    }
  }

  size_t raster_strips(const ImgDesc& d)
  {
    const int rs = (d.H + kStripR - 1) / kStripR;
    const int cs = (d.W + kStripR - 1) / kStripR;
    return (size_t)(rs > cs ? rs : cs);
  }

  size_t raster_mail_words(const ImgDesc& d) { return raster_strips(d) * 2 * 2 * kGhost; }
} // namespace dtk
