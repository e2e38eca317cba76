// Argmin predecessors and the adjoint of the GGDT (gfx950).  No reference
// counterpart: the reference GGDT is forward-only.
//
// Forward (converged D): pred(p) = 8 when D(p) == D0(p) (a root), else the
// first neighbour q in dt::c8 order (dt/include/dt/c8.hpp:19-28) whose
// relaxation (D(q) + e) + g reproduces D(p) exactly and whose (D, index)
// pair is lexicographically below p's, which keeps the forest acyclic.
//
// Backward of L = sum(W * D), rounding treated as exact:
//   G(p) = W(p) + sum over children c of G(c)         (subtree sums)
//   edge q -> p : dL/dI(p) += G(p) * l_grad * sign(I(p) - I(q)), dL/dI(q) -= same
//   root r      : dL/ds(r) = -nu * G(r)               (D0 = nu * (1 - s))
// Subtree sums in three phases over 32 x 32 tiles, so that the pointer
// chasing runs in LDS (a climb through global memory spends a memory
// latency, and a 64-byte transaction, on every 1-byte step):
//   A  per tile: local subtree sums of W over the tile's own edges, to G; a
//      pixel whose parent lies in another tile is an exit, and its local
//      sum starts its coarse word in cw; for every tile-border pixel, the
//      exit (or root) that its in-tile path ends at, to xp;
//   B  over the exits only: the same climb on the coarse forest in which an
//      exit's parent is the exit that its real parent's in-tile path ends
//      at; an exit's coarse word ends as its full subtree sum C;
//   C  per tile: each border pixel's inflow (C summed over its children in
//      other tiles) is added along its in-tile path, which completes G.
// The climbs of A and B are leaf-started: a pixel's count of children
// still to arrive and the sum of their subtree sums (a double) share one
// 64-bit word, updated by compare-and-swap, and the last child to arrive
// continues with the complete sum (read from the value it wrote itself).  Sums and the
// gradients built from them are kept in double; a subtree can hold a
// million pixels, and a gradient is a difference of such sums.  The coarse
// chains are about 64 times shorter than the pixel chains, and the tile
// climbs run on LDS latencies.  The kernels take a table of images
// (blockIdx.z), so a whole batch runs in one launch of each.
#include <hip/hip_runtime.h>

#include "dtk_launch.hpp"
#include "dtk_tile.hpp"

namespace dtk
{
  // dt::c8 offsets of predecessor code k, packed like the tile neighbours'.
  constexpr uint32_t kOffPackX = 0x1A90u, kOffPackY = 0xA901u;
  __host__ __device__ constexpr int off_x(int k) { return unpack_off(kOffPackX, k); }
  __host__ __device__ constexpr int off_y(int k) { return unpack_off(kOffPackY, k); }
  static_assert(off_x(0) == -1 && off_x(1) == -1 && off_x(2) == 0 && off_x(3) == 1 && off_x(4) == 1 &&
                  off_x(5) == 1 && off_x(6) == 0 && off_x(7) == -1,
                "dt::c8 x offsets");
  static_assert(off_y(0) == 0 && off_y(1) == -1 && off_y(2) == -1 && off_y(3) == -1 && off_y(4) == 0 &&
                  off_y(5) == 1 && off_y(6) == 1 && off_y(7) == 1,
                "dt::c8 y offsets");

  __device__ __forceinline__ float rd_f(const unsigned char* base, long long pitch, int x, int y)
  {
    return reinterpret_cast<const float*>(base + (long long)y * pitch)[x];
  }

  template <int KIND>
  __device__ __forceinline__ void pred_pixel(const ImgDesc& d, const KParams& p, int c8, int x, int y)
  {
    if (x >= d.W || y >= d.H)
      return;
    const unsigned char* Db = reinterpret_cast<const unsigned char*>(d.D);
    const float          Dp = rd_f(Db, d.D_pitch, x, y);
    float                d0;
    if (p.mask_kind == MASK_U8)
      d0 = p.v * (float)(d.mask[(long long)y * d.mask_pitch + x] == 0);
    else
      d0 = p.v * (1.f - rd_f(d.mask, d.mask_pitch, x, y));
    unsigned char code = 8;
    if (!(Dp == d0))
    {
      // Every neighbour's D and pixel are loaded first, from clamped (in
      // bounds) addresses, so the loads are in flight together; then the
      // neighbours are tested in dt::c8 order and the first match wins.
      const uint32_t  ip  = load_pix<KIND>(d.img + (long long)y * d.img_pitch, x);
      const long long pid = (long long)y * d.W + x;
      float           Dq[8];
      uint32_t        iq[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        const int nx = min(max(x + off_x(k), 0), d.W - 1), ny = min(max(y + off_y(k), 0), d.H - 1);
        Dq[k]        = rd_f(Db, d.D_pitch, nx, ny);
        iq[k]        = load_pix<KIND>(d.img + (long long)ny * d.img_pitch, nx);
      }
#pragma unroll
      for (int k = 7; k >= 0; --k)
      {
        const bool diag = off_x(k) != 0 && off_y(k) != 0;
        const int  nx = x + off_x(k), ny = y + off_y(k);
        const bool in = (!diag || c8) && nx >= 0 && ny >= 0 && nx < d.W && ny < d.H;
        const long long qid = (long long)ny * d.W + nx;
        // lambda = 0 has no image term (the forward kernels do not read the
        // image), so a NaN or infinite float pixel changes nothing there.
        float           g   = p.l_grad == 0.f ? 0.f : p.l_grad * pix_dist<KIND>(ip, iq[k]);
        g                   = (g == g) ? g : DTK_INF;
        float cur           = Dq[k] + (diag ? p.e_diag : p.e_axis);
        cur                 = cur + g;
        // Walking k downwards, the last match kept is the first in c8 order.
        if (in && (Dq[k] < Dp || (Dq[k] == Dp && qid < pid)) && cur == Dp)
          code = (unsigned char)k;
      }
    }
    d.pred[(long long)y * d.pred_pitch + x] = code;
  }

  template <int KIND>
  __global__ void k_pred(ImgDesc d, KParams p, int c8)
  {
    pred_pixel<KIND>(d, p, c8, (int)(blockIdx.x * blockDim.x + threadIdx.x), (int)blockIdx.y);
  }

  // A batch of images in one launch: image blockIdx.z of the table.
  template <int KIND>
  __global__ void k_pred_batched(const ImgDesc* tab, KParams p, int c8)
  {
    pred_pixel<KIND>(tab[blockIdx.z], p, c8, (int)(blockIdx.x * blockDim.x + threadIdx.x), (int)blockIdx.y);
  }

  // ---- tiled subtree sums --------------------------------------------------
#ifndef DTK_BACK_TILE
#define DTK_BACK_TILE 32
#endif
  constexpr int      kBT      = DTK_BACK_TILE; // tile side of the subtree sums (32 or 64)
  constexpr int      kBTPix   = kBT * kBT;
#ifndef DTK_BACK_THREADS
#define DTK_BACK_THREADS 256
#endif
  constexpr int      kBThreads = DTK_BACK_THREADS;  // threads of a tile block (a multiple of kBT)
  constexpr int      kBTRows   = kBThreads / kBT;   // tile rows one pass of the block covers
  static_assert(kBThreads % kBT == 0 && kBTPix % kBThreads == 0, "tile block shape");
  static_assert(4 * kBT - 4 <= 256 && 4 * kBT - 4 <= kBThreads, "one thread per border position");
  constexpr int      kBorder  = 4 * kBT - 4; // border positions of a tile
  constexpr uint32_t kNoExit  = 0xFFFFFFFFu;
  // Sum words: a double whose lowest mantissa bits are replaced by a count
  // (the sum keeps 52 - bits of its 52 fraction bits: relative error below
  // 2^-40, far under the float32 outputs).  Tile words keep the count of
  // children still to arrive (at most 8) in 4 bits; coarse words keep it in
  // 11 bits (an exit has fewer than 4 * 64 * 3 coarse children) and bit 11
  // flags that the exit has any.
  constexpr unsigned long long kTileCount   = 0xFull;
  constexpr unsigned long long kCoarseCount = 0x7FFull;
  constexpr unsigned long long kHasKids     = 0x800ull;
  constexpr unsigned long long kCoarseLow   = 0xFFFull;
  __device__ __forceinline__ double word_sum(unsigned long long w, unsigned long long low)
  {
    return __longlong_as_double((long long)(w & ~low));
  }
  __device__ __forceinline__ unsigned long long sum_bits(double v, unsigned long long low)
  {
    return (unsigned long long)__double_as_longlong(v) & ~low;
  }

  // Border position b (0 .. kBorder-1) of a tile: top row, bottom row, then
  // the left and right columns without their corners.
  __device__ __forceinline__ void border_pos(int b, int& lx, int& ly)
  {
    if (b < kBT)
    {
      lx = b;
      ly = 0;
    }
    else if (b < 2 * kBT)
    {
      lx = b - kBT;
      ly = kBT - 1;
    }
    else if (b < 3 * kBT - 2)
    {
      lx = 0;
      ly = b - 2 * kBT + 1;
    }
    else
    {
      lx = kBT - 1;
      ly = b - (3 * kBT - 2) + 1;
    }
  }

  // Predecessor code of (x, y) in image d: 0..7, or 8 for a root (codes
  // above 7 and codes that point outside the image are roots).
  __device__ __forceinline__ int pred_code(const ImgDesc& d, int x, int y)
  {
    int c = d.pred[(long long)y * d.pred_pitch + x];
    if (c >= 8)
      return 8;
    const int px = x + off_x(c), py = y + off_y(c);
    return (px < 0 || py < 0 || px >= d.W || py >= d.H) ? 8 : c;
  }

  // The tile's pixels: in-image codes (9 outside) and per-pixel values.
  // Thread t owns the tile pixels t, t + kBThreads, ... (column t % kBT).
  struct TileGeom
  {
    int x0, y0, tw, th, lx, ly0;
  };
  __device__ __forceinline__ bool tile_geom(const ImgDesc& d, TileGeom& t)
  {
    t.x0 = (int)blockIdx.x * kBT;
    t.y0 = (int)blockIdx.y * kBT;
    if (t.x0 >= d.W || t.y0 >= d.H)
      return false; // the tile lies past this image (the grid covers the largest)
    t.tw  = min(kBT, d.W - t.x0);
    t.th  = min(kBT, d.H - t.y0);
    t.lx  = (int)threadIdx.x & (kBT - 1);
    t.ly0 = (int)threadIdx.x / kBT;
    return true;
  }

  // Phase A for one tile: local subtree sums of W over the tile's own edges
  // (to G, every pixel; an exit's also to its coarse word in cw) and, for
  // every border pixel, the exit (or root) that its in-tile path ends at
  // (to xp).
  __global__ void __launch_bounds__(kBThreads) k_tile_sums(const BackJob* jobs)
  {
    __shared__ unsigned char      s_code[kBTPix]; // 0..7, 8 = root, 9 = outside the image
    __shared__ double             s_val[kBTPix];  // weight, then the pixel's subtree sum
    __shared__ unsigned long long s_word[kBTPix]; // sum of the arrived children | count still to arrive
    __shared__ unsigned char      s_kids[kBTPix]; // children in the tile (fixed: s_word's count runs down)
    const BackJob& j = jobs[blockIdx.z];
    const ImgDesc& d = j.d;
    TileGeom       tg;
    if (!tile_geom(d, tg))
      return;
    const int x0 = tg.x0, y0 = tg.y0, lx = tg.lx, ly0 = tg.ly0;
    for (int i = 0; i < kBTPix / kBThreads; ++i)
    {
      const int ly = ly0 + kBTRows * i, li = ly * kBT + lx;
      int       c  = 9;
      double    w  = 0.0;
      if (lx < tg.tw && ly < tg.th)
      {
        c = pred_code(d, x0 + lx, y0 + ly);
        w = rd_f(j.gd, j.gd_pitch, x0 + lx, y0 + ly);
      }
      s_code[li] = (unsigned char)c;
      s_val[li]  = w;
    }
    __syncthreads();
    // Children inside the tile: pixel c is a child when its code k points
    // here, c = p - off[k].
    for (int i = 0; i < kBTPix / kBThreads; ++i)
    {
      const int ly = ly0 + kBTRows * i, li = ly * kBT + lx;
      if (s_code[li] == 9)
        continue;
      uint32_t n = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        const int cx = lx - off_x(k), cy = ly - off_y(k);
        if (cx >= 0 && cy >= 0 && cx < kBT && cy < kBT)
          n += s_code[cy * kBT + cx] == k ? 1u : 0u;
      }
      s_word[li] = n; // sum +0.0
      s_kids[li] = (unsigned char)n;
    }
    __syncthreads();
    // Leaf-started climbs, one flat loop per thread: a thread whose climb
    // ends takes its next leaf at once, so a wave runs about as long as its
    // longest climb rather than the sum of each pixel slot's longest.
    {
      int    slot = 0;  // next of the thread's own pixels to look at
      int    p    = -1; // the climb's current pixel
      int    c    = 8;
      double g    = 0.0;
      for (;;)
      {
        if (p < 0)
        {
          while (slot < kBTPix / kBThreads)
          {
            const int li = (ly0 + kBTRows * slot) * kBT + lx;
            ++slot;
            if (s_code[li] != 9 && s_kids[li] == 0)
            {
              p = li;
              c = s_code[li];
              g = s_val[li];
              break;
            }
          }
          if (p < 0)
            break; // no leaf left
        }
        s_val[p] = g; // the pixel's local subtree sum
        if (c >= 8)
        {
          p = -1; // a root
          continue;
        }
        const int px = (p & (kBT - 1)) + off_x(c), py = p / kBT + off_y(c);
        if (px < 0 || py < 0 || px >= kBT || py >= kBT)
        {
          // An exit: its parent lies in another tile.  Its coarse word
          // starts as the local sum with no coarse child.
          j.cw[(long long)(y0 + p / kBT) * d.W + x0 + (p & (kBT - 1))] = sum_bits(g, kCoarseLow);
          p = -1;
          continue;
        }
        const int q = py * kBT + px;
        // The parent's child count, code and weight do not change until it
        // is finished (by the thread that arrives last): loaded together.
        const int    kq = s_kids[q];
        const int    cq = s_code[q];
        const double wq = s_val[q];
        if (kq == 1)
          g = wq + g; // the only child is the last to arrive: no exchange
        else
        {
          // First guess: no sibling has arrived yet (count kq, sum +0.0).
          unsigned long long old = (unsigned long long)kq, nw;
          for (;;)
          {
            nw = sum_bits(word_sum(old, kTileCount) + g, kTileCount) | ((old & kTileCount) - 1ull);
            if (__hip_atomic_compare_exchange_strong(&s_word[q], &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
              break;
          }
          if ((nw & kTileCount) != 0ull)
          {
            p = -1; // a sibling is still climbing: the last one continues
            continue;
          }
          g = wq + word_sum(nw, kTileCount);
        }
        p = q;
        c = cq;
      }
    }
    __syncthreads();
    for (int i = 0; i < kBTPix / kBThreads; ++i)
    {
      const int ly = ly0 + kBTRows * i;
      if (lx < tg.tw && ly < tg.th)
        j.G[(long long)(y0 + ly) * d.W + x0 + lx] = s_val[ly * kBT + lx];
    }
    if ((int)threadIdx.x < kBorder)
    {
      // The end of each border pixel's in-tile path: the exit's image index,
      // or kNoExit when the path ends at a root (or runs in a cycle, which
      // no forest from dtk_ggdt_pred has).
      int bx, by;
      border_pos((int)threadIdx.x, bx, by);
      if (bx < tg.tw && by < tg.th)
      {
        int      p   = by * kBT + bx;
        uint32_t out = kNoExit;
        for (int step = 0; step < kBTPix; ++step)
        {
          const int c = s_code[p];
          if (c >= 8)
            break;
          const int px = (p & (kBT - 1)) + off_x(c), py = (p / kBT) + off_y(c);
          if (px < 0 || py < 0 || px >= kBT || py >= kBT)
          {
            out = (uint32_t)(y0 + p / kBT) * (uint32_t)d.W + (uint32_t)(x0 + (p & (kBT - 1)));
            break;
          }
          p = py * kBT + px;
        }
        j.xp[(long long)(y0 + by) * d.W + x0 + bx] = out;
      }
    }
  }

  // Phase C for one tile: G = local sum + every inflow on the pixel's
  // in-tile path from the border.  A border pixel's inflow In is the sum of
  // C over its children in other tiles (exits, final after phase B); it is
  // added to each pixel from there up to the path's end.
  __global__ void __launch_bounds__(kBThreads) k_tile_inflow(const BackJob* jobs)
  {
    __shared__ unsigned char s_code[kBTPix];
    __shared__ double        s_val[kBTPix];
    const BackJob& j = jobs[blockIdx.z];
    const ImgDesc& d = j.d;
    TileGeom       tg;
    if (!tile_geom(d, tg))
      return;
    const int x0 = tg.x0, y0 = tg.y0, lx = tg.lx, ly0 = tg.ly0;
    for (int i = 0; i < kBTPix / kBThreads; ++i)
    {
      const int ly = ly0 + kBTRows * i, li = ly * kBT + lx;
      int       c  = 9;
      double    v  = 0.0;
      if (lx < tg.tw && ly < tg.th)
      {
        c = pred_code(d, x0 + lx, y0 + ly);
        v = j.G[(long long)(y0 + ly) * d.W + x0 + lx];
      }
      s_code[li] = (unsigned char)c;
      s_val[li]  = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < kBorder)
    {
      int bx, by;
      border_pos((int)threadIdx.x, bx, by);
      double in = 0.0;
      if (bx < tg.tw && by < tg.th)
      {
        // Children outside the tile, in dt::c8 order; their codes first.
        const int x = x0 + bx, y = y0 + by;
        int       cc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
          const int cx = bx - off_x(k), cy = by - off_y(k), gx = x - off_x(k), gy = y - off_y(k);
          const bool outside = cx < 0 || cy < 0 || cx >= kBT || cy >= kBT;
          cc[k] = (outside && gx >= 0 && gy >= 0 && gx < d.W && gy < d.H) ? pred_code(d, gx, gy) : 9;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (cc[k] == k)
            in += word_sum(j.cw[(long long)(y - off_y(k)) * d.W + x - off_x(k)], kCoarseLow);
      }
      if (in != 0.0)
      {
        int p = by * kBT + bx;
        for (int step = 0; step < kBTPix; ++step)
        {
          __hip_atomic_fetch_add(&s_val[p], in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const int c = s_code[p];
          if (c >= 8)
            break;
          const int px = (p & (kBT - 1)) + off_x(c), py = (p / kBT) + off_y(c);
          if (px < 0 || py < 0 || px >= kBT || py >= kBT)
            break; // the exit: its parent is in another tile
          p = py * kBT + px;
        }
      }
    }
    __syncthreads();
    for (int i = 0; i < kBTPix / kBThreads; ++i)
    {
      const int ly = ly0 + kBTRows * i;
      if (lx < tg.tw && ly < tg.th)
        j.G[(long long)(y0 + ly) * d.W + x0 + lx] = s_val[ly * kBT + lx];
    }
  }

  // The exit at border position threadIdx.x of this tile, if any: its image
  // index (*e) and the exit its parent's in-tile path ends at (*t, kNoExit
  // when that path ends at a root).
  __device__ __forceinline__ bool exit_at(const BackJob& j, int b, uint32_t* e, uint32_t* t)
  {
    const ImgDesc& d  = j.d;
    const int      x0 = (int)blockIdx.x * kBT, y0 = (int)blockIdx.y * kBT;
    if (b >= kBorder || x0 >= d.W || y0 >= d.H)
      return false;
    int bx, by;
    border_pos(b, bx, by);
    const int x = x0 + bx, y = y0 + by;
    if (x >= d.W || y >= d.H)
      return false;
    const int c = pred_code(d, x, y);
    if (c >= 8)
      return false;
    const int px = bx + off_x(c), py = by + off_y(c);
    if (px >= 0 && py >= 0 && px < kBT && py < kBT)
      return false; // the parent is in this tile
    *e = (uint32_t)y * (uint32_t)d.W + (uint32_t)x;
    *t = j.xp[(long long)(y + off_y(c)) * d.W + x + off_x(c)];
    if (*t != kNoExit && *t >= (uint32_t)d.W * (uint32_t)d.H)
      *t = kNoExit; // phase A writes only image indices here; anything else is not followed
    return true;
  }

  // Phase B, first step: coarse children per exit (count and flag).
  __global__ void __launch_bounds__(256) k_coarse_count(const BackJob* jobs)
  {
    const BackJob& j = jobs[blockIdx.z];
    uint32_t       e, t;
    if (!exit_at(j, (int)threadIdx.x, &e, &t) || t == kNoExit)
      return;
    __hip_atomic_fetch_or(&j.cw[t], kHasKids, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&j.cw[t], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // Phase B, second step: climbs from the coarse leaves.  An exit's word is
  // final once its count is 0: its low half is then C, the exit's full
  // subtree sum.
  __global__ void __launch_bounds__(256) k_coarse_climb(const BackJob* jobs)
  {
    const BackJob& j = jobs[blockIdx.z];
    const ImgDesc& d = j.d;
    uint32_t       e, t;
    if (!exit_at(j, (int)threadIdx.x, &e, &t))
      return;
    const unsigned long long w0 = __hip_atomic_load(&j.cw[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w0 & kHasKids)
      return; // not a coarse leaf (the flag does not change here; the count does)
    double g = word_sum(w0, kCoarseLow);
    for (;;)
    {
      if (t == kNoExit)
        return; // the flow ends at a root, inside the parent's tile (phase C adds it there)
      unsigned long long old = __hip_atomic_load(&j.cw[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned long long nw;
      for (;;)
      {
        nw = sum_bits(word_sum(old, kCoarseLow) + g, kCoarseLow) | (old & kHasKids) |
             (((old & kCoarseCount) - 1ull) & kCoarseCount);
        if (__hip_atomic_compare_exchange_strong(&j.cw[t], &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
          break;
      }
      if ((nw & kCoarseCount) != 0ull)
        return; // a coarse sibling is still climbing
      g = word_sum(nw, kCoarseLow);
      // Continue from exit t: its parent, then the exit that parent's path ends at.
      const int x = (int)(t % (uint32_t)d.W), y = (int)(t / (uint32_t)d.W);
      const int c = pred_code(d, x, y); // an exit: c < 8, the parent in another tile
      if (c >= 8)
        return;
      t = j.xp[(long long)(y + off_y(c)) * d.W + x + off_x(c)];
      if (t >= (uint32_t)d.W * (uint32_t)d.H)
        return; // kNoExit, or not an image index (phase A writes only those)
    }
  }

  __device__ __forceinline__ float sgnf(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

  template <int KIND>
  __device__ __forceinline__ float img_val(const ImgDesc& d, int x, int y)
  {
    return __uint_as_float(load_pix<IMG_F32>(d.img + (long long)y * d.img_pitch, x));
  }

  __global__ void k_grad(const BackJob* jobs, KParams p)
  {
    const BackJob&  j        = jobs[blockIdx.z];
    const ImgDesc&  d        = j.d;
    const double*   G        = j.G;
    unsigned char*  gi       = j.gi;
    const long long gi_pitch = j.gi_pitch;
    unsigned char*  gm       = j.gm;
    const long long gm_pitch = j.gm_pitch;
    const int       x        = blockIdx.x * blockDim.x + threadIdx.x;
    const int       y        = blockIdx.y;
    if (x >= d.W || y >= d.H)
      return;
    const long long id   = (long long)y * d.W + x;
    const float     ip   = img_val<IMG_F32>(d, x, y);
    int             code = d.pred[(long long)y * d.pred_pitch + x];
    if (code < 8 && (x + off_x(code) < 0 || y + off_y(code) < 0 || x + off_x(code) >= d.W || y + off_y(code) >= d.H))
      code = 8; // invalid code: treated as a root
    // In double, like the subtree sums: a gradient is a difference of sums
    // that can be far larger than it; only the output is rounded to float.
    const double lg  = (double)p.l_grad;
    double       acc = 0.0;
    if (code < 8)
    {
      const float iq = img_val<IMG_F32>(d, x + off_x(code), y + off_y(code));
      acc += G[id] * lg * (double)sgnf(ip - iq);
    }
    // The eight possible children: codes, G and pixels loaded first from
    // clamped addresses, then summed in k order for the real children.
    unsigned char pc[8];
    double        gc[8];
    float         ic[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = min(max(x - off_x(k), 0), d.W - 1), cy = min(max(y - off_y(k), 0), d.H - 1);
      pc[k]        = d.pred[(long long)cy * d.pred_pitch + cx];
      gc[k]        = G[(long long)cy * d.W + cx];
      ic[k]        = img_val<IMG_F32>(d, cx, cy);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = x - off_x(k), cy = y - off_y(k);
      if (cx >= 0 && cy >= 0 && cx < d.W && cy < d.H && pc[k] == k)
        acc -= gc[k] * lg * (double)sgnf(ic[k] - ip);
    }
    reinterpret_cast<float*>(gi + (long long)y * gi_pitch)[x] = (float)acc;
    if (gm)
      reinterpret_cast<float*>(gm + (long long)y * gm_pitch)[x] = (code >= 8) ? (float)(-(double)p.v * G[id]) : 0.f;
  }

  hipError_t launch_pred_batched(const ImgDesc* tab, int n, int maxW, int maxH, const KParams& p, bool c8,
                                 hipStream_t s)
  {
    if (n <= 0 || maxW <= 0 || maxH <= 0)
      return hipSuccess;
    const dim3 grid((maxW + 255) / 256, maxH, n), blk(256);
    switch (p.img_kind)
    {
    case IMG_U8:
      hipLaunchKernelGGL(k_pred_batched<IMG_U8>, grid, blk, 0, s, tab, p, (int)c8);
      break;
    case IMG_RGB8:
      hipLaunchKernelGGL(k_pred_batched<IMG_RGB8>, grid, blk, 0, s, tab, p, (int)c8);
      break;
    default:
      hipLaunchKernelGGL(k_pred_batched<IMG_F32>, grid, blk, 0, s, tab, p, (int)c8);
      break;
    }
    return hipGetLastError();
  }

  hipError_t launch_pred(const ImgDesc& d, const KParams& p, bool c8, hipStream_t s)
  {
    if (d.W <= 0 || d.H <= 0)
      return hipSuccess;
    const dim3 grid((d.W + 255) / 256, d.H), blk(256);
    switch (p.img_kind)
    {
    case IMG_U8:
      hipLaunchKernelGGL(k_pred<IMG_U8>, grid, blk, 0, s, d, p, (int)c8);
      break;
    case IMG_RGB8:
      hipLaunchKernelGGL(k_pred<IMG_RGB8>, grid, blk, 0, s, d, p, (int)c8);
      break;
    default:
      hipLaunchKernelGGL(k_pred<IMG_F32>, grid, blk, 0, s, d, p, (int)c8);
      break;
    }
    return hipGetLastError();
  }

  // jobs: n device entries; the grid covers the widest and tallest image
  // (maxW x maxH), and every block of a smaller image past its edge returns.
  hipError_t launch_backward(const BackJob* jobs, int n, int maxW, int maxH, const KParams& p, hipStream_t s)
  {
    if (n <= 0 || maxW <= 0 || maxH <= 0)
      return hipSuccess;
    const dim3 grid((maxW + 255) / 256, maxH, n), blk(256);
    // Each kernel reads what the one before wrote: stop at the first failed launch.
    unsigned gx = 0, gy = 0;
    if (!backward_tiles(n, maxW, maxH, kBT, &gx, &gy))
      return hipErrorInvalidConfiguration;
    const dim3 tiles(gx, gy, n);
    hipLaunchKernelGGL(k_tile_sums, tiles, dim3(kBThreads), 0, s, jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_coarse_count, tiles, blk, 0, s, jobs);
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_coarse_climb, tiles, blk, 0, s, jobs);
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_tile_inflow, tiles, dim3(kBThreads), 0, s, jobs);
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_grad, grid, blk, 0, s, jobs, p);
    return hipGetLastError();
  }

  // ---- self-test ---------------------------------------------------------
  __global__ void k_selftest_dpp(int* out)
  {
    const int   lane = threadIdx.x;
    const float v    = (float)(lane + 1);
    const float a    = shr_z(v);           // expect lane     (value of lane-1), 0 at lane 0
    const float b    = shl_z(v);           // expect lane + 2 (value of lane+1), 0 at lane 63
    const float c    = shr_keep(-1.f, v);  // lane 0 keeps -1
    const float rl   = rol_f(v);           // value of lane (lane+1) % 64
    const float rr   = ror_f(v);           // value of lane (lane+63) % 64
    const bool  ok   = a == (lane == 0 ? 0.f : (float)lane) && b == (lane == 63 ? 0.f : (float)(lane + 2)) &&
                    c == (lane == 0 ? -1.f : (float)lane) && rl == (float)((lane + 1) % 64 + 1) &&
                    rr == (float)((lane + 63) % 64 + 1);
    // Integer shifts feeding an integer subtract (the uint8 cost pattern).
    const uint32_t u  = (uint32_t)(lane * 37 % 256);
    const uint32_t ul = shl_u(u), ur = shr_u(u);
    const int      dl = (int)u - (int)ul, dr = (int)u - (int)ur;
    const int      el = (int)u - (lane == 63 ? 0 : (int)((lane + 1) * 37 % 256));
    const int      er = (int)u - (lane == 0 ? 0 : (int)((lane - 1) * 37 % 256));
    const bool     ok2 = dl == el && dr == er && pix_dist<IMG_U8>(u, ul) == (float)(el < 0 ? -el : el);
    // The RGB8 squared distance (dot products) against channel differences,
    // on pixel pairs that include every channel extreme.
    bool ok3 = true;
    for (int k = 0; k < 64; ++k)
    {
      const uint32_t pa = ((uint32_t)(lane * 97 + 13 * k) & 255u) | (((uint32_t)(lane * 5 + k * 71) & 255u) << 8) |
                          (((uint32_t)(255 - lane * 3 - k) & 255u) << 16);
      const uint32_t pb = ((uint32_t)(k * 4 + (lane & 1) * 255) & 255u) | ((uint32_t)((lane + k) & 1 ? 255 : 0) << 8) |
                          (((uint32_t)(lane * k) & 255u) << 16);
      const int dr2 = (int)(pa & 255u) - (int)(pb & 255u);
      const int dg2 = (int)((pa >> 8) & 255u) - (int)((pb >> 8) & 255u);
      const int db2 = (int)((pa >> 16) & 255u) - (int)((pb >> 16) & 255u);
      ok3 &= rgb_dist2(pa, pb) == (uint32_t)(dr2 * dr2 + dg2 * dg2 + db2 * db2);
    }
    const uint64_t all = __ballot(ok && ok2 && ok3);
    if (lane == 0)
      out[0] = all == ~0ull ? 1 : 0;
  }

  __global__ void k_selftest_sqrt(float* out, int n)
  {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
      out[i] = sqrt_int_rn((uint32_t)i); // the RGB8 edge-cost square root
  }

  hipError_t launch_selftest(int* dpp_out, float* sqrt_out, hipStream_t s)
  {
    hipLaunchKernelGGL(k_selftest_dpp, dim3(1), dim3(64), 0, s, dpp_out);
    const int n = 3 * 255 * 255 + 1;
    hipLaunchKernelGGL(k_selftest_sqrt, dim3((n + 255) / 256), dim3(256), 0, s, sqrt_out, n);
    return hipGetLastError();
  }
} // namespace dtk
