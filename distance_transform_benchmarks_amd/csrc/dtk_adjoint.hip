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
// Subtree sums without a sort: every pixel counts its children; one thread
// starts at each leaf and climbs while it is the last child to arrive at
// the parent, so each pixel is finished exactly once, after all of its
// children.  A pixel's count of children still to arrive and the float sum
// of their subtree sums share one 64-bit word, updated by a relaxed
// compare-and-swap: the last child reads the complete sum from the value it
// wrote itself, so no fence orders two addresses (an acquire-release
// counter beside a separate sum costs an L2 write-back and invalidate per
// step).  Most pixels have a single child, and a single child is always
// the last to arrive: k_child_count also writes an immutable byte per pixel
// (child count << 4 | predecessor code), and a climb through a pixel with
// one child reads only that byte (a plain load) and moves on.  The kernels
// take a table of images (blockIdx.z), so the chains of a whole batch climb
// in one launch.
#include <hip/hip_runtime.h>

#include "dtk_launch.hpp"
#include "dtk_tile.hpp"

namespace dtk
{
  __device__ constexpr int kOffX[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
  __device__ constexpr int kOffY[8] = {0, -1, -1, -1, 0, 1, 1, 1};

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
        const int nx = min(max(x + kOffX[k], 0), d.W - 1), ny = min(max(y + kOffY[k], 0), d.H - 1);
        Dq[k]        = rd_f(Db, d.D_pitch, nx, ny);
        iq[k]        = load_pix<KIND>(d.img + (long long)ny * d.img_pitch, nx);
      }
#pragma unroll
      for (int k = 7; k >= 0; --k)
      {
        const bool diag = kOffX[k] != 0 && kOffY[k] != 0;
        const int  nx = x + kOffX[k], ny = y + kOffY[k];
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

  // Children count of every pixel (gather, no atomics).
  __global__ void k_child_count(const BackJob* jobs)
  {
    const BackJob& j = jobs[blockIdx.z];
    const ImgDesc& d = j.d;
    const int      x = blockIdx.x * blockDim.x + threadIdx.x;
    const int      y = blockIdx.y;
    if (x >= d.W || y >= d.H)
      return;
    // Child c of p has c + off[k] == p and pred(c) == k; the eight codes
    // are loaded first, from clamped addresses.
    unsigned char pc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = min(max(x - kOffX[k], 0), d.W - 1), cy = min(max(y - kOffY[k], 0), d.H - 1);
      pc[k]        = d.pred[(long long)cy * d.pred_pitch + cx];
    }
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = x - kOffX[k], cy = y - kOffY[k];
      n += (cx >= 0 && cy >= 0 && cx < d.W && cy < d.H && pc[k] == k) ? 1u : 0u;
    }
    j.cw[(long long)y * d.W + x] = (unsigned long long)n << 32; // sum +0.0f
    int code = d.pred[(long long)y * d.pred_pitch + x];
    if (code < 8 && (x + kOffX[code] < 0 || y + kOffY[code] < 0 || x + kOffX[code] >= d.W || y + kOffY[code] >= d.H))
      code = 8; // invalid code: treated as a root
    j.meta[(long long)y * d.W + x] = (unsigned char)(n << 4 | (uint32_t)code);
  }

  // Leaf-started climbs computing G.
  // gx > 0: a one-dimensional grid in which image z runs on XCD z % 8
  // (workgroups go to the XCDs in turn), gx x gy blocks per image.  The
  // climbs of one image converge on shared trunks, so keeping an image on
  // one XCD keeps those lines in one L2 instead of fetching them into all
  // eight (diff32: 6.4 -> 5.1 ms per backward).
  __global__ void k_subtree(const BackJob* jobs, int gx, int gy, int n)
  {
    int bx = (int)blockIdx.x, by = (int)blockIdx.y, bz = (int)blockIdx.z;
    if (gx > 0)
    {
      const unsigned b = blockIdx.x, xcd = b & 7u, local = b >> 3;
      const unsigned per = (unsigned)gx * (unsigned)gy, slot = local / per, rem = local - slot * per;
      bz = (int)(xcd + 8u * slot);
      if (bz >= n)
        return;
      bx = (int)(rem % (unsigned)gx);
      by = (int)(rem / (unsigned)gx);
    }
    const BackJob&       j        = jobs[bz];
    const ImgDesc&       d        = j.d;
    const unsigned char* gd       = j.gd;
    const long long      gd_pitch = j.gd_pitch;
    unsigned long long*  cw       = j.cw;
    float*               G        = j.G;
    int                  x        = bx * blockDim.x + threadIdx.x;
    int                  y        = by;
    if (x >= d.W || y >= d.H)
      return;
    long long            id   = (long long)y * d.W + x;
    const unsigned char* meta = j.meta; // immutable here: the words in cw change
    uint32_t             m    = meta[id];
    if ((m >> 4) != 0u)
      return; // not a leaf
    float g = rd_f(gd, gd_pitch, x, y); // a leaf: G = W
    for (;;)
    {
      G[id]          = g;
      const int code = (int)(m & 15u);
      if (code >= 8)
        return; // root (k_child_count marked invalid codes as roots)
      x += kOffX[code];
      y += kOffY[code];
      id = (long long)y * d.W + x;
      m  = meta[id];
      if ((m >> 4) == 1u)
      {
        // The only child: the last to arrive, and the sum is g alone.
        g = rd_f(gd, gd_pitch, x, y) + g;
        continue;
      }
      // One fewer child to wait for, g added to the sum.  Every failed
      // exchange means a sibling's succeeded, and a pixel has at most eight
      // children, so the loop ends.
      unsigned long long old = __hip_atomic_load(&cw[id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned long long nw;
      for (;;)
      {
        const uint32_t left = (uint32_t)(old >> 32) - 1u;
        const float    sum  = __uint_as_float((uint32_t)old) + g;
        nw                  = ((unsigned long long)left << 32) | __float_as_uint(sum);
        if (__hip_atomic_compare_exchange_strong(&cw[id], &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
          break;
      }
      if ((uint32_t)(nw >> 32) != 0u)
        return; // a sibling is still climbing: the last one continues
      g = rd_f(gd, gd_pitch, x, y) + __uint_as_float((uint32_t)nw);
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
    const float*    G        = j.G;
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
    if (code < 8 && (x + kOffX[code] < 0 || y + kOffY[code] < 0 || x + kOffX[code] >= d.W || y + kOffY[code] >= d.H))
      code = 8; // invalid code: treated as a root
    float acc = 0.f;
    if (code < 8)
    {
      const float iq = img_val<IMG_F32>(d, x + kOffX[code], y + kOffY[code]);
      acc += G[id] * p.l_grad * sgnf(ip - iq);
    }
    // The eight possible children: codes, G and pixels loaded first from
    // clamped addresses, then summed in k order for the real children.
    unsigned char pc[8];
    float         gc[8], ic[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = min(max(x - kOffX[k], 0), d.W - 1), cy = min(max(y - kOffY[k], 0), d.H - 1);
      pc[k]        = d.pred[(long long)cy * d.pred_pitch + cx];
      gc[k]        = G[(long long)cy * d.W + cx];
      ic[k]        = img_val<IMG_F32>(d, cx, cy);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
      const int cx = x - kOffX[k], cy = y - kOffY[k];
      if (cx >= 0 && cy >= 0 && cx < d.W && cy < d.H && pc[k] == k)
        acc -= gc[k] * p.l_grad * sgnf(ic[k] - ip);
    }
    reinterpret_cast<float*>(gi + (long long)y * gi_pitch)[x] = acc;
    if (gm)
      reinterpret_cast<float*>(gm + (long long)y * gm_pitch)[x] = (code >= 8) ? -p.v * G[id] : 0.f;
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
    hipLaunchKernelGGL(k_child_count, grid, blk, 0, s, jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    // Each kernel reads what the one before wrote: stop at the first failed launch.
    if (const unsigned long long blocks = subtree_grid_1d(n, maxW, maxH))
      hipLaunchKernelGGL(k_subtree, dim3((unsigned)blocks), blk, 0, s, jobs, (int)grid.x, (int)grid.y, n);
    else
      hipLaunchKernelGGL(k_subtree, grid, blk, 0, s, jobs, 0, 0, n);
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
    const uint64_t all = __ballot(ok && ok2);
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
