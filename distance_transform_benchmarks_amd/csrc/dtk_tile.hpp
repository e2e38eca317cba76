// Wave64 tile solver for the GGDT (device code, gfx950).
//
// Reference counterpart: the block solve of dt/src/generalised_distance_transform_blocks.cu:8-112
// and dt/src/generalised_distance_transform_task.cu:13-127, with the per-tile
// passes of dt/src/dt_block_passes.cuh:11-127.  The relaxation arithmetic is
// the raster's (dt/src/generalised_distance_transform.cu:34-56).
//
// Region and layouts.  A tile owns 62 x (RH-2) pixels; one wave loads the
// 64 x RH region around it (owned pixels plus a one-pixel halo).
//
//   column layout : lane i = region column i, register r = region row r.
//                   Column sweeps step through the registers; the three
//                   upper (or lower) neighbours come from lanes i-1, i, i+1
//                   through DPP wave shifts.
//   row layout    : the 64 columns are cut into S segments of SW = 64/S
//                   columns; lane (h, r) = h*RH + r holds row r of segment h
//                   in registers k = 1..SW (column SW*h + k - 1).  Registers
//                   0 and SW+1 hold the neighbouring segments' edge columns
//                   as read at the start of the row phase (or +inf at the
//                   region border); they are read-only during the phase.
//   Layout changes go through one LDS buffer of RH x 65 floats.
//
// Edge costs are precomputed once per visit, per undirected edge, as floats
// (+inf for an edge that leaves the image or the region, which also makes
// the DPP values that wrap between segments or off the wave harmless):
//   column layout: N, NW, NE of every pixel (edges to the row above);
//   row layout   : W, NW, SW of every pixel (edges to the column on the left).
// A relaxation is pulled when the receiver owns the edge cost and pushed
// (the sender adds its own cost, then the candidate is shifted) otherwise.
//
// Halo pixels are relaxed like owned pixels: every value the solver ever
// holds is the float sum of a real path, so a smaller halo value is still a
// valid upper bound.  Only owned pixels are written back.
//
// Convergence inside a visit: stop after four consecutive sweeps that changed
// nothing, where a backward row sweep counts as quiet only if the forward row
// sweep of the same phase was quiet too (otherwise its segment-edge copies
// may be stale).  Such a window certifies every edge of the region.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "dtk_internal.hpp"

namespace dtk
{
#define DTK_INF __builtin_huge_valf()

  // ---- DPP wave shifts (gfx9 wave_shr:1 / wave_shl:1) --------------------
  // shr: lane i receives lane i-1; shl: lane i receives lane i+1.
  __device__ __forceinline__ float shr_z(float v)
  {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
  }
  __device__ __forceinline__ float shl_z(float v)
  {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
  }
  // Lanes without a source keep `old`.
  __device__ __forceinline__ float shr_keep(float old, float v)
  {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xF, 0xF, false));
  }
  __device__ __forceinline__ float shl_keep(float old, float v)
  {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xF, 0xF, false));
  }
  // Wave rotates by one lane: rol gives lane i the value of lane i+1 (lane
  // 63 that of lane 0), ror that of lane i-1 (lane 0 that of lane 63).  The
  // backward sweeps use them where the sending edge lane always holds +inf,
  // so the wrapped value is a non-candidate, as a kept old value would be,
  // without the copy that a keep-old shift needs.
  __device__ __forceinline__ float rol_f(float v)
  {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));
  }
  __device__ __forceinline__ float ror_f(float v)
  {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x13C, 0xF, 0xF, false));
  }
  // Integer wave shifts.  The result is copied through an opaque VALU move so
  // the DPP is never folded into an integer VOP2 instruction: measured on
  // gfx950, a folded v_subrev_u32_dpp with wave_shl:1 produced wrong values,
  // while a plain v_mov_b32_dpp followed by the subtract is correct
  // (dtk_selftest checks both the shifts and this subtract pattern).  The asm
  // is not volatile: it has no side effect, so the compiler may share one
  // copy between identical shifts (a row's W edge and the next row's NW edge
  // shift the same pixels), which saves a tenth of the tile kernels' VALU
  // instructions.
  __device__ __forceinline__ uint32_t opaque_u(uint32_t v)
  {
    uint32_t r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
  }
  __device__ __forceinline__ uint32_t shr_u(uint32_t v)
  {
    return opaque_u((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true));
  }
  __device__ __forceinline__ uint32_t shl_u(uint32_t v)
  {
    return opaque_u((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true));
  }

  // Minimum of non-negative floats through their bit patterns.  Every value
  // the solver handles lies in [+0, +inf] (D0 >= 0, costs >= 0), where the
  // unsigned order of the bits equals the float order, and a NaN (bits above
  // +inf) never wins -- the same outcome as the strict compares of the
  // reference (a NaN candidate never replaces D).  Unlike fminf this needs no
  // NaN canonicalisation.
  __device__ __forceinline__ float minf_nn(float a, float b)
  {
    return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
  }
  __device__ __forceinline__ float min3f(float a, float b, float c)
  {
    return __uint_as_float(min(min(__float_as_uint(a), __float_as_uint(b)), __float_as_uint(c)));
  }
  __device__ __forceinline__ bool lt_nn(float a, float b) { return __float_as_uint(a) < __float_as_uint(b); }

  // Loads through the global address space (generic pointers would compile
  // to FLAT loads, which also wait on the LDS counter).
  typedef const __attribute__((address_space(1))) unsigned char* gbytes;
  template <typename T>
  __device__ __forceinline__ T gload(const void* base, long long off)
  {
    return *reinterpret_cast<const __attribute__((address_space(1))) T*>((gbytes)base + off);
  }

  __device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

  // Index of this wave in its workgroup, as a wave-uniform (scalar) value:
  // row addresses and bounds tests derived from it then use scalar
  // instructions instead of being computed per lane.
  __device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

  // ---- pixels and edge costs -----------------------------------------------
  template <int KIND>
  __device__ __forceinline__ uint32_t load_pix(const unsigned char* row, int x)
  {
    if constexpr (KIND == IMG_U8)
      return gload<unsigned char>(row, x);
    else if constexpr (KIND == IMG_RGB8)
    {
      const long long o = 3ll * x;
      return (uint32_t)gload<unsigned char>(row, o) | ((uint32_t)gload<unsigned char>(row, o + 1) << 8) |
             ((uint32_t)gload<unsigned char>(row, o + 2) << 16);
    }
    else
      return gload<uint32_t>(row, 4ll * x);
  }

  // Correctly rounded square root of an integer 0 <= n <= 3 * 255^2 (the
  // exact squared colour distance), equal to sqrtf((float)n).  v_sqrt_f32
  // is accurate to 1 ulp; on gfx950 it never returns more than the
  // correctly rounded root on this domain (checked exhaustively for every
  // n < 2^18), so the root is s or its upper float neighbour, and the exact
  // fma residual n - s_up * s picks it (the upper half of the standard
  // two-sided correction of a 1-ulp square root).  dtk_selftest compares
  // every n of the domain with the host's sqrtf on each run.
  __device__ __forceinline__ float sqrt_int_rn(uint32_t n)
  {
    const float x    = (float)n;
    const float s    = __builtin_amdgcn_sqrtf(x);
    const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
    return __builtin_fmaf(-s_up, s, x) > 0.f ? s_up : s;
  }

  // Exact squared distance of two RGB8 pixels packed r | g << 8 | b << 16
  // (top byte zero): |a|^2 + |b|^2 - 2 a.b with three v_dot4_u32_u8, in
  // place of extracting and squaring three channel differences.  No term
  // exceeds 2 * 3 * 255^2, so the unsigned arithmetic is exact.
  __device__ __forceinline__ uint32_t rgb_dist2(uint32_t a, uint32_t b)
  {
    const uint32_t ab = __builtin_amdgcn_udot4(a, b, 0u, false);
    const uint32_t aa = __builtin_amdgcn_udot4(a, a, 0u, false);
    const uint32_t bb = __builtin_amdgcn_udot4(b, b, 0u, false);
    return aa + bb - 2u * ab;
  }

  // |I(p) - I(q)|: uint8 abs difference (dt/src/utils.cuh:37-41); RGB8:
  // correctly rounded sqrtf of the exact integer squared distance; float32:
  // fabsf of the rounded difference.
  template <int KIND>
  __device__ __forceinline__ float pix_dist(uint32_t a, uint32_t b)
  {
    if constexpr (KIND == IMG_U8)
    {
      const int d = (int)a - (int)b;
      return (float)(d < 0 ? -d : d);
    }
    else if constexpr (KIND == IMG_RGB8)
      return sqrt_int_rn(rgb_dist2(a, b));
    else
      return fabsf(__uint_as_float(a) - __uint_as_float(b));
  }

  // Cost term of an edge at lambda > 0: g = l_grad * dist, +inf when not
  // valid.  (Lambda 0 takes the chamfer costs of setup_costs_const.)  A NaN
  // g (float images only; lambda is finite, so integer kinds give finite g)
  // is stored as +inf: a NaN candidate never wins a strict compare, and
  // neither does +inf, so both leave D unchanged.
  template <int KIND>
  __device__ __forceinline__ float edge_cost(bool valid, uint32_t a, uint32_t b, const KParams& p)
  {
    float g = p.l_grad * pix_dist<KIND>(a, b);
    if constexpr (KIND == IMG_F32)
      g = (g == g) ? g : DTK_INF;
    return valid ? g : DTK_INF;
  }

  struct VisitResult
  {
    uint32_t border;   // bit k: this visit improved a pixel of neighbour k (see nbr_dx/nbr_dy)
    int      sweeps;   // sweeps run
    bool     converged;
    bool     changed;  // any owned pixel changed
    uint32_t cyc_setup; // shader-clock cycles: loads + edge costs
    uint32_t cyc_sweep; //   sweeps and layout changes
    uint32_t cyc_store; //   border check and write-back
  };

  // Neighbour tiles in border-bit order: L, R, T, B, TL, TR, BL, BR.
  // Offsets of a direction k (0..7) from a table packed two bits per entry
  // (offset + 1): looking one up by a per-lane k is a shift, where an
  // indexed __device__ array would be a load from memory.
  __host__ __device__ constexpr int unpack_off(uint32_t packed, int k) { return (int)((packed >> (2 * k)) & 3u) - 1; }
  constexpr uint32_t kNbrPackX = 0x8858u, kNbrPackY = 0xA085u;
  __host__ __device__ constexpr int nbr_dx(int k) { return unpack_off(kNbrPackX, k); }
  __host__ __device__ constexpr int nbr_dy(int k) { return unpack_off(kNbrPackY, k); }
  static_assert(nbr_dx(0) == -1 && nbr_dx(1) == 1 && nbr_dx(2) == 0 && nbr_dx(3) == 0 && nbr_dx(4) == -1 &&
                  nbr_dx(5) == 1 && nbr_dx(6) == -1 && nbr_dx(7) == 1,
                "neighbour x offsets: L, R, T, B, TL, TR, BL, BR");
  static_assert(nbr_dy(0) == 0 && nbr_dy(1) == 0 && nbr_dy(2) == -1 && nbr_dy(3) == 1 && nbr_dy(4) == -1 &&
                  nbr_dy(5) == -1 && nbr_dy(6) == 1 && nbr_dy(7) == 1,
                "neighbour y offsets: L, R, T, B, TL, TR, BL, BR");

  template <int RH>
  __device__ __forceinline__ const float* drow(const ImgDesc& d, int y)
  {
    return reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(d.D) + (long long)y * d.D_pitch);
  }

  // The transpose buffer T belongs to one wave, and the LDS operations of a
  // wave execute in order, so a transpose needs no workgroup barrier: this
  // only keeps the compiler from moving LDS accesses across it.
  __device__ __forceinline__ void lds_wave_sync()
  {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }

  // Column layout -> LDS -> row layout of one array.  k = 0..SW+1.
  template <int S>
  __device__ __forceinline__ void col_to_row(const float (&c)[64 / S], float (&rw)[64 / S + 2], float* T, int lane)
  {
    constexpr int RH = 64 / S, SW = 64 / S;
#pragma unroll
    for (int r = 0; r < RH; ++r)
      T[r * 65 + lane] = c[r];
    lds_wave_sync();
    const int h = lane / RH, rr = lane % RH;
#pragma unroll
    for (int k = 0; k < SW + 2; ++k)
    {
      const int  col = SW * h + k - 1;
      const bool ok  = (k >= 1 && k <= SW) || (k == 0 ? h > 0 : h < S - 1);
      const float v  = T[rr * 65 + (ok ? col : 0)];
      rw[k]          = ok ? v : DTK_INF;
    }
    lds_wave_sync();
  }

  template <int S>
  __device__ __forceinline__ void row_to_col(const float (&rw)[64 / S + 2], float (&c)[64 / S], float* T, int lane)
  {
    constexpr int RH = 64 / S, SW = 64 / S;
    const int     h = lane / RH, rr = lane % RH;
#pragma unroll
    for (int k = 1; k <= SW; ++k)
      T[rr * 65 + SW * h + k - 1] = rw[k];
    lds_wave_sync();
#pragma unroll
    for (int r = 0; r < RH; ++r)
      c[r] = T[r * 65 + lane];
    lds_wave_sync();
  }

  // ---- edge costs seen by the sweeps ------------------------------------------
  // Per-pixel costs of a region (column layout N / NW / NE, row layout
  // W / NW / SW), +inf on edges that leave the image or the region.
  template <int S>
  struct TileCosts
  {
    static constexpr int RH = 64 / S, SW = 64 / S;
    float                cN_[RH], cNW_[RH], cNE_[RH];
    float                rW_[SW + 2], rNW_[SW + 2], rSW_[SW + 2];
    __device__ __forceinline__ float cN(int r) const { return cN_[r]; }
    __device__ __forceinline__ float cNW(int r) const { return cNW_[r]; }
    __device__ __forceinline__ float cNE(int r) const { return cNE_[r]; }
    __device__ __forceinline__ float rW(int k) const { return rW_[k]; }
    __device__ __forceinline__ float rNW(int k) const { return rNW_[k]; }
    __device__ __forceinline__ float rSW(int k) const { return rSW_[k]; }
  };

  // lambda == 0: every edge costs its Euclidean step, so a wave keeps two
  // scalars and four per-lane values instead of six arrays (and runs more
  // waves per SIMD).  Only edges that would take a value across the region's
  // edge or between row segments are +inf: the wave-shift sources that do
  // not exist read +inf (column sweeps) or are cut by rowNW / rowSW (row
  // sweeps), and the region's outer slots hold +inf.  Pixels outside the
  // image are not masked: with chamfer costs a path through them is longer
  // by at least 2 - sqrt(2) than its projection on the image border, far
  // beyond float32 rounding, so it never wins and never reaches the stored
  // pixels (and the halo checks that wake neighbours skip them).
  struct ConstCosts
  {
    float ea, ed, colNW, colNE, rowNW, rowSW;
    __device__ __forceinline__ float cN(int) const { return ea; }
    __device__ __forceinline__ float cNW(int) const { return colNW; }
    __device__ __forceinline__ float cNE(int) const { return colNE; }
    __device__ __forceinline__ float rW(int) const { return ea; }
    __device__ __forceinline__ float rNW(int) const { return rowNW; }
    __device__ __forceinline__ float rSW(int) const { return rowSW; }
  };

  // ---- sweeps --------------------------------------------------------------
  // Column forward: row r pulls from row r-1 (dt/src/generalised_distance_transform.cu:59-99, Forward).
  // edge: changes in the rows a stacked neighbour wave also holds (1, RH-2,
  // RH-1; row 0 is only read).
  template <int RH, bool C8, bool TWO, class CC>
  __device__ __forceinline__ uint64_t sweep_col_fwd(float (&D)[RH], const CC& cc, float ea, float ed, uint64_t& edge)
  {
    uint64_t ch = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 1; r < RH; ++r)
    {
      const float up = D[r - 1];
      const float c  = (TWO ? up + ea : up) + cc.cN(r);
      float       m  = c;
      if constexpr (C8)
      {
        const float t = TWO ? up + ed : up;
        const float a = shr_z(t) + cc.cNW(r);
        const float b = shl_z(t) + cc.cNE(r);
        m             = min3f(a, b, c);
      }
      const uint64_t bal = ballot(lt_nn(m, D[r]));
      ch |= bal;
      if (r == 1 || r >= RH - 2)
        edge |= bal;
      D[r] = minf_nn(m, D[r]);
    }
    return ch;
  }

  // Column backward: row r receives from row r+1.  The diagonal candidates
  // are formed by the sender (its NW / NE cost) and shifted.
  template <int RH, bool C8, bool TWO, class CC>
  __device__ __forceinline__ uint64_t sweep_col_bwd(float (&D)[RH], const CC& cc, float ea, float ed)
  {
    uint64_t ch = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = RH - 2; r >= 0; --r)
    {
      const float dn = D[r + 1];
      const float c  = (TWO ? dn + ea : dn) + cc.cN(r + 1);
      float       m  = c;
      if constexpr (C8)
      {
        const float t   = TWO ? dn + ed : dn;
        const float pNW = t + cc.cNW(r + 1); // (i, r+1) -> (i-1, r); +inf at lane 0
        const float pNE = t + cc.cNE(r + 1); // (i, r+1) -> (i+1, r); +inf at lane 63
        const float a   = rol_f(pNW);
        const float b   = ror_f(pNE);
        m               = min3f(a, b, c);
      }
      ch |= ballot(lt_nn(m, D[r]));
      D[r] = minf_nn(m, D[r]);
    }
    return ch;
  }

  // Row forward: slot k pulls from slot k-1 (dt/src/generalised_distance_transform.cu:17-57, Forward).
  // edge: changes in a segment's first and last columns (slots 1 and SW),
  // which the neighbouring segments hold as copies taken at the transpose.
  template <int SW, bool C8, bool TWO, class CC>
  __device__ __forceinline__ uint64_t sweep_row_fwd(float (&R)[SW + 2], const CC& cc, float ea, float ed,
                                                    uint64_t& edge)
  {
    uint64_t ch = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 1; k <= SW; ++k)
    {
      const float lf = R[k - 1];
      const float c  = (TWO ? lf + ea : lf) + cc.rW(k);
      float       m  = c;
      if constexpr (C8)
      {
        const float t = TWO ? lf + ed : lf;
        const float a = shr_z(t) + cc.rNW(k);
        const float b = shl_z(t) + cc.rSW(k);
        m             = min3f(a, b, c);
      }
      const uint64_t bal = ballot(lt_nn(m, R[k]));
      ch |= bal;
      if (k == 1 || k == SW)
        edge |= bal;
      R[k] = minf_nn(m, R[k]);
    }
    return ch;
  }

  // Lanes of the row layout (lane = segment * RH + row) that hold a row a
  // stacked neighbour wave also holds: rows 0, 1, RH-2, RH-1 of every segment.
  template <int S>
  __device__ constexpr uint64_t shared_row_lanes()
  {
    constexpr int RH = 64 / S;
    uint64_t      m  = 0;
    for (int h = 0; h < S; ++h)
      m |= (1ull << (h * RH)) | (1ull << (h * RH + 1)) | (1ull << (h * RH + RH - 2)) | (1ull << (h * RH + RH - 1));
    return m;
  }

  // Row backward: slot k receives from slot k+1.
  template <int SW, bool C8, bool TWO, class CC>
  __device__ __forceinline__ uint64_t sweep_row_bwd(float (&R)[SW + 2], const CC& cc, float ea, float ed)
  {
    uint64_t ch = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = SW; k >= 1; --k)
    {
      const float rt = R[k + 1];
      const float c  = (TWO ? rt + ea : rt) + cc.rW(k + 1);
      float       m  = c;
      if constexpr (C8)
      {
        const float t   = TWO ? rt + ed : rt;
        const float pSW = t + cc.rSW(k + 1); // (c+1, r) -> (c, r+1); +inf in a segment's last row (lane 63)
        const float pNW = t + cc.rNW(k + 1); // (c+1, r) -> (c, r-1); +inf in a segment's row 0 (lane 0)
        const float a   = ror_f(pSW);
        const float b   = rol_f(pNW);
        m               = min3f(a, b, c);
      }
      ch |= ballot(lt_nn(m, R[k]));
      R[k] = minf_nn(m, R[k]);
    }
    return ch;
  }

  // ---- certification check ---------------------------------------------------
  // A clean pair of column sweeps leaves every edge it relaxes satisfied (N, S
  // and the four diagonals); only the straight W and E edges remain.  This
  // check tests them with the row sweeps' own candidate arithmetic, all pixels
  // at once, in the row layout: slot k against slot k-1 (a forward row
  // sweep's straight candidate) and slot k+1 (a backward one's).  Slots 0 and
  // SW+1 are the neighbouring segments' edge columns as read at the
  // transpose, so they are current.  When no pixel would be lowered the
  // region is at its fixed point, the state in which the next pair of row
  // sweeps would change nothing.  Returns the same value in every lane.  (The
  // mirror check of N and S after a clean row pair costs more than the column
  // pair it saves, measured on RGB8 4096^2.)
  template <int SW, bool TWO, class CC>
  __device__ __forceinline__ bool row_we_violated(const float (&R)[SW + 2], const CC& cc, float ea)
  {
    bool bad = false;
#pragma unroll
    for (int k = 1; k <= SW; ++k)
    {
      const float lf = R[k - 1], rt = R[k + 1];
      bad |= lt_nn((TWO ? lf + ea : lf) + cc.rW(k), R[k]) | lt_nn((TWO ? rt + ea : rt) + cc.rW(k + 1), R[k]);
    }
    return ballot(bad) != 0;
  }

  // ---- cost precomputation --------------------------------------------------
  // Image values of the region in the column layout (0 outside the image).
  template <int KIND, int RH>
  __device__ __forceinline__ void load_region_image(const ImgDesc& d, int X0, int Y0, int lane, uint32_t (&I)[RH])
  {
    // Every load is issued unconditionally from a clamped (in-bounds)
    // address and masked afterwards, so all rows are in flight together.
    const int  gx  = X0 + lane;
    const bool cin = gx >= 0 && gx < d.W;
    const int  cx  = min(max(gx, 0), d.W - 1);
#pragma unroll
    for (int r = 0; r < RH; ++r)
    {
      const int      gy = Y0 + r;
      const int      cy = min(max(gy, 0), d.H - 1);
      const uint32_t v  = load_pix<KIND>(d.img + (long long)cy * d.img_pitch, cx);
      I[r]              = (cin && gy >= 0 && gy < d.H) ? v : 0u;
    }
  }

  // Which neighbour an edge array refers to (relative to pixel (i, r)).
  enum EdgeDir : int
  {
    E_W  = 0, // (i-1, r)
    E_N  = 1, // (i, r-1)
    E_NW = 2, // (i-1, r-1)
    E_NE = 3, // (i+1, r-1)
    E_SW = 4, // (i-1, r+1)
  };

  // Cost of the edge from every pixel of the region to its DIR neighbour
  // (column layout), +inf when either end is outside the image or region.
  template <int KIND, int RH, int DIR>
  __device__ __forceinline__ void edge_costs(const ImgDesc& d, const KParams& p, int X0, int Y0, int lane,
                                             const uint32_t (&I)[RH], float (&out)[RH])
  {
    const int  gx  = X0 + lane;
    const bool cin = gx >= 0 && gx < d.W;
    const bool inL = lane >= 1 && gx - 1 >= 0 && gx - 1 < d.W;
    const bool inR = lane <= 62 && gx + 1 >= 0 && gx + 1 < d.W;
#pragma unroll
    for (int r = 0; r < RH; ++r)
    {
      const int  gy   = Y0 + r;
      const bool here = cin && gy >= 0 && gy < d.H;
      if constexpr (DIR == E_W)
        out[r] = edge_cost<KIND>(here && inL, I[r], shr_u(I[r]), p);
      else if constexpr (DIR == E_SW)
      {
        if (r <= RH - 2)
        {
          const bool rinD = gy + 1 >= 0 && gy + 1 < d.H;
          out[r]          = edge_cost<KIND>(here && inL && rinD, I[r], shr_u(I[r + 1]), p);
        }
        else
          out[r] = DTK_INF;
      }
      else
      {
        if (r >= 1)
        {
          const bool     rinU = gy - 1 >= 0 && gy - 1 < d.H;
          const uint32_t u    = I[r - 1];
          if constexpr (DIR == E_N)
            out[r] = edge_cost<KIND>(here && rinU, I[r], u, p);
          else if constexpr (DIR == E_NW)
            out[r] = edge_cost<KIND>(here && rinU && inL, I[r], shr_u(u), p);
          else
            out[r] = edge_cost<KIND>(here && rinU && inR, I[r], shl_u(u), p);
        }
        else
          out[r] = DTK_INF;
      }
    }
  }

  // All cost arrays of a visit: column layout N / NW / NE, row layout
  // W / NW / SW.  Computed one array at a time to bound register pressure.
  template <int KIND, int S, bool C8>
  __device__ __forceinline__ void setup_costs(const ImgDesc& d, const KParams& p, int X0, int Y0, int lane, float* T,
                                              float (&cN)[64 / S], float (&cNW)[64 / S], float (&cNE)[64 / S],
                                              float (&rW)[64 / S + 2], float (&rNW)[64 / S + 2],
                                              float (&rSW)[64 / S + 2])
  {
    constexpr int RH = 64 / S, SW = 64 / S;
    uint32_t      I[RH];
    load_region_image<KIND, RH>(d, X0, Y0, lane, I);
    {
      float tmp[RH];
      edge_costs<KIND, RH, E_W>(d, p, X0, Y0, lane, I, tmp);
      col_to_row<S>(tmp, rW, T, lane);
    }
    if constexpr (C8)
    {
      edge_costs<KIND, RH, E_NE>(d, p, X0, Y0, lane, I, cNE);
      {
        // The SW edge of (i, r) is the NE edge of (i - 1, r + 1): one lane
        // shift instead of a second cost evaluation.  Lane 0 and the last
        // row have no SW neighbour in the region.
        float tmp[RH];
#pragma unroll
        for (int r = 0; r < RH - 1; ++r)
        {
          const float sw = shr_z(cNE[r + 1]);
          tmp[r]         = lane >= 1 ? sw : DTK_INF;
        }
        tmp[RH - 1] = DTK_INF;
        col_to_row<S>(tmp, rSW, T, lane);
      }
      edge_costs<KIND, RH, E_NW>(d, p, X0, Y0, lane, I, cNW);
      col_to_row<S>(cNW, rNW, T, lane);
    }
    else
    {
#pragma unroll
      for (int r = 0; r < RH; ++r)
        cNW[r] = cNE[r] = DTK_INF;
#pragma unroll
      for (int k = 0; k < SW + 2; ++k)
        rNW[k] = rSW[k] = DTK_INF;
    }
    edge_costs<KIND, RH, E_N>(d, p, X0, Y0, lane, I, cN);
  }

  // ---- cached edge costs ------------------------------------------------------
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(1))) unsigned char* gbytes_w;

  // Writes the cost records of rows r0 .. r0+n-1 of a tile (n <= NR,
  // wave-uniform) for column gx = X0 + lane, lanes 1..62 (lanes 0 and 63
  // only supply neighbour values).  The NR+1 image rows (with the one above)
  // are loaded before any is used.  Record = {W, N, NW, NE} (C8) or {W, N}
  // (C4), each the cost of the edge to that neighbour as edge_cost()
  // defines it, +inf when the neighbour is outside the image.  Edge costs
  // are symmetric (|a - b| and the squared colour distance are), so the
  // solver reads the cost of (p, q) from either end.
  template <int KIND, bool C8, int NR>
  __device__ __forceinline__ void write_cost_group(const ImgDesc& d, const KParams& p, int X0, int Y0, int r0, int n,
                                                   int lane)
  {
    const int  gx  = X0 + lane;
    const int  cx  = min(max(gx, 0), d.W - 1);
    const bool own = lane >= 1 && lane <= kTileW && gx < d.W;
    const bool inL = gx - 1 >= 0;
    const bool inR = gx + 1 < d.W;
    uint32_t   I[NR + 1];
#pragma unroll
    for (int k = 0; k <= NR; ++k)
    {
      const int cy = min(max(Y0 + r0 - 1 + k, 0), d.H - 1);
      I[k]         = load_pix<KIND>(d.img + (long long)cy * d.img_pitch, cx);
    }
#pragma unroll
    for (int k = 1; k <= NR; ++k)
    {
      if (k > n)
        break;
      const int      gy   = Y0 + r0 - 1 + k;
      const uint32_t cur  = I[k], up = I[k - 1];
      const bool     rinU = gy - 1 >= 0;
      const float    cw   = edge_cost<KIND>(inL, cur, shr_u(cur), p);
      const float    cn   = edge_cost<KIND>(rinU, cur, up, p);
      if constexpr (C8)
      {
        const float cnw = edge_cost<KIND>(rinU && inL, cur, shr_u(up), p);
        const float cne = edge_cost<KIND>(rinU && inR, cur, shl_u(up), p);
        if (own && gy < d.H)
        {
          f4 v = {cw, cn, cnw, cne};
          *reinterpret_cast<__attribute__((address_space(1))) f4*>((gbytes_w)d.cost + (long long)gy * d.cost_pitch +
                                                                   16ll * gx) = v;
        }
      }
      else if (own && gy < d.H)
      {
        f2 v = {cw, cn};
        *reinterpret_cast<__attribute__((address_space(1))) f2*>((gbytes_w)d.cost + (long long)gy * d.cost_pitch +
                                                                 8ll * gx) = v;
      }
    }
  }

  // lambda == 0: every edge costs its Euclidean step (e_axis / e_diag), +inf
  // where an end lies outside the image or the region; no image access.
  template <int S, bool C8>
  __device__ __forceinline__ void setup_costs_const(const ImgDesc& d, const KParams& p, int X0, int Y0, int lane,
                                                    float* T, float (&cN)[64 / S], float (&cNW)[64 / S],
                                                    float (&cNE)[64 / S], float (&rW)[64 / S + 2],
                                                    float (&rNW)[64 / S + 2], float (&rSW)[64 / S + 2])
  {
    constexpr int RH  = 64 / S, SW = 64 / S;
    const int     gx  = X0 + lane;
    const bool    cin = gx >= 0 && gx < d.W;
    const bool    inL = lane >= 1 && gx - 1 >= 0 && gx - 1 < d.W;
    const bool    inR = lane <= 62 && gx + 1 >= 0 && gx + 1 < d.W;
    const float   ea = p.e_axis, ed = p.e_diag;
    {
      float tmp[RH];
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const int  gy   = Y0 + r;
        const bool here = cin && gy >= 0 && gy < d.H;
        tmp[r]          = (here && inL) ? ea : DTK_INF;
      }
      col_to_row<S>(tmp, rW, T, lane);
    }
#pragma unroll
    for (int r = 0; r < RH; ++r)
    {
      const int  gy   = Y0 + r;
      const bool here = cin && gy >= 0 && gy < d.H;
      const bool rinU = r >= 1 && gy - 1 >= 0 && gy - 1 < d.H;
      cN[r]           = (here && rinU) ? ea : DTK_INF;
      cNW[r]          = (C8 && here && rinU && inL) ? ed : DTK_INF;
      cNE[r]          = (C8 && here && rinU && inR) ? ed : DTK_INF;
    }
    if constexpr (C8)
    {
      col_to_row<S>(cNW, rNW, T, lane);
      float tmp[RH];
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const int  gy   = Y0 + r;
        const bool here = cin && gy >= 0 && gy < d.H;
        const bool rinD = r <= RH - 2 && gy + 1 >= 0 && gy + 1 < d.H;
        tmp[r]          = (here && inL && rinD) ? ed : DTK_INF;
      }
      col_to_row<S>(tmp, rSW, T, lane);
    }
    else
    {
#pragma unroll
      for (int k = 0; k < SW + 2; ++k)
        rNW[k] = rSW[k] = DTK_INF;
    }
  }

  // lambda > 0 with cached costs: one 16-byte (C8) or 8-byte (C4) load per
  // row and lane, then region masking.  The SW edge of (i, r) is the NE edge
  // of (i - 1, r + 1), moved one lane right.
  template <int S, bool C8>
  __device__ __forceinline__ void setup_costs_cached(const ImgDesc& d, int X0, int Y0, int lane, float* T,
                                                     float (&cN)[64 / S], float (&cNW)[64 / S], float (&cNE)[64 / S],
                                                     float (&rW)[64 / S + 2], float (&rNW)[64 / S + 2],
                                                     float (&rSW)[64 / S + 2])
  {
    constexpr int RH  = 64 / S, SW = 64 / S;
    const int     gx  = X0 + lane;
    const bool    cin = gx >= 0 && gx < d.W;
    const int     cx  = min(max(gx, 0), d.W - 1);
    if constexpr (C8)
    {
      f4 C[RH];
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const int cy = min(max(Y0 + r, 0), d.H - 1);
        C[r]         = gload<f4>(d.cost, (long long)cy * d.cost_pitch + 16ll * cx);
      }
      bool here[RH];
#pragma unroll
      for (int r = 0; r < RH; ++r)
        here[r] = cin && Y0 + r >= 0 && Y0 + r < d.H;
      {
        float tmp[RH];
#pragma unroll
        for (int r = 0; r < RH; ++r)
          tmp[r] = (here[r] && lane >= 1) ? C[r].x : DTK_INF;
        col_to_row<S>(tmp, rW, T, lane);
      }
#pragma unroll
      for (int r = 0; r < RH; ++r)
        cNW[r] = (here[r] && r >= 1 && lane >= 1) ? C[r].z : DTK_INF;
      col_to_row<S>(cNW, rNW, T, lane);
      {
        float tmp[RH];
#pragma unroll
        for (int r = 0; r < RH - 1; ++r)
        {
          const float ne = here[r + 1] ? C[r + 1].w : DTK_INF;
          const float sw = shr_z(ne);
          tmp[r]         = (here[r] && lane >= 1) ? sw : DTK_INF;
        }
        tmp[RH - 1] = DTK_INF;
        col_to_row<S>(tmp, rSW, T, lane);
      }
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        cNE[r] = (here[r] && r >= 1 && lane <= 62) ? C[r].w : DTK_INF;
        cN[r]  = (here[r] && r >= 1) ? C[r].y : DTK_INF;
      }
    }
    else
    {
      f2 C[RH];
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const int cy = min(max(Y0 + r, 0), d.H - 1);
        C[r]         = gload<f2>(d.cost, (long long)cy * d.cost_pitch + 8ll * cx);
      }
      {
        float tmp[RH];
#pragma unroll
        for (int r = 0; r < RH; ++r)
        {
          const bool here = cin && Y0 + r >= 0 && Y0 + r < d.H;
          tmp[r]          = (here && lane >= 1) ? C[r].x : DTK_INF;
        }
        col_to_row<S>(tmp, rW, T, lane);
      }
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const bool here = cin && Y0 + r >= 0 && Y0 + r < d.H;
        cN[r]           = (here && r >= 1) ? C[r].y : DTK_INF;
        cNW[r] = cNE[r] = DTK_INF;
      }
#pragma unroll
      for (int k = 0; k < SW + 2; ++k)
        rNW[k] = rSW[k] = DTK_INF;
    }
  }

  // ---- one visit of one tile --------------------------------------------------
  // Boundary-row exchange between the NW stacked waves of a tile.  Wave w's
  // rows 0 and 1 are the same image rows as wave w-1's rows RH-2 and RH-1
  // (its halo and first owned row), so after each pair of sweeps both copies
  // of those rows are merged by min through LDS (X: 2 parities x NW x 4 rows
  // x 64 floats; F: 2 x NW flag words).  Returns the OR over the waves of
  // their flags for the pair just swept (the same in every wave).
  template <int RH, int NW>
  __device__ __forceinline__ uint32_t exchange_rows(float (&Dc)[RH], float* X, uint32_t* F, int w, int lane, int par,
                                                    uint32_t flags, bool& any)
  {
    float* Xp = X + par * (NW * 4 * 64);
    Xp[(w * 4 + 0) * 64 + lane] = Dc[0];
    Xp[(w * 4 + 1) * 64 + lane] = Dc[1];
    Xp[(w * 4 + 2) * 64 + lane] = Dc[RH - 2];
    Xp[(w * 4 + 3) * 64 + lane] = Dc[RH - 1];
    if (lane == 0)
      F[par * NW + w] = flags;
    __syncthreads();
    bool lowered = false;
    if (w > 0)
    {
      const float a = Xp[((w - 1) * 4 + 2) * 64 + lane];
      const float b = Xp[((w - 1) * 4 + 3) * 64 + lane];
      lowered |= lt_nn(a, Dc[0]) | lt_nn(b, Dc[1]);
      Dc[0] = minf_nn(Dc[0], a);
      Dc[1] = minf_nn(Dc[1], b);
    }
    if (w < NW - 1)
    {
      const float a = Xp[((w + 1) * 4 + 0) * 64 + lane];
      const float b = Xp[((w + 1) * 4 + 1) * 64 + lane];
      lowered |= lt_nn(a, Dc[RH - 2]) | lt_nn(b, Dc[RH - 1]);
      Dc[RH - 2] = minf_nn(Dc[RH - 2], a);
      Dc[RH - 1] = minf_nn(Dc[RH - 1], b);
    }
    any |= ballot(lowered) != 0;
    uint32_t all = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k)
      all |= F[par * NW + k];
    return all;
  }

  // OR of a wave-uniform flag over the NW waves of a tile (F[3*NW ..]).  One
  // slot per wave suffices: an exchange barrier separates two votes.
  template <int NW>
  __device__ __forceinline__ bool any_wave(bool v, uint32_t* F, int w, int lane)
  {
    if (lane == 0)
      F[3 * NW + w] = v ? 1u : 0u;
    __syncthreads();
    uint32_t all = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k)
      all |= F[3 * NW + k];
    return all != 0;
  }

  // A tile is solved by NW stacked waves (one 64 x RH region each, RH - 2
  // owned rows each).  T: LDS RH*65 floats of this wave; Sv: LDS 2*RH floats
  // of this wave; X, F: shared exchange buffers (NW > 1 only).  Loads the
  // regions, iterates to the tile's local fixed point (or max_sweeps),
  // writes back the owned pixels and reports which neighbours saw a border
  // pixel improve (border and changed are the same in every wave).
  // CM: 0 = one add per edge with per-pixel costs (lambda 1, or lambda 0
  // with masked costs), 1 = two adds (0 < lambda < 1), 2 = lambda 0 with
  // ConstCosts.
  //
  // KEYS: also report, per neighbour k (border-bit order), the smallest value
  // of this visit's improved halo pixels that belong to neighbour k (float
  // bits; +inf when none): Kw[8 * w + k] for wave w, written before the
  // visit's last workgroup barrier, when Kw is not null (a launch-wide
  // choice: the KEY order queues a woken neighbour by it, the other orders
  // pass null and skip the reductions).
  //
  // AFTER: called with each visit's result once its stores are issued (every
  // thread of the workgroup, the same result in each).  It returns true to
  // visit the tile again at once (the same in every thread): the edge costs
  // stay in registers, the owned pixels in registers equal the ones just
  // stored, and only the one-pixel ring around the tile is loaded again
  // (the caller has ordered those loads after the stores of the neighbours
  // that woke the tile).  The default never visits again.
  struct NoRevisit
  {
    __device__ bool operator()(const VisitResult&) const { return false; }
  };

  // FRESH: with constant costs, test the wakes against the neighbours'
  // current values (below).  The task schedule gains from it; the blocks
  // schedule, whose rounds wait for their slowest tile, does not.
  template <int S, bool C8, int CM, int NW, bool KEYS = false, bool FRESH = false, class After = NoRevisit>
  __device__ __attribute__((always_inline)) VisitResult visit_tile(const ImgDesc& d, const KParams& p, int tx, int ty,
                                                                   float* T, float* Sv, float* X, uint32_t* F,
                                                                   uint32_t* Kw = nullptr, After after = After{})
  {
    constexpr int RH   = 64 / S;
    constexpr int SW   = 64 / S;
    const int     lane = threadIdx.x & 63;
    const int     w    = NW > 1 ? wave_index() : 0;
    const int     X0   = tx * kTileW - 1;
    const int     Y0   = ty * (NW * (RH - 2)) - 1 + w * (RH - 2);
    const int     gx   = X0 + lane;
    const bool    cin  = gx >= 0 && gx < d.W;
    constexpr bool TWO  = CM == 1;
    uint64_t       clk0 = __builtin_amdgcn_s_memtime();

    using Costs = std::conditional_t<CM == 2, ConstCosts, TileCosts<S>>;
    Costs cc;
    if constexpr (CM == 2)
    {
      const int rr = lane % RH;
      cc.ea        = p.e_axis;
      cc.ed        = p.e_diag;
      cc.colNW     = (C8 && lane != 0) ? p.e_diag : DTK_INF;
      cc.colNE     = (C8 && lane != 63) ? p.e_diag : DTK_INF;
      cc.rowNW     = (C8 && rr != 0) ? p.e_diag : DTK_INF;
      cc.rowSW     = (C8 && rr != RH - 1) ? p.e_diag : DTK_INF;
    }
    else if (p.zero_lambda)
      setup_costs_const<S, C8>(d, p, X0, Y0, lane, T, cc.cN_, cc.cNW_, cc.cNE_, cc.rW_, cc.rNW_, cc.rSW_);
    else if (d.cost)
      setup_costs_cached<S, C8>(d, X0, Y0, lane, T, cc.cN_, cc.cNW_, cc.cNE_, cc.rW_, cc.rNW_, cc.rSW_);
    else
      switch (p.img_kind)
      {
      case IMG_U8:
        setup_costs<IMG_U8, S, C8>(d, p, X0, Y0, lane, T, cc.cN_, cc.cNW_, cc.cNE_, cc.rW_, cc.rNW_, cc.rSW_);
        break;
      case IMG_RGB8:
        setup_costs<IMG_RGB8, S, C8>(d, p, X0, Y0, lane, T, cc.cN_, cc.cNW_, cc.cNE_, cc.rW_, cc.rNW_, cc.rSW_);
        break;
      default:
        setup_costs<IMG_F32, S, C8>(d, p, X0, Y0, lane, T, cc.cN_, cc.cNW_, cc.cNE_, cc.rW_, cc.rNW_, cc.rSW_);
        break;
      }

    // D of the region, +inf outside the image.  Loaded at the start of every
    // visit as a min into the registers: the first visit starts from +inf;
    // a kept tile's next visit takes the new values of the ring around the
    // tile (lanes 0 and 63, the top wave's row 0 and the bottom wave's last
    // row).  Every other pixel in memory equals its register (this tile
    // stored it, and its stacked waves merged their shared rows), so the min
    // over the whole region needs no divergent loads.
    float     Dc[RH];
    const int cx = min(max(gx, 0), d.W - 1);
#pragma unroll
    for (int r = 0; r < RH; ++r)
      Dc[r] = DTK_INF;

    const float ea = p.e_axis, ed = p.e_diag;
    VisitResult res;
    for (;;)
    {
      // The column passes through an empty asm at every visit, so the row
      // addresses of the loads and stores are formed where they are used
      // instead of being kept live through the sweeps.
      int             vx     = cx;
      const long long dpitch = d.D_pitch;
      asm volatile("" : "+v"(vx));
#pragma unroll
      for (int r = 0; r < RH; ++r)
      {
        const int   gy = Y0 + r;
        const int   cy = min(max(gy, 0), d.H - 1);
        const float v  = gload<float>(d.D, (long long)cy * dpitch + 4ll * vx);
        Dc[r]          = minf_nn(Dc[r], (cin && gy >= 0 && gy < d.H) ? v : DTK_INF);
      }

      // Halo as loaded: a neighbour needs a new visit exactly when one of its
      // pixels in this tile's halo ends up strictly smaller (this visit then
      // found a shorter path into that neighbour).
      const float init_top = Dc[0];
      const float init_bot = Dc[RH - 1];
      if (lane == 0 || lane == 63)
      {
        const int o = (lane == 63) ? RH : 0;
#pragma unroll
        for (int r = 0; r < RH; ++r)
          Sv[o + r] = Dc[r];
      }

      const uint64_t clk1 = __builtin_amdgcn_s_memtime();
      // A visit needs four quiet sweeps to certify its region, so a cap below
      // four could never converge (the tile would re-queue itself forever).
      const int   cap = p.max_sweeps > 0 ? (p.max_sweeps < 4 ? 4 : p.max_sweeps) : 0x7fffffff;
      int         quiet = 0, sweeps = 0;
      bool        conv = false, any = false;
      float       R[SW + 2];
      // The current stored values of the region's halo pixels, which other
      // tiles own, for the wake test after the sweeps (constant costs only):
      // a neighbour needs a visit when this visit found a value below the one
      // it holds now.  Comparing with the values as loaded at the start of the
      // visit instead also wakes a neighbour whose own concurrent visit has
      // already stored as good a value, for a visit that changes nothing (a
      // fifth of the visits at lambda 0).  Agent-scope loads, not cached in
      // this L2, see the neighbours' latest stores; they are issued before
      // this visit's own stores.  With per-pixel costs the reload is not
      // worth its latency (the loads cannot be overlapped without spilling).
      float       hv[RH];
      auto        now = [&](float loaded, int r) {
        if constexpr (FRESH && CM == 2)
          return minf_nn(loaded, hv[r]);
        else
          return loaded;
      };
      auto        reload_halo = [&]() {
#pragma unroll
        for (int r = 0; r < RH; ++r)
        {
          const int cy = min(max(Y0 + r, 0), d.H - 1);
          hv[r]        = __hip_atomic_load(reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(d.D) +
                                                                          (long long)cy * dpitch) + vx,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      };
      if constexpr (NW == 1)
      {
        for (;;)
        {
          uint64_t edge_unused = 0;
          uint64_t ch          = sweep_col_fwd<RH, C8, TWO>(Dc, cc, ea, ed, edge_unused);
          ++sweeps;
          any |= ch != 0;
          quiet = ch ? 0 : quiet + 1;
          if (quiet >= 4)
          {
            conv = true;
            break;
          }
          if (sweeps >= cap)
            break;
          ch = sweep_col_bwd<RH, C8, TWO>(Dc, cc, ea, ed);
          ++sweeps;
          any |= ch != 0;
          quiet = ch ? 0 : quiet + 1;
          if (quiet >= 4)
          {
            conv = true;
            break;
          }
          if (sweeps >= cap)
            break;

          col_to_row<S>(Dc, R, T, lane);
          const uint64_t cf = sweep_row_fwd<SW, C8, TWO>(R, cc, ea, ed, edge_unused);
          ++sweeps;
          any |= cf != 0;
          quiet = cf ? 0 : quiet + 1;
          if (quiet >= 4 || sweeps >= cap)
          {
            conv = quiet >= 4;
            row_to_col<S>(R, Dc, T, lane);
            break;
          }
          const uint64_t cb = sweep_row_bwd<SW, C8, TWO>(R, cc, ea, ed);
          ++sweeps;
          any |= cb != 0;
          quiet = (cb | cf) ? 0 : quiet + 1;
          row_to_col<S>(R, Dc, T, lane);
          if (quiet >= 4)
          {
            conv = true;
            break;
          }
          if (sweeps >= cap)
            break;
        }
      }
      else
      {
        int par = 0;
        if constexpr (CM != 2)
        {
          // Per-pixel costs (lambda > 0): one exchange per column pair and
          // row pair together.  A wave whose column pair is clean and whose
          // W / E check passes is at its region's fixed point and skips the
          // row pair; otherwise it runs the row pair and is at the fixed
          // point only if neither pair changed anything.  Either way it
          // changed no shared row then, so when every wave reports a fixed
          // point the merge changes nothing and the tile is stable.  Half the
          // workgroup barriers of the exchange after every pair (each costs
          // the wait for the slowest of the four waves), and the W / E
          // certification needs no vote of its own.
          int iters = 0;
          for (;;)
          {
            uint64_t       edge = 0;
            const uint64_t cf   = sweep_col_fwd<RH, C8, TWO>(Dc, cc, ea, ed, edge);
            const uint64_t cb   = sweep_col_bwd<RH, C8, TWO>(Dc, cc, ea, ed);
            sweeps += 2;
            ++iters;
            any |= (cf | cb) != 0;
            col_to_row<S>(Dc, R, T, lane);
            bool fixed = (cb | edge) == 0 && !row_we_violated<SW, TWO>(R, cc, ea);
            if (!fixed)
            {
              uint64_t       redge = 0;
              const uint64_t rf    = sweep_row_fwd<SW, C8, TWO>(R, cc, ea, ed, redge);
              const uint64_t rb    = sweep_row_bwd<SW, C8, TWO>(R, cc, ea, ed);
              row_to_col<S>(R, Dc, T, lane);
              sweeps += 2;
              any |= (rf | rb) != 0;
              fixed = (cf | cb | rf | rb) == 0;
            }
            const uint32_t all = exchange_rows<RH, NW>(Dc, X, F, w, lane, par, fixed ? 0u : 1u, any);
            par ^= 1;
            if (!(all & 1u))
            {
              conv = true;
              break;
            }
            // The cap counts four sweeps per exchange in every wave, so the
            // waves leave the loop together.
            if (4 * iters >= cap)
              break;
          }
        }
        else
        {
          // Constant costs (lambda = 0, 32-row waves, which gain from merging
          // more often): sweep pairs (column down+up, row right+left) in every
          // wave at once, the shared rows merged after each pair.  A forward
          // sweep leaves its direction's edges satisfied except where it read
          // a copy that it then changed: rows a neighbour wave also holds, and
          // (row layout) a segment's first and last columns.  So a pair whose
          // backward sweep changed nothing and whose forward sweep did not
          // touch those leaves the tile stable in its direction ("clean"), and
          // one pair of the other direction without any change then certifies
          // the tile.  Before the first pair the tile counts as not clean,
          // which makes an untouched tile take two quiet pairs.  After a clean
          // column pair the W / E check (row_we_violated, voted over the
          // waves) can certify the tile in place of the row pair.
          bool prev_clean = false;
          for (;;)
          {
            uint64_t       edge = 0;
            const uint64_t cf   = sweep_col_fwd<RH, C8, TWO>(Dc, cc, ea, ed, edge);
            const uint64_t cb   = sweep_col_bwd<RH, C8, TWO>(Dc, cc, ea, ed);
            sweeps += 2;
            any |= (cf | cb) != 0;
            uint32_t all = exchange_rows<RH, NW>(
              Dc, X, F, w, lane, par, ((cf | cb) != 0 ? 1u : 0u) | ((cb | edge) != 0 ? 2u : 0u), any);
            par ^= 1;
            if (!(all & 1u) && prev_clean)
            {
              conv = true;
              break;
            }
            prev_clean = !(all & 2u);
            if (sweeps >= cap)
              break;
            col_to_row<S>(Dc, R, T, lane);
            // A clean column pair leaves only the W and E edges to check; the
            // column layout is still current, so a pass needs no transpose back.
            if (prev_clean && !any_wave<NW>(row_we_violated<SW, TWO>(R, cc, ea), F, w, lane))
            {
              conv = true;
              break;
            }
            edge              = 0;
            const uint64_t rf = sweep_row_fwd<SW, C8, TWO>(R, cc, ea, ed, edge);
            const uint64_t rb = sweep_row_bwd<SW, C8, TWO>(R, cc, ea, ed);
            row_to_col<S>(R, Dc, T, lane);
            sweeps += 2;
            any |= (rf | rb) != 0;
            edge |= rf & shared_row_lanes<S>();
            all = exchange_rows<RH, NW>(Dc, X, F, w, lane, par,
                                        ((rf | rb) != 0 ? 1u : 0u) | ((rb | edge) != 0 ? 2u : 0u), any);
            par ^= 1;
            if (!(all & 1u) && prev_clean)
            {
              conv = true;
              break;
            }
            prev_clean = !(all & 2u);
            if (sweeps >= cap)
              break;
          }
        }
      }

      const uint64_t clk2 = __builtin_amdgcn_s_memtime();
      res.sweeps    = sweeps;
      res.converged = conv;
      res.changed   = any;
      res.border    = 0;
      uint32_t b    = 0;
      if (any)
      {
        if constexpr (FRESH && CM == 2)
          reload_halo();
        // Write back the owned pixels.
        if (lane >= 1 && lane <= 62 && cin)
        {
#pragma unroll
          for (int r = 1; r < RH - 1; ++r)
          {
            const int gy = Y0 + r;
            if (gy < d.H) // device-scope store: written through, so the
                          // release fence has no dirty L2 lines of D to write back
              __hip_atomic_store(reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d.D) + (long long)gy * dpitch) + vx,
                                 Dc[r], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          }
        }

        // Only pixels inside the image count: with constant costs the pixels
        // outside it start at +inf on every visit and always decrease, which
        // would wake the neighbours (and through them this tile) forever.
        bool     colimp = false;
        uint32_t colkey = 0x7f800000u; // smallest improved halo-column value (lanes 0 and 63)
        if ((lane == 0 || lane == 63) && cin)
        {
          const int o = (lane == 63) ? RH : 0;
#pragma unroll
          for (int r = 1; r < RH - 1; ++r)
          {
            const bool imp = Y0 + r < d.H && lt_nn(Dc[r], now(Sv[o + r], r));
            colimp |= imp;
            if constexpr (KEYS)
              colkey = imp ? min(colkey, __float_as_uint(Dc[r])) : colkey;
          }
        }
        const bool     timp = cin && Y0 >= 0 && lt_nn(Dc[0], now(init_top, 0));
        const bool     bimp = cin && Y0 + RH - 1 < d.H && lt_nn(Dc[RH - 1], now(init_bot, RH - 1));
        const uint64_t top  = ballot(timp);
        const uint64_t bot  = ballot(bimp);
        const uint64_t cols = ballot(colimp);
        if (KEYS && Kw != nullptr)
        {
          // Per wave: L, R from the halo columns; T, TL, TR from row 0 of
          // the top wave; B, BL, BR from the last row of the bottom wave.
          const bool     mid_lane = lane >= 1 && lane <= 62;
          uint32_t       tk = (timp && mid_lane) ? __float_as_uint(Dc[0]) : 0x7f800000u;
          uint32_t       bk = (bimp && mid_lane) ? __float_as_uint(Dc[RH - 1]) : 0x7f800000u;
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1)
          {
            tk = min(tk, (uint32_t)__shfl_xor((int)tk, o));
            bk = min(bk, (uint32_t)__shfl_xor((int)bk, o));
          }
          const uint32_t c0  = (uint32_t)__shfl((int)colkey, 0), c63 = (uint32_t)__shfl((int)colkey, 63);
          const uint32_t t0  = (uint32_t)__shfl((int)(timp ? __float_as_uint(Dc[0]) : 0x7f800000u), 0);
          const uint32_t t63 = (uint32_t)__shfl((int)(timp ? __float_as_uint(Dc[0]) : 0x7f800000u), 63);
          const uint32_t b0  = (uint32_t)__shfl((int)(bimp ? __float_as_uint(Dc[RH - 1]) : 0x7f800000u), 0);
          const uint32_t b63 = (uint32_t)__shfl((int)(bimp ? __float_as_uint(Dc[RH - 1]) : 0x7f800000u), 63);
          if (lane < 8)
          {
            uint32_t k = 0x7f800000u;
            k          = lane == 0 ? c0 : k;
            k          = lane == 1 ? c63 : k;
            if (w == 0)
            {
              k = lane == 2 ? tk : k;
              k = lane == 4 ? t0 : k;
              k = lane == 5 ? t63 : k;
            }
            if (w == NW - 1)
            {
              k = lane == 3 ? bk : k;
              k = lane == 6 ? b0 : k;
              k = lane == 7 ? b63 : k;
            }
            Kw[8 * w + lane] = k;
          }
        }
        const uint64_t mid  = 0x7FFFFFFFFFFFFFFEull; // lanes 1..62
        b |= (uint32_t)(cols & 1u);                 // L : left halo column
        b |= (uint32_t)((cols >> 63) & 1u) << 1;    // R : right halo column
        if (w == 0)
        {
          b |= ((top & mid) != 0 ? 1u : 0u) << 2;   // T : top halo row
          b |= (uint32_t)(top & 1u) << 4;           // TL: top-left corner
          b |= (uint32_t)((top >> 63) & 1u) << 5;   // TR
        }
        if (w == NW - 1)
        {
          b |= ((bot & mid) != 0 ? 1u : 0u) << 3;   // B : bottom halo row
          b |= (uint32_t)(bot & 1u) << 6;           // BL
          b |= (uint32_t)((bot >> 63) & 1u) << 7;   // BR
        }

      }
      if (KEYS && Kw != nullptr)
      {
        if (!any && lane < 8)
          Kw[8 * w + lane] = 0x7f800000u;
      }
      if constexpr (NW == 1)
      {
        res.border = b;
        if constexpr (KEYS)
          lds_wave_sync();
      }
      else
      {
        // Border bits and change flags of all waves (F[2*NW ..]).
        if (lane == 0)
          F[2 * NW + w] = b | (any ? 0x100u : 0u);
        __syncthreads();
        uint32_t all = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k)
          all |= F[2 * NW + k];
        res.border  = all & 0xFFu;
        res.changed = (all & 0x100u) != 0;
      }
      res.cyc_setup = (uint32_t)(clk1 - clk0);
      res.cyc_sweep = (uint32_t)(clk2 - clk1);
      res.cyc_store = (uint32_t)(__builtin_amdgcn_s_memtime() - clk2);
      if (!after(res))
        return res;
      clk0 = __builtin_amdgcn_s_memtime();
    }
  }
} // namespace dtk
