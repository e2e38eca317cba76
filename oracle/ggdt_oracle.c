/*
 * CPU oracle for the generalised geodesic distance transform (GGDT).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path links, loads or
 * calls this file: it is the checker that the HIP kernels are compared with
 * (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 *
 * What it restates, and from where (paths relative to the reference tree):
 *
 *   ora_init            dt/src/utils.cu:5-13
 *                       D(x,y) = v * (mask(x,y) == 0)  (mask != 0 marks a seed)
 *                       soft-mask extension: D0 = nu * (1 - s)      (no reference)
 *   ora_raster          dt/src/generalised_distance_transform.cu:101-134
 *                       iteration = row-fwd, row-bwd, col-fwd, col-bwd passes
 *                       (:129-132), repeated until an iteration in which no
 *                       pixel strictly decreased (:126-134).
 *     row pass          dt/src/generalised_distance_transform.cu:17-57
 *                       x runs 1..W-1 (fwd) or W-2..0 (bwd); neighbours
 *                       dy = -1, 0, +1 in the previous column, out-of-range
 *                       skipped; cost = (D(q) + l_eucl*ld) + l_grad*|I(p)-I(q)|
 *                       evaluated as two rounded float adds (no contraction);
 *                       new = std::min(new, cur); store only if new < D.
 *     column pass       dt/src/generalised_distance_transform.cu:59-99
 *     |I(p)-I(q)|       dt/src/utils.cuh:37-41 (uint8 abs difference)
 *     l_grad, l_eucl    dt/src/generalised_distance_transform.cu:108-111
 *                       (float lambda, 1 - lambda, sqrtf(2))
 *   ora_raster_omp      the same schedule with the pixels of each column step
 *                       (row passes) or row step (column passes) spread over
 *                       OpenMP threads; every step reads only the previous
 *                       step, so results are identical to ora_raster.
 *   ora_dijkstra        independent float32 label-setting Dijkstra with the
 *                       same edge cost; pins the fixed point on tiny inputs.
 *   ora_pred            argmin predecessor codes of a converged D.
 *   ora_backward        reference gradient of sum(W*D) w.r.t. image and soft
 *                       mask through the predecessor forest (double sums).
 *   ora_raster_f64      the raster schedule in double precision, for central
 *                       finite differences.
 *
 * Extensions with no reference counterpart (definitions of this project):
 *   - C4 connectivity: only the axis neighbour (dy = 0 / dx = 0) is relaxed.
 *   - RGB8 cost: sqrtf((float)(dr*dr + dg*dg + db*db)).
 *   - float32 image: fabsf(I(p) - I(q)); an edge whose image term is NaN
 *     (a NaN pixel, or two infinite pixels of the same sign) never wins.
 *   - lambda = 0: no image term at all; every edge costs its chamfer step,
 *     whatever the pixel values (for uint8 and RGB8 pixels 0 * |dI| is +0
 *     anyway; for a NaN or infinite float pixel it would be NaN).
 *   - soft float mask s in [0,1]: D0 = nu * (1 - s).
 *   - fixed pass count: exactly N iterations of the four passes.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_IMG_U8 0
#define ORA_IMG_RGB8 1
#define ORA_IMG_F32 2

#define ORA_MASK_U8 0
#define ORA_MASK_F32 1

typedef struct
{
  int            kind;
  const uint8_t* u8; /* grey: W*H bytes, rgb8: W*H*3 bytes (r,g,b) */
  const float*   f32;
  float          l_grad;
  float          e_axis; /* l_eucl * 1.f            */
  float          e_diag; /* l_eucl * sqrtf(2.f)     */
} cost_ctx;

/* |I(p) - I(q)| with the per-kind definition, as a float. */
static inline float local_dist(const cost_ctx* c, long p, long q)
{
  switch (c->kind)
  {
  case ORA_IMG_U8: {
    const uint8_t a = c->u8[p], b = c->u8[q];
    const uint8_t d = (uint8_t)(a < b ? b - a : a - b);
    return (float)d;
  }
  case ORA_IMG_RGB8: {
    const int dr = (int)c->u8[3 * p + 0] - (int)c->u8[3 * q + 0];
    const int dg = (int)c->u8[3 * p + 1] - (int)c->u8[3 * q + 1];
    const int db = (int)c->u8[3 * p + 2] - (int)c->u8[3 * q + 2];
    const int n  = dr * dr + dg * dg + db * db;
    return sqrtf((float)n);
  }
  default:
    return fabsf(c->f32[p] - c->f32[q]);
  }
}

/* cur = (D(q) + l_eucl*ld) + l_grad*dist, two rounded adds. */
static inline float relax_cost(const cost_ctx* c, float dq, float e, long p, long q)
{
  const float g   = c->l_grad == 0.f ? 0.f : c->l_grad * local_dist(c, p, q);
  float       cur = dq + e;
  cur             = cur + g;
  return cur;
}

static void make_ctx(cost_ctx* c, int kind, const void* img, float lambda)
{
  c->kind         = kind;
  c->u8           = (kind == ORA_IMG_F32) ? NULL : (const uint8_t*)img;
  c->f32          = (kind == ORA_IMG_F32) ? (const float*)img : NULL;
  c->l_grad       = lambda;
  const float l_e = 1.f - lambda;
  c->e_axis       = l_e * 1.f;
  c->e_diag       = l_e * sqrtf(2.f);
}

void ora_init(const void* mask, int mask_kind, float* D, int W, int H, float v)
{
  const long n = (long)W * H;
  if (mask_kind == ORA_MASK_U8)
  {
    const uint8_t* m = (const uint8_t*)mask;
    for (long i = 0; i < n; ++i)
      D[i] = v * (float)(m[i] == 0);
  }
  else
  {
    const float* s = (const float*)mask;
    for (long i = 0; i < n; ++i)
      D[i] = v * (1.f - s[i]);
  }
}

/* One row pass (dt/src/generalised_distance_transform.cu:17-57).  Returns 1
 * if any pixel strictly decreased.  Pixels of column x are independent of
 * each other, so the y loop may be split across threads. */
static int row_pass(const cost_ctx* c, float* D, int W, int H, int forward, int conn8, int nthreads)
{
  const int start = forward ? 1 : W - 2;
  const int end   = forward ? W : -1;
  const int inc   = forward ? 1 : -1;
  const int dx    = -inc;
  int       changed = 0;
  for (int x = start; x != end; x += inc)
  {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : changed) if (nthreads > 1)
    for (int y = 0; y < H; ++y)
    {
      const long p   = (long)y * W + x;
      float      nd  = D[p];
      for (int dy = -1; dy < 2; ++dy)
      {
        if (!conn8 && dy != 0)
          continue;
        const int ny = y + dy;
        if (ny < 0 || ny >= H)
          continue;
        const long  q   = (long)ny * W + x + dx;
        const float cur = relax_cost(c, D[q], dy == 0 ? c->e_axis : c->e_diag, p, q);
        nd              = (cur < nd) ? cur : nd; /* std::min(nd, cur) */
      }
      if (nd < D[p])
      {
        D[p]    = nd;
        changed = 1;
      }
    }
  }
  return changed;
}

/* One column pass (dt/src/generalised_distance_transform.cu:59-99). */
static int col_pass(const cost_ctx* c, float* D, int W, int H, int forward, int conn8, int nthreads)
{
  const int start = forward ? 1 : H - 2;
  const int end   = forward ? H : -1;
  const int inc   = forward ? 1 : -1;
  const int dy    = -inc;
  int       changed = 0;
  for (int y = start; y != end; y += inc)
  {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : changed) if (nthreads > 1)
    for (int x = 0; x < W; ++x)
    {
      const long p  = (long)y * W + x;
      float      nd = D[p];
      for (int dx = -1; dx < 2; ++dx)
      {
        if (!conn8 && dx != 0)
          continue;
        const int nx = x + dx;
        if (nx < 0 || nx >= W)
          continue;
        const long  q   = (long)(y + dy) * W + nx;
        const float cur = relax_cost(c, D[q], dx == 0 ? c->e_axis : c->e_diag, p, q);
        nd              = (cur < nd) ? cur : nd;
      }
      if (nd < D[p])
      {
        D[p]    = nd;
        changed = 1;
      }
    }
  }
  return changed;
}

static int raster_impl(const void* img, int img_kind, const void* mask, int mask_kind, float* D, int W, int H,
                       float lambda, float v, int conn8, int max_iters, int nthreads)
{
  cost_ctx c;
  make_ctx(&c, img_kind, img, lambda);
  ora_init(mask, mask_kind, D, W, H, v);
  if (W <= 0 || H <= 0)
    return 0;
  int iters   = 0;
  int changed = 1;
  while (changed && (max_iters <= 0 || iters < max_iters))
  {
    changed = 0;
    changed |= row_pass(&c, D, W, H, 1, conn8, nthreads);
    changed |= row_pass(&c, D, W, H, 0, conn8, nthreads);
    changed |= col_pass(&c, D, W, H, 1, conn8, nthreads);
    changed |= col_pass(&c, D, W, H, 0, conn8, nthreads);
    ++iters;
  }
  return iters;
}

/* Converged when max_iters <= 0, else exactly min(max_iters, converged)
 * iterations.  Returns the number of iterations run (the last converged
 * iteration is the one that changed nothing). */
int ora_raster(const void* img, int img_kind, const void* mask, int mask_kind, float* D, int W, int H, float lambda,
               float v, int conn8, int max_iters)
{
  return raster_impl(img, img_kind, mask, mask_kind, D, W, H, lambda, v, conn8, max_iters, 1);
}

int ora_raster_omp(const void* img, int img_kind, const void* mask, int mask_kind, float* D, int W, int H,
                   float lambda, float v, int conn8, int max_iters, int nthreads)
{
  return raster_impl(img, img_kind, mask, mask_kind, D, W, H, lambda, v, conn8, max_iters, nthreads);
}

int ora_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* Independent check: float32 label-setting Dijkstra with the same edge    */
/* cost.  f(x) = fl(fl(x + e) + g) is monotone with f(x) >= x, so the      */
/* settled labels are the minimum over paths of left-to-right float sums,  */
/* which is also the unique fixed point every relaxation schedule reaches. */
/* ---------------------------------------------------------------------- */

typedef struct
{
  float d;
  long  i;
} heap_item;

static void heap_push(heap_item* h, long* n, heap_item it)
{
  long k = (*n)++;
  h[k]   = it;
  while (k > 0)
  {
    long p = (k - 1) / 2;
    if (h[p].d < h[k].d || (h[p].d == h[k].d && h[p].i <= h[k].i))
      break;
    heap_item t = h[p];
    h[p]        = h[k];
    h[k]        = t;
    k           = p;
  }
}

static heap_item heap_pop(heap_item* h, long* n)
{
  heap_item top = h[0];
  h[0]          = h[--(*n)];
  long k        = 0;
  for (;;)
  {
    long l = 2 * k + 1, r = l + 1, m = k;
    if (l < *n && (h[l].d < h[m].d || (h[l].d == h[m].d && h[l].i < h[m].i)))
      m = l;
    if (r < *n && (h[r].d < h[m].d || (h[r].d == h[m].d && h[r].i < h[m].i)))
      m = r;
    if (m == k)
      break;
    heap_item t = h[m];
    h[m]        = h[k];
    h[k]        = t;
    k           = m;
  }
  return top;
}

static const int OFF_X[8] = {-1, -1, 0, 1, 1, 1, 0, -1}; /* dt/include/dt/c8.hpp:19-28 order */
static const int OFF_Y[8] = {0, -1, -1, -1, 0, 1, 1, 1};

int ora_dijkstra(const void* img, int img_kind, const void* mask, int mask_kind, float* D, int W, int H, float lambda,
                 float v, int conn8)
{
  cost_ctx c;
  make_ctx(&c, img_kind, img, lambda);
  ora_init(mask, mask_kind, D, W, H, v);
  const long n = (long)W * H;
  if (n == 0)
    return 0;
  unsigned char* done = (unsigned char*)calloc((size_t)n, 1);
  heap_item*     heap = (heap_item*)malloc(sizeof(heap_item) * (size_t)(n * 9 + 1));
  if (!done || !heap)
  {
    free(done);
    free(heap);
    return -1;
  }
  long hn = 0;
  for (long i = 0; i < n; ++i)
  {
    heap_item it = {D[i], i};
    heap_push(heap, &hn, it);
  }
  while (hn > 0)
  {
    heap_item it = heap_pop(heap, &hn);
    if (done[it.i] || it.d != D[it.i])
      continue;
    done[it.i] = 1;
    const int x = (int)(it.i % W), y = (int)(it.i / W);
    for (int k = 0; k < 8; ++k)
    {
      const int diag = (OFF_X[k] != 0 && OFF_Y[k] != 0);
      if (!conn8 && diag)
        continue;
      const int nx = x + OFF_X[k], ny = y + OFF_Y[k];
      if (nx < 0 || ny < 0 || nx >= W || ny >= H)
        continue;
      const long p = (long)ny * W + nx;
      if (done[p])
        continue;
      const float cur = relax_cost(&c, D[it.i], diag ? c.e_diag : c.e_axis, p, it.i);
      if (cur < D[p])
      {
        D[p]          = cur;
        heap_item nit = {cur, p};
        heap_push(heap, &hn, nit);
      }
    }
  }
  free(done);
  free(heap);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Predecessors and the reference gradient.                                */
/* ---------------------------------------------------------------------- */

/* Code 0..7 = index into the c8 offset table above (neighbour q = p + off),
 * 8 = root (D(p) == D0(p)).  Among neighbours, the first k (table order)
 * whose relaxation reproduces D(p) exactly and whose (D, index) pair is
 * lexicographically smaller than p's wins; this keeps the forest acyclic. */
void ora_pred(const void* img, int img_kind, const void* mask, int mask_kind, const float* D, uint8_t* pred, int W,
              int H, float lambda, float v, int conn8)
{
  cost_ctx c;
  make_ctx(&c, img_kind, img, lambda);
  const long n  = (long)W * H;
  float*     D0 = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  ora_init(mask, mask_kind, D0, W, H, v);
  for (long p = 0; p < n; ++p)
  {
    const int x = (int)(p % W), y = (int)(p / W);
    uint8_t   code = 8;
    if (!(D[p] == D0[p]))
    {
      for (int k = 0; k < 8; ++k)
      {
        const int diag = (OFF_X[k] != 0 && OFF_Y[k] != 0);
        if (!conn8 && diag)
          continue;
        const int nx = x + OFF_X[k], ny = y + OFF_Y[k];
        if (nx < 0 || ny < 0 || nx >= W || ny >= H)
          continue;
        const long q = (long)ny * W + nx;
        if (!(D[q] < D[p] || (D[q] == D[p] && q < p)))
          continue;
        const float cur = relax_cost(&c, D[q], diag ? c.e_diag : c.e_axis, p, q);
        if (cur == D[p])
        {
          code = (uint8_t)k;
          break;
        }
      }
    }
    pred[p] = code;
  }
  free(D0);
}

static const float* g_sort_D;
static int          cmp_desc(const void* a, const void* b)
{
  const long ia = *(const long*)a, ib = *(const long*)b;
  const float da = g_sort_D[ia], db = g_sort_D[ib];
  if (da != db)
    return da > db ? -1 : 1;
  return ia > ib ? -1 : (ia < ib ? 1 : 0);
}

/* Gradient of L = sum(Wt * D) for a float32 image and (optionally) a soft
 * float mask, through the predecessor forest:
 *   G(p)      = Wt(p) + sum_{children c of p} G(c)
 *   edge q->p: dL/dI(p) += G(p) * l_grad * sign(I(p) - I(q)), dL/dI(q) -= same
 *   root r   : dL/ds(r) = -nu * G(r)
 * Children have a (D, index) pair strictly greater than their parent, so a
 * decreasing sort of (D, index) visits every child before its parent.
 * Sums are carried in double. */
void ora_backward(const float* img, const float* D, const uint8_t* pred, const float* Wt, double* dimg, double* dmask,
                  double* G, int W, int H, float lambda, float nu)
{
  const long n   = (long)W * H;
  long*      ord = (long*)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
  for (long i = 0; i < n; ++i)
  {
    ord[i] = i;
    G[i]   = Wt[i];
    dimg[i] = 0.0;
    if (dmask)
      dmask[i] = 0.0;
  }
  g_sort_D = D;
  qsort(ord, (size_t)n, sizeof(long), cmp_desc);
  for (long k = 0; k < n; ++k)
  {
    const long p = ord[k];
    const int  code = pred[p];
    if (code == 8)
    {
      if (dmask)
        dmask[p] = -(double)nu * G[p];
      continue;
    }
    const int  x = (int)(p % W), y = (int)(p / W);
    const long q = (long)(y + OFF_Y[code]) * W + (x + OFF_X[code]);
    G[q] += G[p];
    const float  diff = img[p] - img[q];
    const double s    = diff > 0.f ? 1.0 : (diff < 0.f ? -1.0 : 0.0);
    const double gi   = G[p] * (double)lambda * s;
    dimg[p] += gi;
    dimg[q] -= gi;
  }
  free(ord);
}

/* The raster schedule in double precision for a float32 image and soft mask
 * (finite differences).  Converges; returns iterations. */
int ora_raster_f64(const double* img, const double* smask, double* D, int W, int H, double lambda, double nu,
                   int conn8)
{
  const long   n      = (long)W * H;
  const double e_axis = (1.0 - lambda);
  const double e_diag = (1.0 - lambda) * sqrt(2.0);
  for (long i = 0; i < n; ++i)
    D[i] = nu * (1.0 - smask[i]);
  if (n == 0)
    return 0;
  int iters = 0, changed = 1;
  while (changed)
  {
    changed = 0;
    for (int pass = 0; pass < 4; ++pass)
    {
      const int rows = pass < 2, fwd = (pass % 2) == 0;
      const int len   = rows ? W : H;
      const int start = fwd ? 1 : len - 2, end = fwd ? len : -1, inc = fwd ? 1 : -1;
      for (int s = start; s != end; s += inc)
      {
        const int other = rows ? H : W;
        for (int t = 0; t < other; ++t)
        {
          const int  x = rows ? s : t, y = rows ? t : s;
          const long p = (long)y * W + x;
          double     nd = D[p];
          for (int d = -1; d < 2; ++d)
          {
            if (!conn8 && d != 0)
              continue;
            const int qx = rows ? x - inc : x + d, qy = rows ? y + d : y - inc;
            if (qx < 0 || qy < 0 || qx >= W || qy >= H)
              continue;
            const long   q   = (long)qy * W + qx;
            const double cur = D[q] + (d == 0 ? e_axis : e_diag) + (lambda == 0.0 ? 0.0 : lambda * fabs(img[p] - img[q]));
            if (cur < nd)
              nd = cur;
          }
          if (nd < D[p])
          {
            D[p]    = nd;
            changed = 1;
          }
        }
      }
    }
    ++iters;
  }
  return iters;
}

/* L = sum(Wt * D_f64) at the unperturbed inputs and, for each pixel in
 * idx[0..m), with its image value moved by +eps and -eps and its soft-mask
 * value by +eps_mask and -eps_mask: out[5k + 0..3] = L(I+), L(I-), L(s+),
 * L(s-); out[5k + 4] = L (the same for every k).  The one-sided slopes tell
 * whether a step straddles a kink of the piecewise-linear map. */
void ora_fd_sides(const float* img, const float* smask, const float* Wt, int W, int H, double lambda, double nu,
                  int conn8, double eps, double eps_mask, const long* idx, int m, double* out)
{
  const long n = (long)W * H;
  if (n <= 0)
    return;
  double* I = (double*)calloc((size_t)n, sizeof(double));
  double* S = (double*)calloc((size_t)n, sizeof(double));
  double* D = (double*)calloc((size_t)n, sizeof(double));
  for (long i = 0; i < n; ++i)
  {
    I[i] = img[i];
    S[i] = smask[i];
  }
  double l0 = 0;
  ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
  for (long i = 0; i < n; ++i)
    l0 += (double)Wt[i] * D[i];
  for (int k = 0; k < m; ++k)
  {
    const long p = idx[k];
    for (int side = 0; side < 4; ++side)
    {
      double*      v  = side < 2 ? &I[p] : &S[p];
      const double v0 = *v;
      const double st = side < 2 ? eps : eps_mask;
      *v              = (side % 2 == 0) ? v0 + st : v0 - st;
      ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
      double l = 0;
      for (long i = 0; i < n; ++i)
        l += (double)Wt[i] * D[i];
      out[5 * k + side] = l;
      *v                = v0;
    }
    out[5 * k + 4] = l0;
  }
  free(I);
  free(S);
  free(D);
}

/* Central differences of L = sum(Wt * D_f64) w.r.t. image pixels and mask
 * pixels listed in idx[0..m).  out_img[k], out_mask[k] receive dL/dI, dL/ds. */
void ora_fd_grad(const float* img, const float* smask, const float* Wt, int W, int H, double lambda, double nu,
                 int conn8, double eps, double eps_mask, const long* idx, int m, double* out_img,
                 double* out_mask)
{
  const long n  = (long)W * H;
  double*    I  = (double*)malloc(sizeof(double) * (size_t)n);
  double*    S  = (double*)malloc(sizeof(double) * (size_t)n);
  double*    D  = (double*)malloc(sizeof(double) * (size_t)n);
  for (long i = 0; i < n; ++i)
  {
    I[i] = img[i];
    S[i] = smask[i];
  }
  for (int k = 0; k < m; ++k)
  {
    const long   p = idx[k];
    double       lp, lm;
    const double i0 = I[p];
    I[p]            = i0 + eps;
    ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
    lp = 0;
    for (long i = 0; i < n; ++i)
      lp += (double)Wt[i] * D[i];
    I[p] = i0 - eps;
    ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
    lm = 0;
    for (long i = 0; i < n; ++i)
      lm += (double)Wt[i] * D[i];
    I[p]       = i0;
    out_img[k] = (lp - lm) / (2 * eps);

    const double s0 = S[p];
    S[p]            = s0 + eps_mask;
    ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
    lp = 0;
    for (long i = 0; i < n; ++i)
      lp += (double)Wt[i] * D[i];
    S[p] = s0 - eps_mask;
    ora_raster_f64(I, S, D, W, H, lambda, nu, conn8);
    lm = 0;
    for (long i = 0; i < n; ++i)
      lm += (double)Wt[i] * D[i];
    S[p]        = s0;
    out_mask[k] = (lp - lm) / (2 * eps_mask);
  }
  free(I);
  free(S);
  free(D);
}
