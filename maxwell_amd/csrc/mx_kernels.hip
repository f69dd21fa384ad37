// mx_kernels.hip -- gfx950 kernels for the Bloch-periodic Nedelec eigensolver.
//
// Every kernel works on a block of nv vectors (grid.y = vector) so that one
// launch covers all vectors of an eigenvector block or of a batch of Bloch
// problems.  Threads map to grid nodes (or patches, or coarse nodes) with the
// i index fastest, so consecutive lanes of a 64-wide wavefront touch
// consecutive complex values of a grid row.  Values outside the periodic cell
// are reached through the Bloch relation u(i+nx, j) = xf u(i, j),
// u(i, j+ny) = yf u(i, j).
//
// Reference behaviour restated here (file:line in the reference):
//   curl-curl + mass stencil ............ edge2d.c:1335-1650
//   AFW vertex-patch solve .............. edge2d.c:2029-2654
//   four-colour sweep order ............. edge2d.c:3043-3480
//   edge prolongation / restriction ..... edge2d.c:5583-5705
//   node Laplacian, GS, transfers ....... node2d.c:516-1273
//   gradient / adjoint gradient ......... node2d.c:1275-1350
//   Ritz Gram entries ................... exp_bloch2d.c:992-996
//   block update ........................ exp_bloch2d.c:1006-1010
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>

#include "mxh.h"

#define MXH_DEV __device__ __forceinline__

static hipStream_t g_stream = nullptr;
static int g_device = -1;  // the device g_stream belongs to (current at mxh_init)

// Scratch of mxh_gram_static, the entry point that allocates nothing: the
// Gram kernel's partial-sum sets in a file-scope device array.  Every device
// the library runs on holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for it on
// that device (g_scratch_ptr, set by mxh_init with the device current).
#define MXH_MAX_DEVICES 64
static __device__ __align__(256) double scratch[MXH_GRAM_SCRATCH_BYTES / sizeof(double)];
static void* g_scratch_ptr[MXH_MAX_DEVICES];

// ----------------------------------------------------------------- complex
template <typename T>
struct alignas(2 * sizeof(T)) C {
  T r, i;
};

template <typename T> MXH_DEV C<T> cmk(T r, T i) { C<T> z; z.r = r; z.i = i; return z; }
template <typename T> MXH_DEV C<T> operator+(C<T> a, C<T> b) { return cmk(a.r + b.r, a.i + b.i); }
template <typename T> MXH_DEV C<T> operator-(C<T> a, C<T> b) { return cmk(a.r - b.r, a.i - b.i); }
template <typename T> MXH_DEV C<T> operator*(C<T> a, C<T> b)
{
  return cmk(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r);
}
template <typename T> MXH_DEV C<T> operator*(T s, C<T> a) { return cmk(s * a.r, s * a.i); }
template <typename T> MXH_DEV C<T> cconj(C<T> a) { return cmk(a.r, -a.i); }
// acc += conj(a) b as four fused multiply-adds
template <typename T> MXH_DEV void cmac_conj(C<T>& acc, C<T> a, C<T> b)
{
  acc.r = fma(a.r, b.r, acc.r);
  acc.r = fma(a.i, b.i, acc.r);
  acc.i = fma(a.r, b.i, acc.i);
  acc.i = fma(-a.i, b.r, acc.i);
}
template <typename T> MXH_DEV C<T> cdiv(C<T> a, C<T> b)
{
  const T d = T(1) / (b.r * b.r + b.i * b.i);
  return cmk((a.r * b.r + a.i * b.i) * d, (a.i * b.r - a.r * b.i) * d);
}
template <typename T> MXH_DEV C<T> czero() { return cmk(T(0), T(0)); }
template <typename T> MXH_DEV C<T>& operator+=(C<T>& a, C<T> b) { a.r += b.r; a.i += b.i; return a; }
template <typename T> MXH_DEV C<T>& operator-=(C<T>& a, C<T> b) { a.r -= b.r; a.i -= b.i; return a; }
template <typename T, typename S> MXH_DEV C<T> ccast(C<S> a) { return cmk(T(a.r), T(a.i)); }

// ----------------------------------------------------------------- geometry
template <typename T>
struct Geo {
  int nx, ny;
  C<T> xf, yf;

  // value at unwrapped (i, j) of a row-major nx x ny array, with Bloch phase
  MXH_DEV C<T> at(const C<T>* a, int i, int j) const
  {
    C<T> ph = cmk(T(1), T(0));
    bool w = false;
    if (i < 0) { i += nx; ph = cconj(xf); w = true; }
    else if (i >= nx) { i -= nx; ph = xf; w = true; }
    if (j < 0) { j += ny; ph = w ? ph * cconj(yf) : cconj(yf); w = true; }
    else if (j >= ny) { j -= ny; ph = w ? ph * yf : yf; w = true; }
    C<T> v = a[i + (size_t)j * nx];
    return w ? ph * v : v;
  }
  // stored value at wrapped (i, j) += conj(phase) * d  (inverse of `at`)
  MXH_DEV void add(C<T>* a, int i, int j, C<T> d) const
  {
    C<T> ph = cmk(T(1), T(0));
    bool w = false;
    if (i < 0) { i += nx; ph = xf; w = true; }
    else if (i >= nx) { i -= nx; ph = cconj(xf); w = true; }
    if (j < 0) { j += ny; ph = w ? ph * yf : yf; w = true; }
    else if (j >= ny) { j -= ny; ph = w ? ph * cconj(yf) : cconj(yf); w = true; }
    size_t k = i + (size_t)j * nx;
    a[k] += w ? ph * d : d;
  }
};

template <typename T>
MXH_DEV T eta_at(const T* eta, int nx, int ny, int i, int j)
{
  if (i < 0) i += nx; else if (i >= nx) i -= nx;
  if (j < 0) j += ny; else if (j >= ny) j -= ny;
  return eta[i + (size_t)j * nx];
}

template <typename T>
MXH_DEV Geo<T> make_geo(const mxh_level& lv, const mxh_bloch& bl, int v)
{
  Geo<T> g;
  g.nx = lv.nx;
  g.ny = lv.ny;
  if (bl.per_problem) {
    const int gv = bl.vmap ? bl.vmap[v] : v;
    const double* q = bl.per_problem + 4 * (gv / bl.vpp);
    g.xf = cmk(T(q[0]), T(q[1]));
    g.yf = cmk(T(q[2]), T(q[3]));
  } else {
    g.xf = cmk(T(bl.xr), T(bl.xi));
    g.yf = cmk(T(bl.yr), T(bl.yi));
  }
  return g;
}

// Edge vector view: x-edges at x, y-edges at x + nx*ny.
template <typename T>
struct EdgeV {
  const C<T>* x;
  const C<T>* y;
};

template <typename T>
MXH_DEV EdgeV<T> edge_view(const void* base, long long stride, int v, size_t N)
{
  EdgeV<T> e;
  e.x = (const C<T>*)base + (size_t)v * stride;
  e.y = e.x + N;
  return e;
}

// Curl of cell (i,j) in unwrapped coordinates (edge2d.c:1505-1514).
template <typename T>
MXH_DEV C<T> cell_curl(const Geo<T>& g, const EdgeV<T>& e, int i, int j)
{
  return g.at(e.x, i, j) - g.at(e.x, i, j + 1) - g.at(e.y, i, j) + g.at(e.y, i + 1, j);
}

// (alpha K + beta M) x at x-edge (i,j) and y-edge (i,j), unwrapped (i,j) allowed.
template <typename T>
MXH_DEV void apply_at(const Geo<T>& g, const EdgeV<T>& e, const T* eta, int i, int j,
                      C<T> axy, C<T> mx, C<T> my, C<T>& ox, C<T>& oy)
{
  C<T> x00 = g.at(e.x, i, j), x0m = g.at(e.x, i, j - 1), x0p = g.at(e.x, i, j + 1);
  C<T> xm0 = g.at(e.x, i - 1, j), xmp = g.at(e.x, i - 1, j + 1);
  C<T> y00 = g.at(e.y, i, j), ym0 = g.at(e.y, i - 1, j), yp0 = g.at(e.y, i + 1, j);
  C<T> y0m = g.at(e.y, i, j - 1), ypm = g.at(e.y, i + 1, j - 1);
  T e00 = eta_at(eta, g.nx, g.ny, i, j);
  T e0m = eta_at(eta, g.nx, g.ny, i, j - 1);
  T em0 = eta_at(eta, g.nx, g.ny, i - 1, j);
  C<T> c00 = x00 - x0p - y00 + yp0;  // cell (i,j)
  C<T> c0m = x0m - x00 - y0m + ypm;  // cell (i,j-1)
  C<T> cm0 = xm0 - xmp - ym0 + y00;  // cell (i-1,j)
  C<T> kx = e00 * c00 - e0m * c0m;
  C<T> ky = em0 * cm0 - e00 * c00;
  ox = axy * kx + mx * (T(4) * x00 + x0m + x0p);
  oy = axy * ky + my * (T(4) * y00 + ym0 + yp0);
}

// K x (without 1/(hx hy)) and the M stencils at x-edge (i,j) and y-edge (i,j),
// neighbours read through the Bloch wrap (frame points of the grid).
template <typename T>
MXH_DEV void stencil2(const Geo<T>& g, const C<T>* ex, const C<T>* ey, const T* eta, int i, int j, C<T>& kx,
                      C<T>& ky, C<T>& mxs, C<T>& mys, C<T>& x00, C<T>& y00)
{
  x00 = g.at(ex, i, j);
  y00 = g.at(ey, i, j);
  const C<T> x0m = g.at(ex, i, j - 1), x0p = g.at(ex, i, j + 1);
  const C<T> xm0 = g.at(ex, i - 1, j), xmp = g.at(ex, i - 1, j + 1);
  const C<T> ym0 = g.at(ey, i - 1, j), yp0 = g.at(ey, i + 1, j);
  const C<T> y0m = g.at(ey, i, j - 1), ypm = g.at(ey, i + 1, j - 1);
  const T e00 = eta_at(eta, g.nx, g.ny, i, j), e0m = eta_at(eta, g.nx, g.ny, i, j - 1), em0 = eta_at(eta, g.nx, g.ny, i - 1, j);
  const C<T> c00 = x00 - x0p - y00 + yp0;  // cell (i,j)
  const C<T> c0m = x0m - x00 - y0m + ypm;  // cell (i,j-1)
  const C<T> cm0 = xm0 - xmp - ym0 + y00;  // cell (i-1,j)
  kx = e00 * c00 - e0m * c0m;
  ky = em0 * cm0 - e00 * c00;
  mxs = T(4) * x00 + x0m + x0p;
  mys = T(4) * y00 + ym0 + yp0;
}

template <typename T>
MXH_DEV void coefs(const mxh_level& lv, const mxh_coef& cf, int v, C<T>& axy, C<T>& mx, C<T>& my)
{
  double ar = cf.ar, ai = cf.ai, br = cf.br, bi = cf.bi;
  if (cf.per_vector) {
    const double* q = cf.per_vector + 4 * v;
    ar = q[0]; ai = q[1]; br = q[2]; bi = q[3];
  }
  double s = 1.0 / (lv.hx * lv.hy), sx = lv.hy / lv.hx / 6.0, sy = lv.hx / lv.hy / 6.0;
  axy = cmk(T(ar * s), T(ai * s));
  mx = cmk(T(br * sx), T(bi * sx));
  my = cmk(T(br * sy), T(bi * sy));
}

// ----------------------------------------------------------------- reductions
MXH_DEV double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
MXH_DEV double wave_max(double v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum over a 256-thread block; result valid in thread 0.
MXH_DEV double block_sum256(double v, double* sh)
{
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  return r;
}
MXH_DEV double block_max256(double v, double* sh)
{
  v = wave_max(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  return r;
}

// out[w] = sum_b part[b * width + w] (mode 0) or the max (mode 1): one
// wavefront per output, lane l takes b = l, l + 64, ..., then a fixed
// butterfly -- a fixed order, so the result is deterministic, and 64 loads
// in flight per output instead of a serial chain.
__global__ void k_reduce_partials(const double* __restrict__ part, int nblk, int width,
                                  double* __restrict__ out, int mode)
{
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (w >= width) return;  // wave-uniform
  double s = 0.0;
  if (mode == 0) {
    for (int b = l; b < nblk; b += 64) s += part[(size_t)b * width + w];
    s = wave_sum(s);
  } else {
    for (int b = l; b < nblk; b += 64) s = fmax(s, part[(size_t)b * width + w]);
    s = wave_max(s);
  }
  if (l == 0) out[w] = s;
}

static const int TPB = 256;

static inline unsigned nblocks(long long n, int tpb = TPB) { return (unsigned)((n + tpb - 1) / tpb); }

// k_reduce_partials over `width` outputs: four outputs per 256-thread block
static inline unsigned reduce_blocks(long long width) { return nblocks(width, TPB / 64); }

// Blocks per vector of the row-mapped kernels.
static inline unsigned row_blocks(int ny) { return (unsigned)(ny < 128 ? (ny < 1 ? 1 : ny) : 128); }
// The lifting right-hand side walks columns down its band and reads two
// FP64 and two FP32 rows per band start: taller bands (16 rows at 512) pay
// that start less often (768 vectors at 512^2: 2.75 ms against 3.02 ms for
// 8-row bands, which were 10% faster than 4-row bands; short bands suit the
// defect and the lifting update better).
static inline unsigned lift_blocks(int ny) { return (unsigned)(ny < 32 ? (ny < 1 ? 1 : ny) : 32); }
// Rows of block blockIdx.x in a row-mapped kernel: a contiguous band, so the
// neighbouring rows a stencil reads were read by the same block just before.
MXH_DEV int row_first(int ny)
{
  const int rpb = (ny + (int)gridDim.x - 1) / (int)gridDim.x;
  return min(ny, (int)blockIdx.x * rpb);
}
MXH_DEV int row_end(int ny)
{
  const int rpb = (ny + (int)gridDim.x - 1) / (int)gridDim.x;
  return min(ny, ((int)blockIdx.x + 1) * rpb);
}

// The frame of an nx x ny grid (rows 0 and ny - 1, columns 0 and nx - 1):
// the points whose stencils wrap.  The row-mapped kernels run the interior
// with direct indexing and the frame in a second, small launch with the
// wrapped form, so the interior kernel carries no registers for the wrap.
__host__ __device__ inline long long frame_points(int nx, int ny)
{
  return ny > 2 ? 2LL * nx + 2LL * (ny - 2) : (long long)nx * ny;
}
MXH_DEV void frame_point(int k, int nx, int ny, int& i, int& j)
{
  if (ny <= 2) { i = k % nx; j = k / nx; }
  else if (k < nx) { i = k; j = 0; }
  else if (k < 2 * nx) { i = k - nx; j = ny - 1; }
  else { const int q = k - 2 * nx; i = (q & 1) ? nx - 1 : 0; j = 1 + (q >> 1); }
}
static inline unsigned frame_blocks(const mxh_level* lv)
{
  const long long b = (frame_points(lv->nx, lv->ny) + 255) / 256;
  return (unsigned)(b > 8 ? 8 : (b < 1 ? 1 : b));
}

// Number of blocks used for grid-stride reductions (partials per vector).
static inline unsigned red_blocks(long long n)
{
  long long b = (n + TPB - 1) / TPB;
  if (b > 512) b = 512;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ----------------------------------------------------------------- addeval
template <typename T>
__global__ __launch_bounds__(256) void k_addeval(mxh_level lv, mxh_bloch bl, const C<T>* X, long long xs,
                                                 C<T>* Y, long long ys, const T* eta, mxh_coef cf,
                                                 int accumulate)
{
  const int v = blockIdx.y;
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  int i = (int)(idx % lv.nx), j = (int)(idx / lv.nx);
  Geo<T> g = make_geo<T>(lv, bl, v);
  C<T> axy, mx, my;
  coefs<T>(lv, cf, v, axy, mx, my);
  EdgeV<T> e = edge_view<T>(X, xs, v, N);
  C<T> ox, oy;
  apply_at(g, e, eta, i, j, axy, mx, my, ox, oy);
  C<T>* y = Y + (size_t)v * ys;
  if (accumulate) {
    y[idx] += ox;
    y[N + idx] += oy;
  } else {
    y[idx] = ox;
    y[N + idx] = oy;
  }
}

// PINVIT defect r = (K + (1 - lam) M) e in FP64, stored in precision R, plus
// per-block partial sums of |r|^2 and |M e|^2 over the block's band of rows.
template <typename R>
MXH_DEV void defect_point(const Geo<double>& g, const C<double>* ex, const C<double>* ey, const double* eta, int i,
                          int j, double s, double sx, double sy, double be, C<R>* r, size_t N, double& rr,
                          double& mm)
{
  C<double> kx, ky, mxs, mys, x00, y00;
  stencil2(g, ex, ey, eta, i, j, kx, ky, mxs, mys, x00, y00);
  const C<double> mxv = sx * mxs, myv = sy * mys;
  const C<double> ox = s * kx + be * mxv, oy = s * ky + be * myv;
  const size_t k = i + (size_t)j * g.nx;
  r[k] = ccast<R>(ox);
  r[N + k] = ccast<R>(oy);
  rr += ox.r * ox.r + ox.i * ox.i + oy.r * oy.r + oy.i * oy.i;
  mm += mxv.r * mxv.r + mxv.i * mxv.i + myv.r * myv.r + myv.i * myv.i;
}

template <typename R, bool FRAME>
__global__ __launch_bounds__(256) void k_defect(mxh_level lv, mxh_bloch bl, const C<double>* E, long long es,
                                                const int* emap, C<R>* Rv, long long rs, const double* eta,
                                                const double* lam, double* part, int slot0)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  const int gv = emap ? emap[v] : v;
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const Geo<double> g = make_geo<double>(lv, bl, v);
  const double s = 1.0 / (lv.hx * lv.hy), sx = lv.hy / lv.hx / 6.0, sy = lv.hx / lv.hy / 6.0;
  const double be = 1.0 - lam[gv];
  const EdgeV<double> e = edge_view<double>(E, es, gv, N);
  C<R>* r = Rv + (size_t)v * rs;
  double rr = 0.0, mm = 0.0;
  if (FRAME) {
    const int nf = (int)frame_points(nx, ny);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) {
      int i, j;
      frame_point(k, nx, ny, i, j);
      defect_point<R>(g, e.x, e.y, eta, i, j, s, sx, sy, be, r, N, rr, mm);
    }
  } else {
    // each thread walks its columns down the band, keeping the values the
    // next row's stencil shares with this one (half the loads per point)
    const int j0 = max(1, row_first(ny)), j1 = min(ny - 1, row_end(ny));
    const C<double>* ex = e.x;
    const C<double>* ey = e.y;
    if (j0 < j1)
      for (int i = 1 + (int)threadIdx.x; i < nx - 1; i += blockDim.x) {
        size_t k = i + (size_t)j0 * nx;
        C<double> x0m = ex[k - nx], x00 = ex[k], xm0 = ex[k - 1];
        C<double> y0m = ey[k - nx], ypm = ey[k - nx + 1];
        double e0m = eta[k - nx];
        for (int j = j0; j < j1; j++, k += nx) {
          const C<double> x0p = ex[k + nx], xmp = ex[k + nx - 1];
          const C<double> y00 = ey[k], ym0 = ey[k - 1], yp0 = ey[k + 1];
          const double e00 = eta[k], em0 = eta[k - 1];
          const C<double> c00 = x00 - x0p - y00 + yp0;  // cell (i,j)
          const C<double> c0m = x0m - x00 - y0m + ypm;  // cell (i,j-1)
          const C<double> cm0 = xm0 - xmp - ym0 + y00;  // cell (i-1,j)
          const C<double> kx = e00 * c00 - e0m * c0m, ky = em0 * cm0 - e00 * c00;
          const C<double> mxv = sx * (4.0 * x00 + x0m + x0p), myv = sy * (4.0 * y00 + ym0 + yp0);
          const C<double> ox = s * kx + be * mxv, oy = s * ky + be * myv;
          r[k] = ccast<R>(ox);
          r[N + k] = ccast<R>(oy);
          rr += ox.r * ox.r + ox.i * ox.i + oy.r * oy.r + oy.i * oy.i;
          mm += mxv.r * mxv.r + mxv.i * mxv.i + myv.r * myv.r + myv.i * myv.i;
          x0m = x00;
          x00 = x0p;
          xm0 = xmp;
          y0m = y00;
          ypm = yp0;
          e0m = e00;
        }
      }
  }
  double a = block_sum256(rr, sh);
  double b = block_sum256(mm, sh);
  if (threadIdx.x == 0) {
    size_t w = (size_t)(slot0 + blockIdx.x) * (2 * gridDim.y) + 2 * v;
    part[w] = a;
    part[w + 1] = b;
  }
}

// ----------------------------------------------------------------- AFW local solve
// Exact solve of the 4x4 patch system
//   [[a0, 0, -eLL, eLU], [0, a1, eRL, -eRU], [-eLL, eRL, a2, 0], [eLU, -eRU, 0, a3]] z = r
// by eliminating z0, z1 (diagonal block) and solving the 2x2 Schur complement
// for z2, z3.  a0..a3 carry the reference's doubled mass diagonal.
// Reciprocals of the local solve's pivots, which only feed the multigrid
// preconditioner: in FP64 the hardware estimate refined by two Newton steps
// (within an ulp or two of IEEE division, at half its instructions); in FP32
// the hardware reciprocal (1 ulp).  The pivots are positive and of order
// 1 / (hx hy).
MXH_DEV double recip(double a)
{
  double r = __builtin_amdgcn_rcp(a);
  r = fma(fma(-a, r, 1.0), r, r);
  return fma(fma(-a, r, 1.0), r, r);
}
MXH_DEV float recip(float a) { return __builtin_amdgcn_rcpf(a); }

template <typename T>
MXH_DEV void afw_local_solve(T eLL, T eRL, T eLU, T eRU, T mx, T my, C<T> r0, C<T> r1, C<T> r2, C<T> r3, C<T>& z0,
                             C<T>& z1, C<T>& z2, C<T>& z3)
{
  const T a0 = eLL + eLU + T(8) * mx, a1 = eRL + eRU + T(8) * mx;
  const T a2 = eLL + eRL + T(8) * my, a3 = eLU + eRU + T(8) * my;
  const T p = recip(a0), q = recip(a1);
  const T c20 = eLL * p, c21 = eRL * q, c30 = eLU * p, c31 = eRU * q;
  const T s22 = a2 - eLL * c20 - eRL * c21;
  const T s33 = a3 - eLU * c30 - eRU * c31;
  const T s23 = eLL * c30 + eRL * c31;
  const T idet = recip(s22 * s33 - s23 * s23);
  const C<T> s2 = r2 + c20 * r0 - c21 * r1;
  const C<T> s3 = r3 - c30 * r0 + c31 * r1;
  z2 = idet * (s33 * s2 - s23 * s3);
  z3 = idet * (s22 * s3 - s23 * s2);
  z0 = p * r0 + c20 * z2 - c30 * z3;
  z1 = q * r1 - c21 * z2 + c31 * z3;
}

// ----------------------------------------------------------------- AFW smoother
// Exact solve on the star of node (i,j) with the reference's patch matrix
// alpha K_patch + 2 beta M_patch (edge2d.c:2038-2062) and the exact residual
// b - (alpha K + beta M) x.  Unknowns: u0 = x(i-1,j), u1 = x(i,j),
// u2 = y(i,j-1), u3 = y(i,j).
template <typename T>
__global__ __launch_bounds__(256) void k_afw_phase(mxh_level lv, mxh_bloch bl, const C<T>* B, long long bs,
                                                   C<T>* X, long long xs, const T* eta, double alpha,
                                                   double beta, int ci, int cj)
{
  const int v = blockIdx.y;
  const int hx2 = lv.nx >> 1;
  const size_t np = (size_t)hx2 * (lv.ny >> 1);
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int i = 2 * (int)(p % hx2) + ci, j = 2 * (int)(p / hx2) + cj;
  const size_t N = (size_t)lv.nx * lv.ny;
  Geo<T> g = make_geo<T>(lv, bl, v);
  C<T>* xx = X + (size_t)v * xs;
  C<T>* xy = xx + N;
  const C<T>* bx = B + (size_t)v * bs;
  const C<T>* by = bx + N;
  const T axy = T(alpha / (lv.hx * lv.hy));
  const T mx = T(beta * lv.hy / lv.hx / 6.0), my = T(beta * lv.hx / lv.hy / 6.0);

  C<T> xLm = g.at(xx, i - 1, j - 1), xL0 = g.at(xx, i - 1, j), xLp = g.at(xx, i - 1, j + 1);
  C<T> xRm = g.at(xx, i, j - 1), xR0 = g.at(xx, i, j), xRp = g.at(xx, i, j + 1);
  C<T> yLm = g.at(xy, i - 1, j - 1), yCm = g.at(xy, i, j - 1), yRm = g.at(xy, i + 1, j - 1);
  C<T> yL0 = g.at(xy, i - 1, j), yC0 = g.at(xy, i, j), yR0 = g.at(xy, i + 1, j);

  const T eLL = axy * eta_at(eta, lv.nx, lv.ny, i - 1, j - 1);
  const T eRL = axy * eta_at(eta, lv.nx, lv.ny, i, j - 1);
  const T eLU = axy * eta_at(eta, lv.nx, lv.ny, i - 1, j);
  const T eRU = axy * eta_at(eta, lv.nx, lv.ny, i, j);

  C<T> qLL = eLL * (xLm - xL0 - yLm + yCm);
  C<T> qRL = eRL * (xRm - xR0 - yCm + yRm);
  C<T> qLU = eLU * (xL0 - xLp - yL0 + yC0);
  C<T> qRU = eRU * (xR0 - xRp - yC0 + yR0);

  C<T> r0 = g.at(bx, i - 1, j) - (qLU - qLL) - mx * (T(4) * xL0 + xLm + xLp);
  C<T> r1 = g.at(bx, i, j) - (qRU - qRL) - mx * (T(4) * xR0 + xRm + xRp);
  C<T> r2 = g.at(by, i, j - 1) - (qLL - qRL) - my * (T(4) * yCm + yLm + yRm);
  C<T> r3 = g.at(by, i, j) - (qLU - qRU) - my * (T(4) * yC0 + yL0 + yR0);

  C<T> z0, z1, z2, z3;
  afw_local_solve(eLL, eRL, eLU, eRU, mx, my, r0, r1, r2, r3, z0, z1, z2, z3);

  g.add(xx, i - 1, j, z0);
  xx[i + (size_t)j * lv.nx] += z1;
  g.add(xy, i, j - 1, z2);
  xy[i + (size_t)j * lv.nx] += z3;
}

// ----------------------------------------------------------------- residual + restriction
template <typename T>
MXH_DEV void resid_at(const Geo<T>& g, const EdgeV<T>& b, const EdgeV<T>& x, const T* eta, int i, int j,
                      C<T> axy, C<T> mx, C<T> my, C<T>& rx, C<T>& ry)
{
  C<T> ax, ay;
  apply_at(g, x, eta, i, j, axy, mx, my, ax, ay);
  rx = g.at(b.x, i, j) - ax;
  ry = g.at(b.y, i, j) - ay;
}

template <typename T>
__global__ __launch_bounds__(256) void k_resid_restrict(mxh_level lf, mxh_bloch bl, const C<T>* BF, long long bfs,
                                                        const C<T>* XF, long long xfs, const T* eta,
                                                        C<T>* BC, long long bcs, C<T>* XC, long long xcs)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny, NF = (size_t)lf.nx * lf.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> g = make_geo<T>(lf, bl, v);
  EdgeV<T> b = edge_view<T>(BF, bfs, v, NF), x = edge_view<T>(XF, xfs, v, NF);
  const double s = 1.0 / (lf.hx * lf.hy);
  const C<T> axy = cmk(T(s), T(0));
  const C<T> mx = cmk(T(lf.hy / lf.hx / 6.0), T(0)), my = cmk(T(lf.hx / lf.hy / 6.0), T(0));
  C<T> rx, ry, sx = czero<T>(), sy = czero<T>();
  const int i0 = 2 * I, j0 = 2 * J;
  // x-edges: (i0,j0),(i0+1,j0) weight 1/2; rows j0-1 and j0+1 weight 1/4
  resid_at(g, b, x, eta, i0, j0, axy, mx, my, rx, ry);
  sx += T(0.5) * rx;
  sy += T(0.5) * ry;  // y(i0, j0)
  resid_at(g, b, x, eta, i0 + 1, j0, axy, mx, my, rx, ry);
  sx += T(0.5) * rx;
  sy += T(0.25) * ry;  // y(i0+1, j0)
  resid_at(g, b, x, eta, i0, j0 + 1, axy, mx, my, rx, ry);
  sx += T(0.25) * rx;
  sy += T(0.5) * ry;  // y(i0, j0+1)
  resid_at(g, b, x, eta, i0 + 1, j0 + 1, axy, mx, my, rx, ry);
  sx += T(0.25) * rx;
  sy += T(0.25) * ry;  // y(i0+1, j0+1)
  resid_at(g, b, x, eta, i0, j0 - 1, axy, mx, my, rx, ry);
  sx += T(0.25) * rx;
  resid_at(g, b, x, eta, i0 + 1, j0 - 1, axy, mx, my, rx, ry);
  sx += T(0.25) * rx;
  resid_at(g, b, x, eta, i0 - 1, j0, axy, mx, my, rx, ry);
  sy += T(0.25) * ry;
  resid_at(g, b, x, eta, i0 - 1, j0 + 1, axy, mx, my, rx, ry);
  sy += T(0.25) * ry;
  C<T>* bc = BC + (size_t)v * bcs;
  C<T>* xc = XC + (size_t)v * xcs;
  bc[idx] = sx;
  bc[NC + idx] = sy;
  xc[idx] = czero<T>();
  xc[NC + idx] = czero<T>();
}

// ----------------------------------------------------------------- transfers
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(mxh_level lf, mxh_bloch bl, const C<T>* CV, long long cs,
                                                 C<T>* FV, long long fs, T alpha)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny, NF = (size_t)lf.nx * lf.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> gc = make_geo<T>(lf, bl, v);
  gc.nx = cnx;
  gc.ny = cny;
  const C<T>* cx = CV + (size_t)v * cs;
  const C<T>* cy = cx + NC;
  C<T>* fx = FV + (size_t)v * fs;
  C<T>* fy = fx + NF;
  C<T> a0 = cx[idx], a1 = gc.at(cx, I, J + 1);
  C<T> b0 = cy[idx], b1 = gc.at(cy, I + 1, J);
  const size_t r0 = (size_t)(2 * J) * lf.nx, r1 = r0 + lf.nx;
  const int i0 = 2 * I;
  C<T> h = (T(0.5) * alpha) * a0, q = (T(0.25) * alpha) * (a0 + a1);
  fx[r0 + i0] += h;
  fx[r0 + i0 + 1] += h;
  fx[r1 + i0] += q;
  fx[r1 + i0 + 1] += q;
  h = (T(0.5) * alpha) * b0;
  q = (T(0.25) * alpha) * (b0 + b1);
  fy[r0 + i0] += h;
  fy[r1 + i0] += h;
  fy[r0 + i0 + 1] += q;
  fy[r1 + i0 + 1] += q;
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict(mxh_level lf, mxh_bloch bl, const C<T>* FV, long long fs,
                                                  C<T>* CV, long long cs, T alpha)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny, NF = (size_t)lf.nx * lf.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> g = make_geo<T>(lf, bl, v);
  const C<T>* fx = FV + (size_t)v * fs;
  const C<T>* fy = fx + NF;
  const int i0 = 2 * I, j0 = 2 * J;
  C<T> sx = T(0.5) * (g.at(fx, i0, j0) + g.at(fx, i0 + 1, j0)) +
            T(0.25) * (g.at(fx, i0, j0 - 1) + g.at(fx, i0 + 1, j0 - 1) + g.at(fx, i0, j0 + 1) +
                       g.at(fx, i0 + 1, j0 + 1));
  C<T> sy = T(0.5) * (g.at(fy, i0, j0) + g.at(fy, i0, j0 + 1)) +
            T(0.25) * (g.at(fy, i0 - 1, j0) + g.at(fy, i0 - 1, j0 + 1) + g.at(fy, i0 + 1, j0) +
                       g.at(fy, i0 + 1, j0 + 1));
  C<T>* cx = CV + (size_t)v * cs;
  cx[idx] += alpha * sx;
  cx[NC + idx] += alpha * sy;
}

// ----------------------------------------------------------------- coarse dense solve
// x = L^{-H} L^{-1} b with L lower, column-major n x n (double complex).
// One block per vector, n <= blockDim.x; column-oriented substitution.
// Per-k coarse set-up on the device (exp_bloch2d.c:799-809): one workgroup per
// problem assembles the n0 x n0 coarse matrices from cell matrices -- the edge
// K + M with the reference's phases (edge2d.c:5870-6110) and the Q1 node
// Laplacian with the Gamma fill of node2d.c:1454-1458 -- in LDS, factors them
// (lower Cholesky, column by column) and writes the factors column-major.
// Each entry sums its cell contributions in the order of a serial assembly
// (cells row-major, element rows, element columns).  info[p] = 0, or the
// column (1-based) of a non-positive pivot.
MXH_DEV C<double> coarse_wrap(int nx, int ny, C<double> xf, C<double> yf, int& i, int& j)
{
  C<double> ph = cmk(1.0, 0.0);
  if (i >= nx) { i -= nx; ph = ph * xf; }
  if (j >= ny) { j -= ny; ph = ph * yf; }
  return ph;
}

MXH_DEV void coarse_cholesky(C<double>* a, int n, double* dsh, int* info)
{
  for (int c = 0; c < n; c++) {
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = a[c + c * n].r;
      for (int k = 0; k < c; k++) {
        const C<double> v = a[c + k * n];
        s -= v.r * v.r + v.i * v.i;
      }
      if (!(s > 0.0) && *info == 0) *info = c + 1;
      const double d = s > 0.0 ? sqrt(s) : 1.0;
      dsh[0] = d;
      a[c + c * n] = cmk(d, 0.0);
    }
    __syncthreads();
    const double d = dsh[0];
    for (int r = c + 1 + (int)threadIdx.x; r < n; r += blockDim.x) {
      C<double> t = a[r + c * n];
      for (int k = 0; k < c; k++) t -= a[r + k * n] * cconj(a[c + k * n]);
      a[r + c * n] = cmk(t.r / d, t.i / d);
    }
    for (int r = threadIdx.x; r < c; r += blockDim.x) a[r + c * n] = czero<double>();
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_coarse_setup(int n0, const double* bloch, const double* eta, double* Ae,
                                                      double* An, int* info)
{
  extern __shared__ double smem[];
  const int p = blockIdx.x, nx = n0, ny = n0, N = n0 * n0, ne = 2 * N;
  C<double>* a = (C<double>*)smem;
  C<double>* b = a + ne * ne;
  double* dsh = (double*)(b + N * N);
  __shared__ int inf;
  if (threadIdx.x == 0) inf = 0;
  const double* q = bloch + 4 * p;
  const C<double> xf = cmk(q[0], q[1]), yf = cmk(q[2], q[3]);
  const double hx = 1.0 / n0, hy = 1.0 / n0;
  const double axy = 1.0 / hx / hy, mx = hy / hx / 6.0, my = hx / hy / 6.0;
  // edge matrix K + M
  for (int e = threadIdx.x; e < ne * ne; e += blockDim.x) {
    const int r = e % ne, c = e / ne;
    C<double> acc = czero<double>();
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int gi[4];
        C<double> ph[4];
        const int di[4] = {0, 0, 0, 1}, dj[4] = {0, 1, 0, 0}, off[4] = {0, 0, N, N};
        bool hit = false;
        for (int k = 0; k < 4; k++) {
          int ii = i + di[k], jj = j + dj[k];
          ph[k] = coarse_wrap(nx, ny, xf, yf, ii, jj);
          gi[k] = off[k] + ii + jj * nx;
          hit |= gi[k] == r;
        }
        if (!hit) continue;
        const double sgn[4] = {1.0, -1.0, -1.0, 1.0};
        const double et = eta[i + j * nx];
        for (int kp = 0; kp < 4; kp++) {
          if (gi[kp] != r) continue;
          for (int kq = 0; kq < 4; kq++) {
            if (gi[kq] != c) continue;
            double ke = axy * et * sgn[kp] * sgn[kq];
            if (kp < 2 && kq < 2) ke += (kp == kq ? 2.0 : 1.0) * mx;
            if (kp >= 2 && kq >= 2) ke += (kp == kq ? 2.0 : 1.0) * my;
            acc += (cconj(ph[kp]) * cmk(ke, 0.0)) * ph[kq];
          }
        }
      }
    a[r + c * ne] = acc;
  }
  // node matrix: Q1 Laplacian, Gamma fill
  const bool gamma = hypot(xf.r - 1.0, xf.i) < 1e-8 && hypot(yf.r - 1.0, yf.i) < 1e-8;
  const double fill = gamma ? 0.1 / ((double)N * N) : 0.0;
  const double rx = hy / hx, ry = hx / hy;
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int r = e % N, c = e / N;
    C<double> acc = cmk(fill, 0.0);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int gi[4];
        C<double> ph[4];
        for (int k = 0; k < 4; k++) {
          int ii = i + (k & 1), jj = j + (k >> 1);
          ph[k] = coarse_wrap(nx, ny, xf, yf, ii, jj);
          gi[k] = ii + jj * nx;
        }
        for (int kp = 0; kp < 4; kp++) {
          if (gi[kp] != r) continue;
          for (int kq = 0; kq < 4; kq++) {
            if (gi[kq] != c) continue;
            const int px = kp & 1, py = kp >> 1, qx = kq & 1, qy = kq >> 1;
            const double S0 = px == qx ? 1.0 : -1.0, S1 = py == qy ? 1.0 : -1.0;
            const double T0 = px == qx ? 1.0 / 3.0 : 1.0 / 6.0, T1 = py == qy ? 1.0 / 3.0 : 1.0 / 6.0;
            const double ke = rx * S0 * T1 + ry * T0 * S1;
            acc += (cconj(ph[kp]) * cmk(ke, 0.0)) * ph[kq];
          }
        }
      }
    b[r + c * N] = acc;
  }
  __syncthreads();
  coarse_cholesky(a, ne, dsh, &inf);
  coarse_cholesky(b, N, dsh, &inf);
  C<double>* oa = (C<double>*)Ae + (size_t)p * ne * ne;
  C<double>* ob = (C<double>*)An + (size_t)p * N * N;
  for (int e = threadIdx.x; e < ne * ne; e += blockDim.x) oa[e] = a[e];
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) ob[e] = b[e];
  if (threadIdx.x == 0) info[p] = inf;
}

template <typename T>
__global__ __launch_bounds__(256) void k_coarse_solve(int n, const double* Lm, const double* Lper, int vpp,
                                                      const int* vmap, const C<T>* B, long long bs, C<T>* X,
                                                      long long xs)
{
  extern __shared__ double smem[];
  C<double>* z = (C<double>*)smem;
  const int v = blockIdx.x;
  const int gv = vmap ? vmap[v] : v;
  const C<double>* L = (const C<double>*)(Lper ? Lper + (size_t)2 * n * n * (gv / vpp) : Lm);
  const int t = threadIdx.x;
  if (t < n) z[t] = ccast<double>(B[(size_t)v * bs + t]);
  __syncthreads();
  for (int c = 0; c < n; c++) {
    if (t == c) {
      C<double> d = L[c + (size_t)c * n];
      double inv = 1.0 / d.r;  // real positive diagonal
      z[c] = inv * z[c];
    }
    __syncthreads();
    if (t > c && t < n) z[t] -= L[t + (size_t)c * n] * z[c];
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; c--) {
    if (t == c) {
      double inv = 1.0 / L[c + (size_t)c * n].r;
      z[c] = inv * z[c];
    }
    __syncthreads();
    if (t < c) z[t] -= cconj(L[c + (size_t)t * n]) * z[c];
    __syncthreads();
  }
  if (t < n) X[(size_t)v * xs + t] = ccast<T>(z[t]);
}

// ----------------------------------------------------------------- Gram products
// For each block j (problem q = pmap[j]) of m FP64 vectors, the lower
// triangles of gk[a,b] = sum_cells eta conj(c_a) c_b and gm[a,b] = u_a^H M u_b,
// c the cell curl.  The mass matrix couples an x-edge with its neighbours in
// j by the rows mx (1, 4, 1) (a y-edge with its neighbours in i by my (1, 4,
// 1)); with the Bloch shift S (x(j+1), unitary) that circulant factors as
//   4 + S + S^H = 6 (a + b S)^H (a + b S),  a, b = (1 +- 1/sqrt 3) / 2
// (a^2 + b^2 = 2/3, ab = 1/6), so x^H M y = ux^H uy + (y-edge analogue) with
// ux = sqrt(6 mx) (a x(i,j) + b x(i,j+1)): every cell contributes three
// features of its own edges and their upper and right neighbours.  A tile of
// cells is staged in LDS for all m vectors; the m (m+1)/2 pairs times cell
// slices are spread over the 256 threads, slices are summed in LDS at the end.
#define GRAM_TILE_MAX 64
#define GRAM_MMAX 48
#define GRAM_ITEMS 4 /* staged (cell, vector) items per thread and segment */

// 16-byte slots per staged cell of mp vectors: three features each, then one
// or two spare so that the count is odd
__host__ __device__ inline int gram_cell_slots(int mp) { return 3 * mp + ((3 * mp) & 1 ? 2 : 1); }

MXH_DEV void pair_ab(int p, int& a, int& b)
{
  int r = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= p) r++;
  while (r * (r + 1) / 2 > p) r--;
  a = r;
  b = p - r * (r + 1) / 2;
}

// Gram products of each problem's block, register-blocked: thread (group g,
// block pair (A, B)) accumulates the GB x GB pairs a = A GB + ka, b = B GB + kb
// (A >= B) over the cells t = g, g + ngrp, ... of each staged row segment.
// Per cell and vector three features are staged in LDS, the curl c and the
// mass factors ux, uy above (from x(i,j), x(i,j+1), y(i,j), y(i+1,j) only),
// so a pair costs three complex multiply-adds per cell:
//   K_ab += conj(c_a) eta c_b,   M_ab += conj(ux_a) ux_b + conj(uy_a) uy_b.
// A workgroup covers a band of rows of one problem; bands are summed by
// k_reduce_partials in a fixed order.  The segment width T and the item
// count per thread are chosen by mxh_gram; the next segment's loads stay in
// flight in registers while the current one is summed (two waves per SIMD).
// With a column list (cmap: per problem q, cmap[q (m+1)] = mq columns, then
// their indices in the block, increasing) a problem's sums run over those mq
// columns only and land at the pair indices of the list (pair a (a+1)/2 + b
// of list positions a >= b); the rest of its output is left unwritten.
#define GB 3

__global__ __launch_bounds__(256) void k_gram(mxh_level lv, mxh_bloch bl, int m, const int* pmap, const int* cmap,
                                              const C<double>* U0, long long us, const double* eta, double* part,
                                              int T, int rpb)
{
  extern __shared__ double smem[];
  const int jb = blockIdx.y;
  const int prob = pmap ? pmap[jb] : jb;
  const C<double>* U = U0 + (size_t)prob * m * us;
  const int* cl = cmap ? cmap + (size_t)prob * (m + 1) : nullptr;
  const int mq = cl ? cl[0] : m;  // columns summed (block-uniform)
  if (mq < 1 || mq > m) return;   // (the whole block: before any barrier)
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const int nbk = (mq + GB - 1) / GB, nbp = nbk * (nbk + 1) / 2;
  // per cell the features of mp = nbk GB vectors (zero beyond m), then one or
  // two spare complex: an odd number of 16-byte slots between cells keeps the
  // cells a wavefront stages (and the groups read) on different banks
  const int mp = nbk * GB, cs = gram_cell_slots(mp);
  C<double>* sf = (C<double>*)smem;                    // [T][cs]
  double* se = (double*)(sf + (size_t)T * cs);         // [T]
  const Geo<double> g = make_geo<double>(lv, bl, prob * m);
  // the mass factors' weights: sqrt(6 mx) a, sqrt(6 mx) b (x-edges), sqrt(6 my) a, sqrt(6 my) b
  const double fa_ = 0.5 * (1.0 + 1.0 / sqrt(3.0)), fb_ = 0.5 * (1.0 - 1.0 / sqrt(3.0));
  const double sxm = sqrt(lv.hy / lv.hx), sym = sqrt(lv.hx / lv.hy);
  const double xa = sxm * fa_, xb = sxm * fb_, ya = sym * fa_, yb = sym * fb_;
  const int ngrp = (int)blockDim.x / nbp;
  const int tid = threadIdx.x, grp = tid / nbp, bp = tid % nbp;
  int A, Bq;
  pair_ab(bp, A, Bq);
  const bool act = grp < ngrp;
  C<double> K[GB][GB], Mm[GB][GB];
#pragma unroll
  for (int x = 0; x < GB; x++)
#pragma unroll
    for (int y = 0; y < GB; y++) K[x][y] = Mm[x][y] = czero<double>();
  const int j0 = blockIdx.x * rpb, j1 = j0 + rpb < ny ? j0 + rpb : ny;
  // items w = tid + r blockDim (r < GRAM_ITEMS, mp T <= GRAM_ITEMS blockDim) of
  // a full segment: vector a = w / T, cell t = w mod T (consecutive threads on
  // consecutive cells); a >= mp marks no item
  int pa[GRAM_ITEMS], pt[GRAM_ITEMS];
#pragma unroll
  for (int r = 0; r < GRAM_ITEMS; r++) {
    const int w = tid + r * (int)blockDim.x;
    pa[r] = w / T;
    pt[r] = w - pa[r] * T;
  }
  // segments sg = (row j, columns i0 .. i0 + T - 1) of the band, column of
  // segments by column of segments: rows j and j + 1 of one column follow each
  // other, so a thread keeps the x-edge value of row j + 1 it loaded for row j
  // (the curl and the mass factor read x at j and j + 1) and loads one new x
  // row per segment instead of two.  The loads of segment sg + 1 are issued
  // before segment sg is summed, so their latency hides behind the products.
  const int nrow = j1 - j0, nseg = (nx + T - 1) / T, sEnd = nrow * nseg;
  int ia[GRAM_ITEMS], it[GRAM_ITEMS];
  C<double> v[GRAM_ITEMS][4];  // x(i,j), y(i,j), x(i,j+1), y(i+1,j), raw
  // eta of the cell of item 0: the one item of vector 0 a thread can hold
  // (items r >= 1 have w >= blockDim > T cells)
  double ev0 = 0.0;
  // all loads of the thread's items of segment sg (clamped to a real vector,
  // so every address is valid); with next_row, sg is the row after the
  // segment whose raw values v holds, in the same column
  auto load = [&](int sg, bool next_row) {
    const int j = j0 + sg % nrow, i0 = (sg / nrow) * T;
    const int cnt = nx - i0 < T ? nx - i0 : T;
    // the next row; across the periodic boundary with the Bloch factor
    const long long offp = j + 1 < ny ? nx : -(long long)(ny - 1) * nx;
#pragma unroll
    for (int r = 0; r < GRAM_ITEMS; r++) {
      ia[r] = pa[r];
      it[r] = pt[r];
      if (cnt != T) {
        const int w = tid + r * (int)blockDim.x;
        ia[r] = w / cnt;
        it[r] = w - ia[r] * cnt;
      }
      const int a = ia[r] < mq ? ia[r] : mq - 1, i = i0 + it[r];
      const C<double>* ux = U + (size_t)(cl ? cl[1 + a] : a) * us;
      const C<double>* uy = ux + N;
      const long long k = i + (long long)j * nx;
      if (next_row) {
        // x(i, j) was this item's x(i, j + 1) one row up
        v[r][0] = v[r][2];
      } else {
        v[r][0] = ux[k];
      }
      v[r][1] = uy[k];
      v[r][2] = ux[k + offp];
      v[r][3] = uy[i + 1 < nx ? k + 1 : k + 1 - nx];
      if (r == 0) ev0 = eta[k];
    }
  };
  if (sEnd > 0) load(0, false);
  for (int sg = 0; sg < sEnd; sg++) {
    const int j = j0 + sg % nrow, i0 = (sg / nrow) * T;
    const int cnt = nx - i0 < T ? nx - i0 : T;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < GRAM_ITEMS; r++) {
      const int a = ia[r], t = it[r], i = i0 + t;
      if (a >= mp) continue;
      C<double>* f = sf + (size_t)t * cs + a * 3;
      if (a >= mq) {
        f[0] = f[1] = f[2] = czero<double>();
        continue;
      }
      const C<double> x0 = v[r][0], y0 = v[r][1];
      C<double> xp = v[r][2], yp = v[r][3];
      if (j + 1 == ny) xp = g.yf * xp;
      if (i + 1 == nx) yp = g.xf * yp;
      f[0] = x0 - xp - y0 + yp;
      f[1] = xa * x0 + xb * xp;
      f[2] = ya * y0 + yb * yp;
      if (r == 0 && a == 0) se[t] = ev0;
    }
    __syncthreads();
    if (sg + 1 < sEnd) load(sg + 1, (sg + 1) % nrow != 0);
    if (act) {
      for (int t = grp; t < cnt; t += ngrp) {
        const C<double>* f = sf + (size_t)t * cs;
        const double et = se[t];
        C<double> fa[GB][3];
#pragma unroll
        for (int ka = 0; ka < GB; ka++) {
          const int a = A * GB + ka;
#pragma unroll
          for (int q = 0; q < 3; q++) fa[ka][q] = f[a * 3 + q];
        }
#pragma unroll
        for (int kb = 0; kb < GB; kb++) {
          const int b = Bq * GB + kb;
          const C<double>* fb = f + b * 3;
          const C<double> ec = et * fb[0], wx = fb[1], wy = fb[2];
#pragma unroll
          for (int ka = 0; ka < GB; ka++) {
            cmac_conj(K[ka][kb], fa[ka][0], ec);
            cmac_conj(Mm[ka][kb], fa[ka][1], wx);
            cmac_conj(Mm[ka][kb], fa[ka][2], wy);
          }
        }
      }
    }
  }
  // sum the groups through LDS, one pair of the block at a time
  const int np = m * (m + 1) / 2;
  double* out = part + ((size_t)blockIdx.x * gridDim.y + jb) * 4 * np;
  double* red = smem;  // [blockDim.x][4]
#pragma unroll
  for (int ka = 0; ka < GB; ka++)
#pragma unroll
    for (int kb = 0; kb < GB; kb++) {
      __syncthreads();
      red[4 * tid] = K[ka][kb].r;
      red[4 * tid + 1] = K[ka][kb].i;
      red[4 * tid + 2] = Mm[ka][kb].r;
      red[4 * tid + 3] = Mm[ka][kb].i;
      __syncthreads();
      if (tid < nbp) {
        int Ae, Be;
        pair_ab(tid, Ae, Be);
        const int a = Ae * GB + ka, b = Be * GB + kb;
        if (a < mq && b < mq && a >= b) {
          double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
          for (int gq = 0; gq < ngrp; gq++) {
            const double* r = red + 4 * (gq * nbp + tid);
            v0 += r[0];
            v1 += r[1];
            v2 += r[2];
            v3 += r[3];
          }
          double* o = out + 4 * (a * (a + 1) / 2 + b);
          o[0] = v0;
          o[1] = v1;
          o[2] = v2;
          o[3] = v3;
        }
      }
    }
}

// ----------------------------------------------------------------- block combine
// Per problem: out_j = sum_a C[a + j nin] in_a over nin input vectors, for
// nout outputs, every input read once.  Output j goes to slot j of the
// problem's output block, or to slot j + gap from j = nsplit on (the LOBPCG
// layout [E | W | P] writes E and P and skips W).  A thread owns one complex
// entry of every output; NOUT (16 or 32) is the accumulator count.
#define COMBX_NIN_MAX 48
template <int NOUT>
__global__ __launch_bounds__(256) void k_combine_x(long long n, int nin, int nout, int nsplit, int gap,
                                                   const int* pmap, const C<double>* X0, long long xs, int xvpp,
                                                   C<double>* Y0, long long ys, int yvpp, const double* Cm0)
{
  __shared__ C<double> sC[COMBX_NIN_MAX * NOUT];
  const int jb = blockIdx.y;
  const int prob = pmap ? pmap[jb] : jb;
  const C<double>* X = X0 + (size_t)prob * xvpp * xs;
  C<double>* Y = Y0 + (size_t)prob * yvpp * ys;
  const C<double>* Cm = (const C<double>*)Cm0 + (size_t)jb * nin * nout;
  // coefficients as [a][j] with a row of NOUT (zero beyond nout)
  for (int w = threadIdx.x; w < nin * NOUT; w += blockDim.x) {
    const int a = w / NOUT, j = w - a * NOUT;
    sC[w] = j < nout ? Cm[a + (size_t)j * nin] : czero<double>();
  }
  __syncthreads();
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  C<double> acc[NOUT];
#pragma unroll
  for (int q = 0; q < NOUT; q++) acc[q] = czero<double>();
  int a = 0;
  for (; a + 2 <= nin; a += 2) {
    const C<double> u0 = X[(size_t)a * xs + idx], u1 = X[(size_t)(a + 1) * xs + idx];
    const C<double>* c0 = sC + a * NOUT;
#pragma unroll
    for (int q = 0; q < NOUT; q++) acc[q] += c0[q] * u0 + c0[NOUT + q] * u1;
  }
  if (a < nin) {
    const C<double> u0 = X[(size_t)a * xs + idx];
    const C<double>* c0 = sC + a * NOUT;
#pragma unroll
    for (int q = 0; q < NOUT; q++) acc[q] += c0[q] * u0;
  }
#pragma unroll
  for (int q = 0; q < NOUT; q++)
    if (q < nout) Y[(size_t)(q < nsplit ? q : q + gap) * ys + idx] = acc[q];
}

// LOBPCG block update of one problem's basis X = [E | W | P] (3m vectors)
// with the Ritz coefficients c (3m x m, column-major: c[a + q 3m] for output
// q): P_new_q = sum_{a >= m} c[a, q] x_a (the W and P part), then
// E_new_q = P_new_q + sum_{a < m} c[a, q] e_a, written to the E and P slots
// of Y.  Every input is read once and the shared part is summed once (3 m^2
// complex multiply-adds per entry instead of 6 m^2); NM >= m accumulators.
template <int NM>
__global__ __launch_bounds__(256) void k_lobpcg_combine(long long n, int m, const int* pmap, const C<double>* X0,
                                                        long long xs, C<double>* Y0, long long ys, const double* Cm0)
{
  __shared__ C<double> sC[48 * NM];
  // bit a: input a has a non-zero coefficient.  The Ritz step zeroes the
  // rows of the columns it drops (no P at the first step, the W of locked
  // vectors, dependent directions), and those inputs are not read at all.
  __shared__ unsigned long long rows_nz;
  const int jb = blockIdx.y;
  const int prob = pmap ? pmap[jb] : jb;
  const int k = 3 * m;
  const C<double>* X = X0 + (size_t)prob * k * xs;
  C<double>* Y = Y0 + (size_t)prob * k * ys;
  const C<double>* Cm = (const C<double>*)Cm0 + (size_t)jb * k * 2 * m;  // the E_new block of the 3m x 2m matrix
  if (threadIdx.x == 0) rows_nz = 0ull;
  for (int w = threadIdx.x; w < k * NM; w += blockDim.x) {
    const int a = w / NM, q = w - a * NM;
    sC[w] = q < m ? Cm[a + (size_t)q * k] : czero<double>();
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    bool nz = false;
    for (int q = 0; q < m; q++) {
      const C<double> c = sC[threadIdx.x * NM + q];
      nz = nz || c.r != 0.0 || c.i != 0.0;
    }
    if (nz) atomicOr(&rows_nz, 1ull << threadIdx.x);
  }
  __syncthreads();
  const unsigned long long nzm = rows_nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  C<double> acc[NM];
#pragma unroll
  for (int q = 0; q < NM; q++) acc[q] = czero<double>();
  // W and P part (2m inputs, m even), by pairs; an input whose row is zero
  // is not read (its slot may hold anything: the W of a locked vector, a P
  // before the first direction)
  for (int a = m; a < k; a += 2) {
    const bool n0 = (nzm >> a) & 1ull, n1 = (nzm >> (a + 1)) & 1ull;
    if (!n0 && !n1) continue;
    const C<double> u0 = n0 ? X[(size_t)a * xs + idx] : czero<double>();
    const C<double> u1 = n1 ? X[(size_t)(a + 1) * xs + idx] : czero<double>();
    const C<double>* c0 = sC + a * NM;
#pragma unroll
    for (int q = 0; q < NM; q++) acc[q] += c0[q] * u0 + c0[NM + q] * u1;
  }
#pragma unroll
  for (int q = 0; q < NM; q++)
    if (q < m) Y[(size_t)(2 * m + q) * ys + idx] = acc[q];
  for (int a = 0; a < m; a += 2) {
    const bool n0 = (nzm >> a) & 1ull, n1 = (nzm >> (a + 1)) & 1ull;
    if (!n0 && !n1) continue;
    const C<double> u0 = n0 ? X[(size_t)a * xs + idx] : czero<double>();
    const C<double> u1 = n1 ? X[(size_t)(a + 1) * xs + idx] : czero<double>();
    const C<double>* c0 = sC + a * NM;
#pragma unroll
    for (int q = 0; q < NM; q++) acc[q] += c0[q] * u0 + c0[NM + q] * u1;
  }
#pragma unroll
  for (int q = 0; q < NM; q++)
    if (q < m) Y[(size_t)q * ys + idx] = acc[q];
}

#define COMB_MMAX 48
__global__ __launch_bounds__(256) void k_block_combine(long long n, int m, const int* pmap, const C<double>* U0,
                                                       long long us, C<double>* E0, long long es, const double* Cm0)
{
  __shared__ C<double> sC[COMB_MMAX * COMB_MMAX];
  const int jb = blockIdx.y;
  const int prob = pmap ? pmap[jb] : jb;
  const C<double>* U = U0 + (size_t)prob * m * us;
  C<double>* E = E0 + (size_t)prob * m * es;
  const C<double>* Cm = (const C<double>*)Cm0 + (size_t)jb * m * m;
  for (int w = threadIdx.x; w < m * m; w += blockDim.x) sC[w] = Cm[w];
  __syncthreads();
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  for (int j0 = 0; j0 < m; j0 += 16) {
    C<double> acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = czero<double>();
    for (int a = 0; a < m; a++) {
      const C<double> u = U[(size_t)a * us + idx];
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (j0 + q < m) acc[q] += sC[a + (j0 + q) * m] * u;
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (j0 + q < m) E[(size_t)(j0 + q) * es + idx] = acc[q];
  }
}

// ----------------------------------------------- Householder M-orthonormalisation
// The reference's unit vectors (edge2d.c:6154-6210): vector idx is the
// x-edge (i, j) with j = 1, 3, 5, ... taking idx row by row, scaled to unit
// M-norm, s = sqrt(0.25 / mx), mx = hy / (6 hx); they are M-orthonormal (M
// couples an x-edge only to the x-edges of the rows above and below).
MXH_DEV void unit_ij(int idx, int nx, int& i, int& j)
{
  i = idx % nx;
  j = 1 + 2 * (idx / nx);
}

// unit(i,j)^H M y = sqrt(0.25 mx) (4 y(i,j) + y(i,j-1) + y(i,j+1)), the row
// above the last one through the Bloch factor yf
MXH_DEV C<double> unit_mprod(const C<double>* y, int i, int j, int nx, int ny, double mx, C<double> yf)
{
  const C<double> up = j + 1 < ny ? y[i + (size_t)(j + 1) * nx] : yf * y[i];
  return sqrt(0.25 * mx) * (4.0 * y[i + (size_t)j * nx] + y[i + (size_t)(j - 1) * nx] + up);
}

// Step idx of the forward pass, one thread: x = x_idx, a = x^H M x (device),
// mx_ = M x (device, all entries).  Builds the Householder vector in place
// (edge2d.c:6238-6262): alpha = |x|_M with the sign of Re(beta), beta =
// unit^H M x, x += alpha unit, tau = 1 / (alpha^2 + beta alpha) (0 when x
// = 0), and updates M x by alpha M unit (three entries) so that mx_ is the
// Householder vector's M-image.
__global__ void k_hh_build(mxh_level lv, C<double> yf, int idx, C<double>* x, C<double>* mx_, const double* a,
                           double* tau)
{
  const int nx = lv.nx, ny = lv.ny;
  int i, j;
  unit_ij(idx, nx, i, j);
  const double mxc = lv.hy / lv.hx / 6.0, sc = sqrt(0.25 / mxc);
  double alpha = sqrt(a[0]);
  if (alpha == 0.0) {
    tau[2 * idx] = tau[2 * idx + 1] = 0.0;
    return;
  }
  const C<double> beta = unit_mprod(x, i, j, nx, ny, mxc, yf);
  if (beta.r < 0.0) alpha = -alpha;
  x[i + (size_t)j * nx] += cmk(alpha * sc, 0.0);
  // tau = 1 / (alpha^2 + beta alpha)
  const C<double> den = cmk(alpha * alpha + beta.r * alpha, beta.i * alpha);
  const double dd = den.r * den.r + den.i * den.i;
  tau[2 * idx] = den.r / dd;
  tau[2 * idx + 1] = -den.i / dd;
  // M (alpha s unit): 4 mx alpha s at (i, j), mx alpha s at (i, j -+ 1) (x-edges of one column)
  const double u = alpha * sc * mxc;
  mx_[i + (size_t)j * nx] += cmk(4.0 * u, 0.0);
  mx_[i + (size_t)(j - 1) * nx] += cmk(u, 0.0);
  if (j + 1 < ny) mx_[i + (size_t)(j + 1) * nx] += cmk(u, 0.0);
  else mx_[i] += cconj(yf) * cmk(u, 0.0);
}

// y_v += c_v x for v < nv, c_v = sign * t * d_v with t = tau[2 tidx] (conj
// when cj), d_v = dots[v] (device complex): the reflections' rank-one updates
__global__ __launch_bounds__(256) void k_hh_axpy(long long n, int nv, const C<double>* x, C<double>* Y, long long ys,
                                                 const double* tau, int tidx, int cj, const double* dots)
{
  const int v = blockIdx.y;
  C<double> t = cmk(tau[2 * tidx], cj ? -tau[2 * tidx + 1] : tau[2 * tidx + 1]);
  const C<double> c = -1.0 * (t * cmk(dots[2 * v], dots[2 * v + 1]));
  C<double>* y = Y + (size_t)v * ys;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n; k += (size_t)gridDim.x * blockDim.x)
    y[k] += c * x[k];
}

// x_v -= (unit_idx^H M x_v) unit_idx for v < nv (edge2d.c:6281-6283)
__global__ void k_hh_unitfix(mxh_level lv, C<double> yf, int idx, int nv, C<double>* X, long long xs)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int i, j;
  unit_ij(idx, lv.nx, i, j);
  const double mxc = lv.hy / lv.hx / 6.0, sc = sqrt(0.25 / mxc);
  C<double>* x = X + (size_t)v * xs;
  const C<double> p = unit_mprod(x, i, j, lv.nx, lv.ny, mxc, yf);
  x[i + (size_t)j * lv.nx] -= sc * p;
}

// V = [unit_0 .. unit_{k-1}]
__global__ void k_hh_units(mxh_level lv, int k, C<double>* V, long long vs)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= k) return;
  int i, j;
  unit_ij(v, lv.nx, i, j);
  V[(size_t)v * vs + i + (size_t)j * lv.nx] = cmk(sqrt(0.25 / (lv.hy / lv.hx / 6.0)), 0.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(long long n, C<T>* D, long long ds, const C<T>* S, long long ss,
                                                const int* smap)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  D[(size_t)v * ds + idx] = S[(size_t)smap[v] * ss + idx];
}

// ------------------------------------------------ pointwise edge Gauss-Seidel
// One colour of the pointwise (one edge at a time) Gauss-Seidel smoother for
// alpha K + beta M (edge2d.c:3892-5600, the smoother of the Hiptmair cycle):
// x_e += (b - A x)_e / A_ee.  An x-edge couples to x-edges only in the rows
// above and below, a y-edge to y-edges only in the columns left and right,
// so colours 0 / 1 (x-edges of even / odd rows) and 2 / 3 (y-edges of even /
// odd columns) are independent sets on grids with even nx and ny.
template <typename T>
__global__ __launch_bounds__(256) void k_edge_gs_phase(mxh_level lv, mxh_bloch bl, const C<T>* B, long long bs,
                                                       C<T>* X, long long xs, const T* eta, mxh_coef cf,
                                                       int colour)
{
  const int v = blockIdx.y;
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N / 2) return;
  Geo<T> g = make_geo<T>(lv, bl, v);
  C<T> axy, mx, my;
  coefs<T>(lv, cf, v, axy, mx, my);
  C<T>* xv = X + (size_t)v * xs;
  const C<T>* ex = xv;
  const C<T>* ey = xv + N;
  const C<T>* b = B + (size_t)v * bs;
  if (colour < 2) {  // x-edge (i, j), j of parity colour
    const int i = (int)(p % lv.nx), j = 2 * (int)(p / lv.nx) + colour;
    const C<T> x00 = g.at(ex, i, j);
    const C<T> c00 = x00 - g.at(ex, i, j + 1) - g.at(ey, i, j) + g.at(ey, i + 1, j);          // cell (i, j)
    const C<T> c0m = g.at(ex, i, j - 1) - x00 - g.at(ey, i, j - 1) + g.at(ey, i + 1, j - 1);  // cell (i, j-1)
    const T e00 = eta_at(eta, g.nx, g.ny, i, j), e0m = eta_at(eta, g.nx, g.ny, i, j - 1);
    const C<T> ax = axy * (e00 * c00 - e0m * c0m) + mx * (T(4) * x00 + g.at(ex, i, j - 1) + g.at(ex, i, j + 1));
    const C<T> diag = (e00 + e0m) * axy + T(4) * mx;
    const size_t k = i + (size_t)j * lv.nx;
    xv[k] = x00 + cdiv(b[k] - ax, diag);
  } else {  // y-edge (i, j), i of parity colour - 2
    const int hx2 = lv.nx >> 1;
    const int i = 2 * (int)(p % hx2) + (colour - 2), j = (int)(p / hx2);
    const C<T> y00 = g.at(ey, i, j);
    const C<T> c00 = g.at(ex, i, j) - g.at(ex, i, j + 1) - y00 + g.at(ey, i + 1, j);          // cell (i, j)
    const C<T> cm0 = g.at(ex, i - 1, j) - g.at(ex, i - 1, j + 1) - g.at(ey, i - 1, j) + y00;  // cell (i-1, j)
    const T e00 = eta_at(eta, g.nx, g.ny, i, j), em0 = eta_at(eta, g.nx, g.ny, i - 1, j);
    const C<T> ay = axy * (em0 * cm0 - e00 * c00) + my * (T(4) * y00 + g.at(ey, i - 1, j) + g.at(ey, i + 1, j));
    const C<T> diag = (e00 + em0) * axy + T(4) * my;
    const size_t k = N + i + (size_t)j * lv.nx;
    xv[k] = y00 + cdiv(b[k] - ay, diag);
  }
}

// ----------------------------------------------------------------- node space
template <typename T>
struct NodeCoef {
  T ac, ax, ay, ad;
};

template <typename T>
MXH_DEV NodeCoef<T> node_coef(const mxh_level& lv, double alpha)
{
  const double hx = lv.hx, hy = lv.hy;
  NodeCoef<T> c;
  c.ac = T(4.0 * alpha * (hx / hy + hy / hx) / 3.0);
  c.ax = T(alpha * (hx / hy - 2.0 * hy / hx) / 3.0);
  c.ay = T(alpha * (hy / hx - 2.0 * hx / hy) / 3.0);
  c.ad = T(-alpha * (hx / hy + hy / hx) / 6.0);
  return c;
}

// A x at node (i,j) without the diagonal term
template <typename T>
MXH_DEV C<T> node_offdiag(const Geo<T>& g, const C<T>* x, int i, int j, const NodeCoef<T>& c)
{
  return c.ax * (g.at(x, i - 1, j) + g.at(x, i + 1, j)) + c.ay * (g.at(x, i, j - 1) + g.at(x, i, j + 1)) +
         c.ad * (g.at(x, i - 1, j - 1) + g.at(x, i + 1, j - 1) + g.at(x, i - 1, j + 1) + g.at(x, i + 1, j + 1));
}

template <typename T>
__global__ __launch_bounds__(256) void k_node_addeval(mxh_level lv, mxh_bloch bl, const C<T>* X, long long xs,
                                                      C<T>* Y, long long ys, double ar, double ai,
                                                      int accumulate)
{
  const int v = blockIdx.y;
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int i = (int)(idx % lv.nx), j = (int)(idx / lv.nx);
  Geo<T> g = make_geo<T>(lv, bl, v);
  NodeCoef<T> c = node_coef<T>(lv, 1.0);
  const C<T>* x = X + (size_t)v * xs;
  C<T> a = c.ac * x[idx] + node_offdiag(g, x, i, j, c);
  C<T> al = cmk(T(ar), T(ai));
  C<T>* y = Y + (size_t)v * ys;
  if (accumulate) y[idx] += al * a;
  else y[idx] = al * a;
}

template <typename T>
__global__ __launch_bounds__(256) void k_node_gs_phase(mxh_level lv, mxh_bloch bl, const C<T>* B, long long bs,
                                                       C<T>* X, long long xs, double alpha, int ci, int cj)
{
  const int v = blockIdx.y;
  const int hx2 = lv.nx >> 1;
  const size_t np = (size_t)hx2 * (lv.ny >> 1);
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int i = 2 * (int)(p % hx2) + ci, j = 2 * (int)(p / hx2) + cj;
  Geo<T> g = make_geo<T>(lv, bl, v);
  NodeCoef<T> c = node_coef<T>(lv, alpha);
  C<T>* x = X + (size_t)v * xs;
  const C<T>* b = B + (size_t)v * bs;
  const size_t k = i + (size_t)j * lv.nx;
  x[k] = (T(1) / c.ac) * (b[k] - node_offdiag(g, (const C<T>*)x, i, j, c));
}

template <typename T>
__global__ __launch_bounds__(256) void k_node_resid_restrict(mxh_level lf, mxh_bloch bl, const C<T>* BF,
                                                             long long bfs, const C<T>* XF, long long xfs,
                                                             C<T>* BC, long long bcs, C<T>* XC, long long xcs)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> g = make_geo<T>(lf, bl, v);
  NodeCoef<T> c = node_coef<T>(lf, 1.0);
  const C<T>* b = BF + (size_t)v * bfs;
  const C<T>* x = XF + (size_t)v * xfs;
  C<T> s = czero<T>();
  for (int dj = -1; dj <= 1; dj++)
    for (int di = -1; di <= 1; di++) {
      const int i = 2 * I + di, j = 2 * J + dj;
      const T w = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5));
      C<T> r = g.at(b, i, j) - c.ac * g.at(x, i, j) - node_offdiag(g, x, i, j, c);
      s += w * r;
    }
  BC[(size_t)v * bcs + idx] = s;
  XC[(size_t)v * xcs + idx] = czero<T>();
}

template <typename T>
__global__ __launch_bounds__(256) void k_node_prolong(mxh_level lf, mxh_bloch bl, const C<T>* CV, long long cs,
                                                      C<T>* FV, long long fs, double ar, double ai)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> gc = make_geo<T>(lf, bl, v);
  gc.nx = cnx;
  gc.ny = cny;
  const C<T>* c = CV + (size_t)v * cs;
  C<T>* f = FV + (size_t)v * fs;
  const C<T> al = cmk(T(ar), T(ai));
  C<T> c00 = c[idx], c10 = gc.at(c, I + 1, J), c01 = gc.at(c, I, J + 1), c11 = gc.at(c, I + 1, J + 1);
  const size_t r0 = (size_t)(2 * J) * lf.nx + 2 * I, r1 = r0 + lf.nx;
  f[r0] += al * c00;
  f[r0 + 1] += al * (T(0.5) * (c00 + c10));
  f[r1] += al * (T(0.5) * (c00 + c01));
  f[r1 + 1] += al * (T(0.25) * (c00 + c10 + c01 + c11));
}

template <typename T>
__global__ __launch_bounds__(256) void k_node_restrict(mxh_level lf, mxh_bloch bl, const C<T>* FV, long long fs,
                                                       C<T>* CV, long long cs, double ar, double ai)
{
  const int v = blockIdx.y;
  const int cnx = lf.nx >> 1, cny = lf.ny >> 1;
  const size_t NC = (size_t)cnx * cny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const int I = (int)(idx % cnx), J = (int)(idx / cnx);
  Geo<T> g = make_geo<T>(lf, bl, v);
  const C<T>* f = FV + (size_t)v * fs;
  C<T> s = czero<T>();
  for (int dj = -1; dj <= 1; dj++)
    for (int di = -1; di <= 1; di++) {
      const T w = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5));
      s += w * g.at(f, 2 * I + di, 2 * J + dj);
    }
  CV[(size_t)v * cs + idx] += cmk(T(ar), T(ai)) * s;
}

// ----------------------------------------------------------------- gradient lifting
// Lowest Bloch plane wave of the node space, phi(i,j) = exp(i (tx i + ty j)),
// tx = arg(xf) / nx: an exact eigenvector of the constant-coefficient Q1
// Laplacian.  Its eigenvalue mu ~ (k / n)^2 is below FP32 resolution near the
// Gamma point, so the FP32 node solve handles it separately (see k_lift_rhs).
// The kernels below walk rows (block-uniform j) with thread columns
// i = tid, tid + blockDim, ...: phi = row(j) col(i), col advanced by a fixed
// complex step, so a thread evaluates one sincos per row and one per column
// start.
struct PlaneWave {
  double tx, ty;
  C<double> c0, step;  // col(tid), exp(i tx blockDim)
  MXH_DEV PlaneWave(const Geo<double>& g)
  {
    tx = atan2(g.xf.i, g.xf.r) / g.nx;
    ty = atan2(g.yf.i, g.yf.r) / g.ny;
    double s, c;
    sincos(tx * (double)threadIdx.x, &s, &c);
    c0 = cmk(c, s);
    sincos(tx * (double)blockDim.x, &s, &c);
    step = cmk(c, s);
  }
  MXH_DEV C<double> row(int j) const
  {
    double s, c;
    sincos(ty * (double)j, &s, &c);
    return cmk(c, s);
  }
  MXH_DEV C<double> col(int i) const
  {
    double s, c;
    sincos(tx * (double)i, &s, &c);
    return cmk(c, s);
  }
};

// u = e - x (u, e FP64; x of precision T, optional); bg (T) = G^H M u; xg = 0
// (xg optional).
// u must not alias e when x is given: threads read e at neighbouring nodes.
// With phi non-null, per-block partial sums of conj(phi) . (G^H M u) (FP64)
// go to phi (2 doubles per vector and block).
template <typename T>
MXH_DEV C<double> lift_point(const Geo<double>& g, const Geo<T>& gt, const C<double>* ex, const C<double>* ey,
                             const C<T>* xx, const C<T>* xy, int i, int j, double mx, double my, C<double>* u,
                             size_t N)
{
  // u at the 12 edges needed: x(i-1, j-1..j+1), x(i, j-1..j+1),
  // y(i-1..i+1, j-1), y(i-1..i+1, j)
  C<double> X[2][3], Y[3][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) X[a][b] = g.at(ex, i - 1 + a, j - 1 + b);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) Y[a][b] = g.at(ey, i - 1 + a, j - 1 + b);
  if (xx) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) X[a][b] -= ccast<double>(gt.at(xx, i - 1 + a, j - 1 + b));
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) Y[a][b] -= ccast<double>(gt.at(xy, i - 1 + a, j - 1 + b));
  }
  const C<double> mxl = mx * (4.0 * X[0][1] + X[0][0] + X[0][2]);  // (M u) at x(i-1, j)
  const C<double> mxr = mx * (4.0 * X[1][1] + X[1][0] + X[1][2]);  // x(i, j)
  const C<double> myb = my * (4.0 * Y[1][0] + Y[0][0] + Y[2][0]);  // y(i, j-1)
  const C<double> myt = my * (4.0 * Y[1][1] + Y[0][1] + Y[2][1]);  // y(i, j)
  if (u) {
    const size_t k = i + (size_t)j * g.nx;
    u[k] = X[1][1];
    u[N + k] = Y[1][1];
  }
  // G^H: node += incoming - outgoing edges (node2d.c:1312-1350)
  return mxl - mxr + myb - myt;
}

template <typename T, bool FRAME>
__global__ __launch_bounds__(256) void k_lift_rhs(mxh_level lv, mxh_bloch bl, const C<double>* E, long long es,
                                                  const int* emap, const C<T>* XV, long long xs, C<double>* U,
                                                  long long us, const int* umap, C<T>* BG, long long bgs, C<T>* XG,
                                                  long long xgs, double* phi, int slot0)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const Geo<double> g = make_geo<double>(lv, bl, v);
  const Geo<T> gt = make_geo<T>(lv, bl, v);
  const C<double>* ex = E + (size_t)(emap ? emap[v] : v) * es;
  const C<double>* ey = ex + N;
  const C<T>* xx = XV ? XV + (size_t)v * xs : nullptr;
  const C<T>* xy = XV ? xx + N : nullptr;
  C<double>* u = U ? U + (size_t)(umap ? umap[v] : v) * us : nullptr;
  C<T>* bg = BG + (size_t)v * bgs;
  C<T>* xg = XG ? XG + (size_t)v * xgs : nullptr;
  const double mx = lv.hy / lv.hx / 6.0, my = lv.hx / lv.hy / 6.0;
  double pr = 0.0, pi = 0.0;
  PlaneWave pw(g);
  if (FRAME) {
    const int nf = (int)frame_points(nx, ny);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) {
      int i, j;
      frame_point(k, nx, ny, i, j);
      const C<double> d = lift_point<T>(g, gt, ex, ey, xx, xy, i, j, mx, my, u, N);
      const size_t q = i + (size_t)j * nx;
      bg[q] = ccast<T>(d);
      if (xg) xg[q] = czero<T>();
      if (phi) {
        const C<double> ph = pw.row(j) * pw.col(i);
        pr += ph.r * d.r + ph.i * d.i;  // conj(phi) d
        pi += ph.r * d.i - ph.i * d.r;
      }
    }
  } else {
    // each thread walks its columns down a band of rows and keeps the values
    // of u = e - x that the next row's stencil shares with this one (five
    // new values of each of e and x per point instead of twelve)
    const int j0 = max(1, row_first(ny)), j1 = min(ny - 1, row_end(ny));
    auto ux = [&](size_t q) {
      C<double> w = ex[q];
      if (xx) w -= ccast<double>(xx[q]);
      return w;
    };
    auto uy = [&](size_t q) {
      C<double> w = ey[q];
      if (xy) w -= ccast<double>(xy[q]);
      return w;
    };
    const C<double> rstep = phi ? pw.row(1) : czero<double>();
    if (j0 < j1)
      for (int i = 1 + (int)threadIdx.x; i < nx - 1; i += blockDim.x) {
        size_t k = i + (size_t)j0 * nx;
        // Xab = u_x(i - 1 + a, j - 1 + b), Yab = u_y(i - 1 + a, j - 1 + b)
        C<double> X00 = ux(k - nx - 1), X01 = ux(k - 1), X10 = ux(k - nx), X11 = ux(k);
        C<double> Y00 = uy(k - nx - 1), Y10 = uy(k - nx), Y20 = uy(k - nx + 1);
        C<double> rowf = phi ? pw.row(j0) : czero<double>();
        const C<double> cfi = phi ? pw.col(i) : czero<double>();
        for (int j = j0; j < j1; j++, k += nx) {
          const C<double> X02 = ux(k + nx - 1), X12 = ux(k + nx);
          const C<double> Y01 = uy(k - 1), Y11 = uy(k), Y21 = uy(k + 1);
          const C<double> mxl = mx * (4.0 * X01 + X00 + X02);  // (M u) at x(i-1, j)
          const C<double> mxr = mx * (4.0 * X11 + X10 + X12);  // x(i, j)
          const C<double> myb = my * (4.0 * Y10 + Y00 + Y20);  // y(i, j-1)
          const C<double> myt = my * (4.0 * Y11 + Y01 + Y21);  // y(i, j)
          if (u) {
            u[k] = X11;
            u[N + k] = Y11;
          }
          // G^H: node += incoming - outgoing edges (node2d.c:1312-1350)
          const C<double> d = mxl - mxr + myb - myt;
          bg[k] = ccast<T>(d);
          if (xg) xg[k] = czero<T>();
          if (phi) {
            const C<double> ph = rowf * cfi;
            pr += ph.r * d.r + ph.i * d.i;  // conj(phi) d
            pi += ph.r * d.i - ph.i * d.r;
            rowf = rowf * rstep;
          }
          X00 = X01;
          X01 = X02;
          X10 = X11;
          X11 = X12;
          Y00 = Y01;
          Y10 = Y11;
          Y20 = Y21;
        }
      }
  }
  if (phi) {
    double a = block_sum256(pr, sh);
    double b = block_sum256(pi, sh);
    if (threadIdx.x == 0) {
      size_t w = (size_t)(slot0 + blockIdx.x) * (2 * gridDim.y) + 2 * v;
      phi[w] = a;
      phi[w + 1] = b;
    }
  }
}

// Laplacian eigenvalue of the plane wave: ac + 2 ax cos tx + 2 ay cos ty +
// 4 ad cos tx cos ty, written with s = 1 - cos = 2 sin^2(t/2) (no cancellation)
MXH_DEV double plane_wave_mu(const mxh_level& lv, double tx, double ty)
{
  const NodeCoef<double> nc = node_coef<double>(lv, 1.0);
  const double sx = 2.0 * sin(0.5 * tx) * sin(0.5 * tx), sy = 2.0 * sin(0.5 * ty) * sin(0.5 * ty);
  return -2.0 * nc.ax * sx - 2.0 * nc.ay * sy - 4.0 * nc.ad * (sx + sy) + 4.0 * nc.ad * sx * sy;
}

// bg -= (c / N) phi with c = phi^H d from k_lift_rhs; coef[v] = c / (N mu)
// (zero at the Gamma point, where the coarse node matrix carries the
// regularisation of the constant mode instead).
template <typename T>
__global__ __launch_bounds__(256) void k_lift_deflate(mxh_level lv, mxh_bloch bl, C<T>* BG, long long bgs,
                                                      const double* csum, double* coef)
{
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const Geo<double> g = make_geo<double>(lv, bl, v);
  const double cr = csum[2 * v] / N, ci = csum[2 * v + 1] / N;
  const bool gamma = g.xf.r == 1.0 && g.xf.i == 0.0 && g.yf.r == 1.0 && g.yf.i == 0.0;
  PlaneWave pw(g);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double mu = plane_wave_mu(lv, pw.tx, pw.ty);
    coef[2 * v] = gamma ? 0.0 : cr / mu;
    coef[2 * v + 1] = gamma ? 0.0 : ci / mu;
  }
  if (gamma) return;
  C<T>* bg = BG + (size_t)v * bgs;
  const C<double> cc = cmk(cr, ci);
  for (int j = row_first(ny), je = row_end(ny); j < je; j++) {
    const C<double> rowf = cc * pw.row(j);
    C<double> cf = pw.c0;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const size_t k = i + (size_t)j * nx;
      bg[k] = ccast<T>(ccast<double>(bg[k]) - rowf * cf);
      cf = cf * pw.step;
    }
  }
}

// Per-block partial sums of phi^H xg (FP64; 2 doubles per vector and block).
template <typename T>
__global__ __launch_bounds__(256) void k_phi_dot(mxh_level lv, mxh_bloch bl, const C<T>* XG, long long xgs,
                                                 double* part)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const Geo<double> g = make_geo<double>(lv, bl, v);
  const C<T>* xg = XG + (size_t)v * xgs;
  PlaneWave pw(g);
  double pr = 0.0, pi = 0.0;
  for (int j = row_first(ny), je = row_end(ny); j < je; j++) {
    const C<double> rowf = pw.row(j);
    C<double> cf = pw.c0;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const C<double> ph = rowf * cf, d = ccast<double>(xg[i + (size_t)j * nx]);
      pr += ph.r * d.r + ph.i * d.i;
      pi += ph.r * d.i - ph.i * d.r;
      cf = cf * pw.step;
    }
  }
  double a = block_sum256(pr, sh);
  double b = block_sum256(pi, sh);
  if (threadIdx.x == 0) {
    size_t w = (size_t)blockIdx.x * (2 * gridDim.y) + 2 * v;
    part[w] = a;
    part[w + 1] = b;
  }
}

// e (FP64) = u - x + alpha G (xg + a_v phi) with a_v = coef_v - phi^H xg / N
// (x of precision T, optional; u may be e itself: each thread reads and
// writes only its own edges): the plane-wave component of the node correction is replaced by the exact one
// (coef from k_lift_deflate; xg has none in exact arithmetic since its right
// hand side was deflated, so what phi^H xg finds is rounding).
template <typename T>
__global__ __launch_bounds__(256) void k_lift_apply(mxh_level lv, mxh_bloch bl, const C<T>* XG, long long xgs,
                                                    const C<double>* U, long long us, const int* umap,
                                                    const C<T>* XV, long long xs, C<double>* E, long long es,
                                                    const int* emap, double alpha, const double* coef,
                                                    const double* pdot)
{
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const Geo<T> gt = make_geo<T>(lv, bl, v);
  const C<T>* xg = XG + (size_t)v * xgs;
  const C<double>* u = U ? U + (size_t)(umap ? umap[v] : v) * us : nullptr;  // null: u = 0
  const C<T>* xv = XV ? XV + (size_t)v * xs : nullptr;
  C<double>* e = E + (size_t)(emap ? emap[v] : v) * es;
  const Geo<double> g = make_geo<double>(lv, bl, v);
  PlaneWave pw(g);
  C<double> a = czero<double>(), ex1 = czero<double>(), ey1 = czero<double>();
  if (coef) {
    a = cmk(coef[2 * v] - pdot[2 * v] / N, coef[2 * v + 1] - pdot[2 * v + 1] / N);
    double s, c;
    sincos(pw.tx, &s, &c);
    ex1 = cmk(c - 1.0, s);  // phi(i+1,j) / phi(i,j) - 1
    sincos(pw.ty, &s, &c);
    ey1 = cmk(c - 1.0, s);
  }
  for (int j = row_first(ny), je = row_end(ny); j < je; j++) {
    const bool jin = j <= ny - 2;
    const C<double> rowf = coef ? a * pw.row(j) : czero<double>();
    C<double> cf = pw.c0;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const size_t k = i + (size_t)j * nx;
      const C<double> x0 = ccast<double>(xg[k]);
      C<double> gx, gy;
      if (jin && i <= nx - 2) {
        gx = ccast<double>(xg[k + 1]) - x0;
        gy = ccast<double>(xg[k + nx]) - x0;
      } else {
        gx = ccast<double>(gt.at(xg, i + 1, j)) - x0;
        gy = ccast<double>(gt.at(xg, i, j + 1)) - x0;
      }
      if (coef) {
        const C<double> ph = rowf * cf;
        gx += ph * ex1;
        gy += ph * ey1;
        cf = cf * pw.step;
      }
      C<double> u0 = u ? u[k] : czero<double>(), u1 = u ? u[N + k] : czero<double>();
      if (xv) {
        u0 -= ccast<double>(xv[k]);
        u1 -= ccast<double>(xv[N + k]);
      }
      e[k] = u0 + alpha * gx;
      e[N + k] = u1 + alpha * gy;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gradient(mxh_level lv, mxh_bloch bl, const C<T>* VV, long long vs,
                                                  C<T>* EV_, long long es, double ar, double ai)
{
  const int v = blockIdx.y;
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int i = (int)(idx % lv.nx), j = (int)(idx / lv.nx);
  Geo<T> g = make_geo<T>(lv, bl, v);
  const C<T>* x = VV + (size_t)v * vs;
  C<T>* e = EV_ + (size_t)v * es;
  const C<T> al = cmk(T(ar), T(ai));
  C<T> x0 = x[idx];
  e[idx] += al * (g.at(x, i + 1, j) - x0);
  e[N + idx] += al * (g.at(x, i, j + 1) - x0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_adjgradient(mxh_level lv, mxh_bloch bl, const C<T>* EV_, long long es,
                                                     C<T>* VV, long long vs, double ar, double ai)
{
  const int v = blockIdx.y;
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int i = (int)(idx % lv.nx), j = (int)(idx / lv.nx);
  Geo<T> g = make_geo<T>(lv, bl, v);
  const C<T>* ex = EV_ + (size_t)v * es;
  const C<T>* ey = ex + N;
  C<T> d = g.at(ex, i - 1, j) - ex[idx] + g.at(ey, i, j - 1) - ey[idx];
  VV[(size_t)v * vs + idx] += cmk(T(ar), T(ai)) * d;
}

// ----------------------------------------------------------------- BLAS-1
template <typename T>
__global__ __launch_bounds__(256) void k_scale(long long n, C<T>* X, long long xs, double ar, double ai)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  C<T>* x = X + (size_t)v * xs;
  x[idx] = cmk(T(ar), T(ai)) * x[idx];
}

template <typename T>
__global__ __launch_bounds__(256) void k_axpy(long long n, const C<T>* X, long long xs, C<T>* Y, long long ys,
                                              double ar, double ai)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  Y[(size_t)v * ys + idx] += cmk(T(ar), T(ai)) * X[(size_t)v * xs + idx];
}

template <typename T>
__global__ __launch_bounds__(256) void k_dot(long long n, const C<T>* X, long long xs, const C<T>* Y,
                                             long long ys, double* part)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  double sr = 0.0, si = 0.0;
  const C<T>* x = X + (size_t)v * xs;
  const C<T>* y = Y + (size_t)v * ys;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)n;
       idx += (size_t)gridDim.x * blockDim.x) {
    C<double> a = ccast<double>(x[idx]), b = ccast<double>(y[idx]);
    sr += a.r * b.r + a.i * b.i;
    si += a.r * b.i - a.i * b.r;
  }
  double a = block_sum256(sr, sh);
  double b = block_sum256(si, sh);
  if (threadIdx.x == 0) {
    size_t w = (size_t)blockIdx.x * (2 * gridDim.y) + 2 * v;
    part[w] = a;
    part[w + 1] = b;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_norm(long long n, const C<T>* X, long long xs, int kind, double* part)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  double s = 0.0;
  const C<T>* x = X + (size_t)v * xs;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)n;
       idx += (size_t)gridDim.x * blockDim.x) {
    C<double> a = ccast<double>(x[idx]);
    double q = a.r * a.r + a.i * a.i;
    if (kind == 0) s += q;
    else s = fmax(s, sqrt(q));
  }
  double r = kind == 0 ? block_sum256(s, sh) : block_max256(s, sh);
  if (threadIdx.x == 0) part[(size_t)blockIdx.x * gridDim.y + v] = r;
}

// Sums of the x half and of the y half of each vector (per block partials)
template <typename T>
__global__ __launch_bounds__(256) void k_halfsums(long long half, const C<T>* X, long long xs, double* part)
{
  __shared__ double sh[4];
  const int v = blockIdx.y;
  const C<T>* x = X + (size_t)v * xs;
  double a = 0, b = 0, c = 0, d = 0;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)half;
       idx += (size_t)gridDim.x * blockDim.x) {
    C<double> p = ccast<double>(x[idx]), q = ccast<double>(x[half + idx]);
    a += p.r; b += p.i; c += q.r; d += q.i;
  }
  a = block_sum256(a, sh);
  b = block_sum256(b, sh);
  c = block_sum256(c, sh);
  d = block_sum256(d, sh);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)blockIdx.x * gridDim.y + v) * 4;
    o[0] = a; o[1] = b; o[2] = c; o[3] = d;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_center_apply(long long half, C<T>* X, long long xs, const double* sums)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)half) return;
  const double* s = sums + 4 * v;
  C<T>* x = X + (size_t)v * xs;
  x[idx] -= cmk(T(s[0] / half), T(s[1] / half));
  x[half + idx] -= cmk(T(s[2] / half), T(s[3] / half));
}

template <typename D, typename S>
__global__ __launch_bounds__(256) void k_convert(long long n, C<D>* Dv, long long ds, const C<S>* Sv, long long ss)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  Dv[(size_t)v * ds + idx] = ccast<D>(Sv[(size_t)v * ss + idx]);
}

MXH_DEV uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Complex Gaussian entries by Box-Muller from a counter-based hash.
template <typename T>
__global__ __launch_bounds__(256) void k_random(long long n, C<T>* X, long long xs, unsigned long long seed,
                                                int first)
{
  const int v = blockIdx.y;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n) return;
  uint64_t key = splitmix64(seed ^ splitmix64((uint64_t)(v + first) * 0x100000001B3ULL + idx));
  uint64_t k2 = splitmix64(key);
  double u1 = ((key >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((k2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincos(6.283185307179586 * u2, &s, &c);
  X[(size_t)v * xs + idx] = cmk(T(rad * c), T(rad * s));
}

// ----------------------------------------------------------------- adjoint
__global__ __launch_bounds__(256) void k_curl2(mxh_level lv, mxh_bloch bl, int nv, const C<double>* E,
                                               long long es, const double* w, double* grad)
{
  const size_t N = (size_t)lv.nx * lv.ny;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int i = (int)(idx % lv.nx), j = (int)(idx / lv.nx);
  const double s = 1.0 / (lv.hx * lv.hy);
  double acc = 0.0;
  for (int v = 0; v < nv; v++) {
    Geo<double> g = make_geo<double>(lv, bl, v);
    EdgeV<double> e = edge_view<double>(E, es, v, N);
    C<double> c = cell_curl(g, e, i, j);
    acc += w[v] * (c.r * c.r + c.i * c.i) * s;
  }
  grad[idx] += acc;
}

// ----------------------------------------------------------------- tile-fused AFW sweeps
// One workgroup smooths a T x T block of nodes of one vector (T = R - 20).  It
// loads the surrounding R x R region (halo TILE_H) into LDS in the
// unwrapped Bloch frame, runs whole colour sweeps there (patches of one colour
// are independent; a sweep moves the region of exact values inward by 4 nodes
// forward and 5 backward), and writes the owned block back.  Per sweep the
// vector is read and written once instead of once per colour.  One thread
// owns one 2x2 node block, i.e. one patch of every colour.  LDS is split by
// node parity so that a
// colour phase reads consecutive addresses across a wavefront.  x, y and the
// right-hand side live in LDS (4 R^2 complex values); eta in registers.
// Region side R (FP32: 64, 1024 threads, owned 44; FP64: 48, 576 threads,
// owned 28), halo 10: two forward sweeps with the residual and its
// restriction (4 + 4 + 2) or two backward sweeps (5 + 5) stay exact on the
// owned block, so a launch does both sweeps of a V(2,2) half-cycle.  The four
// LDS arrays take 128 KB (FP32) / 144 KB (FP64).  Measured alternative: FP32
// R = 48 with two workgroups per CU was slower (more halo work and register
// spills).
#define TILE_H 10
// Rows by which the exact region shrinks: per forward / backward AFW sweep,
// per nodal sweep, and for a residual with its restriction.
#define TILE_SHRINK_FWD 4
#define TILE_SHRINK_BWD 5
#define TILE_SHRINK_RES 2

template <int R>
MXH_DEV int lidx(int i, int j)
{
  return ((((j & 1) << 1) | (i & 1)) * ((R / 2) * (R / 2))) + (j >> 1) * (R / 2) + (i >> 1);
}

template <typename T> struct TileCfg;
#ifndef MXH_TILE_R_F32
#define MXH_TILE_R_F32 64
#endif
#ifndef MXH_TILE_WAVES_F32
#define MXH_TILE_WAVES_F32 1
#endif
template <> struct TileCfg<float> { static const int R = MXH_TILE_R_F32; };
template <> struct TileCfg<double> { static const int R = 48; };

// wrap (i,j) into the grid for any number of periods; returns the phase
template <typename T>
MXH_DEV C<T> wrap_any(int& i, int& j, int nx, int ny, C<T> xf, C<T> yf)
{
  C<T> ph = cmk(T(1), T(0));
  while (i < 0) { i += nx; ph = ph * cconj(xf); }
  while (i >= nx) { i -= nx; ph = ph * xf; }
  while (j < 0) { j += ny; ph = ph * cconj(yf); }
  while (j >= ny) { j -= ny; ph = ph * yf; }
  return ph;
}

template <typename T>
MXH_DEV C<T> load_any(const C<T>* a, int i, int j, int nx, int ny, C<T> xf, C<T> yf)
{
  C<T> ph = wrap_any(i, j, nx, ny, xf, yf);
  return ph * a[i + (size_t)j * nx];
}

template <typename T>
MXH_DEV T eta_any(const T* e, int i, int j, int nx, int ny)
{
  while (i < 0) i += nx;
  while (i >= nx) i -= nx;
  while (j < 0) j += ny;
  while (j >= ny) j -= ny;
  return e[i + (size_t)j * nx];
}

// The AFW patch update of k_afw_phase on LDS values (unwrapped frame).
template <int R, typename T>
MXH_DEV void tile_patch(C<T>* sx, C<T>* sy, int i, int j, C<T> b0, C<T> b1, C<T> b2, C<T> b3, T eLL, T eRL,
                        T eLU, T eRU, T mx, T my)
{
  const int iL0 = lidx<R>(i - 1, j), iR0 = lidx<R>(i, j), iCm = lidx<R>(i, j - 1), iC0 = iR0;
  C<T> xLm = sx[lidx<R>(i - 1, j - 1)], xL0 = sx[iL0], xLp = sx[lidx<R>(i - 1, j + 1)];
  C<T> xRm = sx[iCm], xR0 = sx[iR0], xRp = sx[lidx<R>(i, j + 1)];
  C<T> yLm = sy[lidx<R>(i - 1, j - 1)], yCm = sy[iCm], yRm = sy[lidx<R>(i + 1, j - 1)];
  C<T> yL0 = sy[iL0], yC0 = sy[iC0], yR0 = sy[lidx<R>(i + 1, j)];
  C<T> qLL = eLL * (xLm - xL0 - yLm + yCm);
  C<T> qRL = eRL * (xRm - xR0 - yCm + yRm);
  C<T> qLU = eLU * (xL0 - xLp - yL0 + yC0);
  C<T> qRU = eRU * (xR0 - xRp - yC0 + yR0);
  C<T> r0 = b0 - (qLU - qLL) - mx * (T(4) * xL0 + xLm + xLp);
  C<T> r1 = b1 - (qRU - qRL) - mx * (T(4) * xR0 + xRm + xRp);
  C<T> r2 = b2 - (qLL - qRL) - my * (T(4) * yCm + yLm + yRm);
  C<T> r3 = b3 - (qLU - qRU) - my * (T(4) * yC0 + yL0 + yR0);
  C<T> z0, z1, z2, z3;
  afw_local_solve(eLL, eRL, eLU, eRU, mx, my, r0, r1, r2, r3, z0, z1, z2, z3);
  sx[iL0] = xL0 + z0;
  sx[iR0] = xR0 + z1;
  sy[iCm] = yCm + z2;
  sy[iC0] = yC0 + z3;
}

// mode bits: 1 = backward colour order, 2 = x starts at zero, 4 = add the
// prolongation of the coarse vector XC at load, 8 = after the sweeps form the
// residual and write its restriction to BC (and zero XC).
template <typename T, int MODE, int R = TileCfg<T>::R>
__global__ __launch_bounds__((R / 2) * (R / 2)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? MXH_TILE_WAVES_F32 : 1))) void k_afw_tile(mxh_level lv, mxh_bloch bl, const C<T>* B, long long bs,
                                                   const C<T>* Xin, C<T>* Xout, long long xs, const T* eta, int nsweep,
                                                   const C<T>* XCin,
                                                   C<T>* BC, C<T>* XC, long long cs, int ntx)
{
  __shared__ C<T> sx[R * R];
  __shared__ C<T> sy[R * R];
  __shared__ C<T> sbx[R * R];
  __shared__ C<T> sby[R * R];
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const size_t N = (size_t)nx * ny;
  const int tileX = blockIdx.x % ntx, tileY = blockIdx.x / ntx;
  const int ox = tileX * (R - 2 * TILE_H) - TILE_H, oy = tileY * (R - 2 * TILE_H) - TILE_H;  // global node of local (0,0)
  const int bx = threadIdx.x % (R / 2), by = threadIdx.x / (R / 2);
  Geo<T> g = make_geo<T>(lv, bl, v);
  const C<T>* gbx = B + (size_t)v * bs;
  const C<T>* gby = gbx + N;
  const C<T>* gx = Xin + (size_t)v * xs;
  const C<T>* gy = gx + N;
  C<T>* ox_ = Xout + (size_t)v * xs;
  C<T>* oy_ = ox_ + N;
  const T axy = T(1.0 / (lv.hx * lv.hy));
  const T mx = T(lv.hy / lv.hx / 6.0), my = T(lv.hx / lv.hy / 6.0);

  // ---- load x (optionally + P xc) for the thread's 2x2 nodes
  const int cnx = nx >> 1, cny = ny >> 1;
  const size_t NC = (size_t)cnx * cny;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int li = 2 * bx + (c & 1), lj = 2 * by + (c >> 1);
    C<T> vx = czero<T>(), vy = czero<T>();
    int gi = ox + li, gj = oy + lj;
    C<T> ph = wrap_any(gi, gj, nx, ny, g.xf, g.yf);
    const size_t k = gi + (size_t)gj * nx;
    sbx[lidx<R>(li, lj)] = ph * gbx[k];
    sby[lidx<R>(li, lj)] = ph * gby[k];
    if (!(MODE & 2)) {
      vx = gx[k];
      vy = gy[k];
      if (MODE & 4) {
        Geo<T> gc = g;
        gc.nx = cnx;
        gc.ny = cny;
        const C<T>* cx = XCin + (size_t)v * cs;
        const C<T>* cy = cx + NC;
        const int I = gi >> 1, J = gj >> 1;
        if (gj & 1) vx += T(0.25) * (gc.at(cx, I, J) + gc.at(cx, I, J + 1));
        else vx += T(0.5) * cx[I + (size_t)J * cnx];
        if (gi & 1) vy += T(0.25) * (gc.at(cy, I, J) + gc.at(cy, I + 1, J));
        else vy += T(0.5) * cy[I + (size_t)J * cnx];
      }
      vx = ph * vx;
      vy = ph * vy;
    }
    sx[lidx<R>(li, lj)] = vx;
    sy[lidx<R>(li, lj)] = vy;
  }
  // ---- eta of the nine cells the thread's four patches touch
  T er[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) er[a][c] = axy * eta_any(eta, ox + 2 * bx - 1 + a, oy + 2 * by - 1 + c, nx, ny);
  __syncthreads();

  // ---- colour sweeps
  for (int s = 0; s < nsweep; s++) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int c = (MODE & 1) ? 3 - k : k;
      const int ci = c & 1, cj = c >> 1;
      const int i = 2 * bx + ci, j = 2 * by + cj;
      if (i >= 1 && i <= R - 2 && j >= 1 && j <= R - 2)
        tile_patch<R>(sx, sy, i, j, sbx[lidx<R>(i - 1, j)], sbx[lidx<R>(i, j)], sby[lidx<R>(i, j - 1)],
                      sby[lidx<R>(i, j)], er[ci][cj],
                   er[ci + 1][cj], er[ci][cj + 1], er[ci + 1][cj + 1], mx, my);
      __syncthreads();
    }
  }

  // ---- write back the owned nodes
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int li = 2 * bx + (c & 1), lj = 2 * by + (c >> 1);
    const int gi = ox + li, gj = oy + lj;
    if (li >= TILE_H && li < TILE_H + (R - 2 * TILE_H) && lj >= TILE_H && lj < TILE_H + (R - 2 * TILE_H) && gi < nx && gj < ny) {
      const size_t k = gi + (size_t)gj * nx;
      ox_[k] = sx[lidx<R>(li, lj)];
      oy_[k] = sy[lidx<R>(li, lj)];
    }
  }

  if (MODE & 8) {
    // fine residual b - (K + M) x at the thread's four nodes
    C<T> rx[4], ry[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int ci = c & 1, cj = c >> 1;
      const int i = 2 * bx + ci, j = 2 * by + cj;
      rx[c] = czero<T>();
      ry[c] = czero<T>();
      if (i >= 1 && i <= R - 2 && j >= 1 && j <= R - 2) {
        C<T> x00 = sx[lidx<R>(i, j)], x0m = sx[lidx<R>(i, j - 1)], x0p = sx[lidx<R>(i, j + 1)];
        C<T> xm0 = sx[lidx<R>(i - 1, j)], xmp = sx[lidx<R>(i - 1, j + 1)];
        C<T> y00 = sy[lidx<R>(i, j)], ym0 = sy[lidx<R>(i - 1, j)], yp0 = sy[lidx<R>(i + 1, j)];
        C<T> y0m = sy[lidx<R>(i, j - 1)], ypm = sy[lidx<R>(i + 1, j - 1)];
        C<T> c00 = x00 - x0p - y00 + yp0, c0m = x0m - x00 - y0m + ypm, cm0 = xm0 - xmp - ym0 + y00;
        C<T> kx = er[ci + 1][cj + 1] * c00 - er[ci + 1][cj] * c0m;
        C<T> ky = er[ci][cj + 1] * cm0 - er[ci + 1][cj + 1] * c00;
        rx[c] = sbx[lidx<R>(i, j)] - kx - mx * (T(4) * x00 + x0m + x0p);
        ry[c] = sby[lidx<R>(i, j)] - ky - my * (T(4) * y00 + ym0 + yp0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int li = 2 * bx + (c & 1), lj = 2 * by + (c >> 1);
      sx[lidx<R>(li, lj)] = rx[c];
      sy[lidx<R>(li, lj)] = ry[c];
    }
    __syncthreads();
    const int nco = (R - 2 * TILE_H) / 2;
    if ((int)threadIdx.x < nco * nco) {
      const int cI = threadIdx.x % nco, cJ = threadIdx.x / nco;
      const int fi = TILE_H + 2 * cI, fj = TILE_H + 2 * cJ;
      const int GI = tileX * nco + cI, GJ = tileY * nco + cJ;
      if (GI < cnx && GJ < cny) {
        C<T> cxv = T(0.5) * (sx[lidx<R>(fi, fj)] + sx[lidx<R>(fi + 1, fj)]) +
                   T(0.25) * (sx[lidx<R>(fi, fj - 1)] + sx[lidx<R>(fi + 1, fj - 1)] + sx[lidx<R>(fi, fj + 1)] +
                              sx[lidx<R>(fi + 1, fj + 1)]);
        C<T> cyv = T(0.5) * (sy[lidx<R>(fi, fj)] + sy[lidx<R>(fi, fj + 1)]) +
                   T(0.25) * (sy[lidx<R>(fi - 1, fj)] + sy[lidx<R>(fi - 1, fj + 1)] + sy[lidx<R>(fi + 1, fj)] +
                              sy[lidx<R>(fi + 1, fj + 1)]);
        C<T>* bc = BC + (size_t)v * cs;
        C<T>* xc = XC + (size_t)v * cs;
        const size_t k = GI + (size_t)GJ * cnx;
        bc[k] = cxv;
        bc[NC + k] = cyv;
        xc[k] = czero<T>();
        xc[NC + k] = czero<T>();
      }
    }
  }
}

// ----------------------------------------------------------------- tile-fused nodal GS
// The nodal counterpart of k_afw_tile: pointwise Gauss-Seidel for the Q1
// Laplacian in four colours; a sweep moves the exact region inward by 4.
// x and b of the R x R region (R = 64) live in LDS.
template <typename T, int MODE>
__global__ __launch_bounds__(1024) void k_node_tile(mxh_level lv, mxh_bloch bl, const C<T>* B, long long bs,
                                                    const C<T>* Xin, C<T>* Xout, long long xs, int nsweep,
                                                    const C<T>* XCin, C<T>* BC, C<T>* XC, long long cs, int ntx)
{
  constexpr int R = 64;
  __shared__ C<T> sx[R * R];
  __shared__ C<T> sb[R * R];
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny;
  const int tileX = blockIdx.x % ntx, tileY = blockIdx.x / ntx;
  const int TT = R - 2 * TILE_H;
  const int ox = tileX * TT - TILE_H, oy = tileY * TT - TILE_H;
  const int bx = threadIdx.x % (R / 2), by = threadIdx.x / (R / 2);
  Geo<T> g = make_geo<T>(lv, bl, v);
  const C<T>* gb = B + (size_t)v * bs;
  const C<T>* gx = Xin + (size_t)v * xs;
  C<T>* gout = Xout + (size_t)v * xs;
  NodeCoef<T> nc = node_coef<T>(lv, 1.0);
  const T iac = T(1) / nc.ac;
  const int cnx = nx >> 1, cny = ny >> 1;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int li = 2 * bx + (c & 1), lj = 2 * by + (c >> 1);
    int gi = ox + li, gj = oy + lj;
    C<T> ph = wrap_any(gi, gj, nx, ny, g.xf, g.yf);
    const size_t k = gi + (size_t)gj * nx;
    sb[lidx<R>(li, lj)] = ph * gb[k];
    C<T> vx = czero<T>();
    if (!(MODE & 2)) {
      vx = gx[k];
      if (MODE & 4) {
        Geo<T> gc = g;
        gc.nx = cnx;
        gc.ny = cny;
        const C<T>* xc = XCin + (size_t)v * cs;
        const int I = gi >> 1, J = gj >> 1;
        const T wx1 = (gi & 1) ? T(0.5) : T(0), wx0 = (gi & 1) ? T(0.5) : T(1);
        const T wy1 = (gj & 1) ? T(0.5) : T(0), wy0 = (gj & 1) ? T(0.5) : T(1);
        C<T> p = (wx0 * wy0) * xc[I + (size_t)J * cnx];
        if (gi & 1) p += (wx1 * wy0) * gc.at(xc, I + 1, J);
        if (gj & 1) p += (wx0 * wy1) * gc.at(xc, I, J + 1);
        if ((gi & 1) && (gj & 1)) p += (wx1 * wy1) * gc.at(xc, I + 1, J + 1);
        vx += p;
      }
      vx = ph * vx;
    }
    sx[lidx<R>(li, lj)] = vx;
  }
  __syncthreads();
  for (int s = 0; s < nsweep; s++) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int c = (MODE & 1) ? 3 - k : k;
      const int i = 2 * bx + (c & 1), j = 2 * by + (c >> 1);
      if (i >= 1 && i <= R - 2 && j >= 1 && j <= R - 2) {
        C<T> off = nc.ax * (sx[lidx<R>(i - 1, j)] + sx[lidx<R>(i + 1, j)]) +
                   nc.ay * (sx[lidx<R>(i, j - 1)] + sx[lidx<R>(i, j + 1)]) +
                   nc.ad * (sx[lidx<R>(i - 1, j - 1)] + sx[lidx<R>(i + 1, j - 1)] + sx[lidx<R>(i - 1, j + 1)] +
                            sx[lidx<R>(i + 1, j + 1)]);
        sx[lidx<R>(i, j)] = iac * (sb[lidx<R>(i, j)] - off);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int li = 2 * bx + (c & 1), lj = 2 * by + (c >> 1);
    const int gi = ox + li, gj = oy + lj;
    if (li >= TILE_H && li < TILE_H + TT && lj >= TILE_H && lj < TILE_H + TT && gi < nx && gj < ny)
      gout[gi + (size_t)gj * nx] = sx[lidx<R>(li, lj)];
  }
  if (MODE & 8) {
    C<T> r[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int i = 2 * bx + (c & 1), j = 2 * by + (c >> 1);
      r[c] = czero<T>();
      if (i >= 1 && i <= R - 2 && j >= 1 && j <= R - 2) {
        C<T> off = nc.ax * (sx[lidx<R>(i - 1, j)] + sx[lidx<R>(i + 1, j)]) +
                   nc.ay * (sx[lidx<R>(i, j - 1)] + sx[lidx<R>(i, j + 1)]) +
                   nc.ad * (sx[lidx<R>(i - 1, j - 1)] + sx[lidx<R>(i + 1, j - 1)] + sx[lidx<R>(i - 1, j + 1)] +
                            sx[lidx<R>(i + 1, j + 1)]);
        r[c] = sb[lidx<R>(i, j)] - nc.ac * sx[lidx<R>(i, j)] - off;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++) sx[lidx<R>(2 * bx + (c & 1), 2 * by + (c >> 1))] = r[c];
    __syncthreads();
    const int nco = TT / 2;
    if ((int)threadIdx.x < nco * nco) {
      const int cI = threadIdx.x % nco, cJ = threadIdx.x / nco;
      const int fi = TILE_H + 2 * cI, fj = TILE_H + 2 * cJ;
      const int GI = tileX * nco + cI, GJ = tileY * nco + cJ;
      if (GI < cnx && GJ < cny) {
        C<T> acc = czero<T>();
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
          for (int di = -1; di <= 1; di++) {
            const T w = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5));
            acc += w * sx[lidx<R>(fi + di, fj + dj)];
          }
        const size_t k = GI + (size_t)GJ * cnx;
        BC[(size_t)v * cs + k] = acc;
        XC[(size_t)v * cs + k] = czero<T>();
      }
    }
  }
}

// ----------------------------------------------------------------- row-marching smoothers
// A workgroup smooths an owned block of W x Hgt nodes of one vector.  It reads
// the (W + 2H) x (Hgt + 2H) region around it row by row, from the bottom, and
// keeps only a ring of rows of x, y and eta in LDS (parity-split columns), so
// the region can be tall: the halo re-read is (1 + 2H/W)(1 + 2H/Hgt) instead
// of the square tile's (R / (R - 2H))^2, and several workgroups share a CU,
// so one workgroup's loads overlap another's patch solves.
//
// The 4-colour sweep is pipelined over rows.  Forward (colours 0,1 on even
// rows, then 2,3 on odd rows): step s runs colour 0 on row 2s together with
// colour 2 on row 2s-3, then colour 1 on row 2s with colour 3 on row 2s-3.
// A colour-2/3 patch on odd row r needs colours 0,1 done on rows r-1 and
// r+1 and must come after every colour-0/1 patch that reads row r, which is
// exactly rows r-1 and r+1; the three-row skew satisfies both, so the result
// is that of the whole-grid colour sweep.  Backward mirrors it (colours 3,2
// on odd row 2s+1 with colours 1,0 on even row 2s-2).
//
// NSW sweeps run in one pass: sweep q trails sweep q-1 by six rows (its
// leading row needs the previous sweep finished on the rows around it, which
// happens one step after that sweep's trailing group passes them), so x, b
// and eta are read once for NSW sweeps.  Wave w runs sweep w/2 on its
// leading (w even) or trailing row group.  The halo H covers the exactness
// margin of NSW sweeps plus the residual (4 per forward sweep, 5 backward, +1).
// After step s, rows f0 = 2s - 4 + bwd - 6 (NSW - 1) and f0 + 1 are final:
// they are written out and (forward, MODE 8) their residual is formed; the
// coarse row f0 / 2 - 1 is then restricted.
#define MARCH_W 128

template <int NSW>
struct MarchCfg {
  static constexpr int H = NSW == 1 ? 6 : 10;     // halo (nodes)
  // rows of the LDS ring: live rows of one step are r0 - 11 .. r0 + 1 with
  // two sweeps (13), r0 - 5 .. r0 + 1 with one (7).  A power of two keeps
  // the slot arithmetic to a mask; FP64 two-sweep AFW passes use the tight
  // 13 rows (with a 3-row residual ring) so that the residual mode fits two
  // workgroups per CU in LDS (the kernels may choose otherwise per mode).
  template <typename T>
  static constexpr int ring() { return NSW == 1 ? 8 : (sizeof(T) == 8 ? 13 : 16); }
  template <typename T>
  static constexpr int rres() { return sizeof(T) == 8 ? 3 : 4; }
  static constexpr int THREADS = MARCH_W * NSW;
};

// LDS slot of ring row r, column c: rows of MARCH_W values, split by column
// parity (a colour phase reads consecutive addresses across a wavefront);
// the row part is wave-uniform, the column part fixed per lane.
MXH_DEV int mcol(int c) { return (c & 1) * (MARCH_W / 2) + (c >> 1); }
template <int RING>
MXH_DEV int rrow(int r)
{
  // r is wave-uniform at every use: keep the modulo in scalar registers
  if ((RING & (RING - 1)) == 0) return r & (RING - 1);
  return __builtin_amdgcn_readfirstlane(((r % RING) + RING) % RING);
}
template <int RING>
MXH_DEV int msl(int r, int c) { return rrow<RING>(r) * MARCH_W + mcol(c); }
// residual rows: three are live (the restriction of coarse row J reads fine
// rows 2J - 1 .. 2J + 1); rings of 3 (FP64) or 4 rows
template <int RRES>
MXH_DEV int rsl(int r, int c) { return rrow<RRES>(r) * MARCH_W + mcol(c); }

// Region-local geometry of a marching tile: local (0,0) is global node
// (ox, oy).  The halo reaches at most H nodes past the grid and the host
// requires nx, ny >= H and a last tile that ends within one period past the
// grid, so one wrap suffices.  Values are loaded raw
// from wrapped addresses and the Bloch phase is applied where they are used,
// so a load is never waited for before the next barrier.
template <typename T>
struct MarchGeo {
  int nx, ny, ox, oy;
  C<T> xf, yf;
  // value at unwrapped local (lc, lr) from the raw value stored at the
  // wrapped address
  MXH_DEV C<T> fix(C<T> raw, int lc, int lr) const
  {
    const int gi = ox + lc, gj = oy + lr;
    if ((unsigned)gi < (unsigned)nx && (unsigned)gj < (unsigned)ny) return raw;
    C<T> ph = cmk(T(1), T(0));
    if (gi < 0) ph = cconj(xf);
    else if (gi >= nx) ph = xf;
    if (gj < 0) ph = ph * cconj(yf);
    else if (gj >= ny) ph = ph * yf;
    return ph * raw;
  }
  // the column part of fix's phase at local column lc (a lane's columns are
  // fixed for a whole pass, so this is formed once per lane)
  MXH_DEV C<T> colph(int lc) const
  {
    const int gi = ox + lc;
    return gi < 0 ? cconj(xf) : (gi >= nx ? xf : cmk(T(1), T(0)));
  }
  // fix with the column phase cph = colph(lc) given; lr wave-uniform.  The
  // same products in the same order as fix (a phase of exactly 1 leaves raw
  // as it is)
  MXH_DEV C<T> fixc(C<T> raw, C<T> cph, int lr) const
  {
    const int gj = oy + lr;
    if (gj < 0) cph = cph * cconj(yf);
    else if (gj >= ny) cph = cph * yf;
    return cph * raw;
  }
};

// Raw loads of one region row at one column (x, y, eta).
template <typename T>
struct MarchRow {
  C<T> x, y;
  T e;
};

// mode bits as for k_afw_tile: 1 backward, 2 x starts at zero, 4 add the
// prolongation of XC at load, 8 (forward only) residual + restriction to BC
// (and XC = 0).
template <typename T, int MODE, int NSW>
__global__ __launch_bounds__(MarchCfg<NSW>::THREADS) void k_afw_march(mxh_level lv, mxh_bloch bl, const C<T>* B,
                                                                     long long bs, const int* bmap,
                                                                     const C<T>* Xin, C<T>* Xout,
                                                                     long long xs, const T* eta, const C<T>* XCin,
                                                                     C<T>* BC, C<T>* XC, long long cs, int ntx,
                                                                     int W, int Hgt)
{
  constexpr int BWD = MODE & 1;
  constexpr bool ZERO = MODE & 2, PROL = MODE & 4, RES = MODE & 8;
  // two-sweep FP32 passes with a residual or a prolongation take the tight
  // 13-row ring and 3 residual rows as FP64 does: their LDS (37-39 KB) then
  // fits four workgroups per CU
  constexpr bool TIGHT = NSW == 2 && (RES || PROL);
  constexpr int H = MarchCfg<NSW>::H, RING = TIGHT ? 13 : MarchCfg<NSW>::template ring<T>();
  constexpr int RRES = TIGHT ? 3 : MarchCfg<NSW>::template rres<T>();
  constexpr int NRR = RES ? RRES * MARCH_W : 1;
  __shared__ C<T> sx[RING * MARCH_W];
  __shared__ C<T> sy[RING * MARCH_W];
  __shared__ T se[RING * MARCH_W];
  __shared__ C<T> rxs[NRR];
  __shared__ C<T> rys[NRR];
  // MODE 4: coarse x- and y-edge rows of the prolongation, 4-row rings of
  // MARCH_W / 2 + 1 coarse columns, staged one step ahead (as in k_node_march)
  constexpr int NCW = MARCH_W / 2 + 1;
  __shared__ C<T> scx[PROL ? 4 * NCW : 1];
  __shared__ C<T> scy[PROL ? 4 * NCW : 1];
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny, cnx = nx >> 1, cny = ny >> 1;
  const size_t N = (size_t)nx * ny, NC = (size_t)cnx * cny;
  const int tileX = blockIdx.x % ntx, tileY = blockIdx.x / ntx;
  const int RW = W + 2 * H, RH = Hgt + 2 * H;
  const Geo<T> g = make_geo<T>(lv, bl, v);
  const MarchGeo<T> mg{nx, ny, tileX * W - H, tileY * Hgt - H, g.xf, g.yf};
  // a region that does not cross the periodic boundary needs no Bloch phases
  const bool inter = mg.ox >= 0 && mg.ox + RW <= nx && mg.oy >= 0 && mg.oy + RH <= ny;
  // the right-hand side of vector v: slot bmap[v] of B when mapped
  const C<T>* gbx = B + (size_t)(bmap ? bmap[v] : v) * bs;
  const C<T>* gby = gbx + N;
  const C<T>* gx = Xin + (size_t)v * xs;
  const C<T>* gy = gx + N;
  const C<T>* cxp = PROL ? XCin + (size_t)v * cs : nullptr;
  const C<T>* cyp = PROL ? cxp + NC : nullptr;
  C<T>* outx = Xout + (size_t)v * xs;
  C<T>* outy = outx + N;
  const T axy = T(1.0 / (lv.hx * lv.hy));
  const T mx = T(lv.hy / lv.hx / 6.0), my = T(lv.hx / lv.hy / 6.0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int sweep = wave >> 1, lead = !(wave & 1);
  const int c = tid & (MARCH_W - 1);  // column for row loads, write-out, residuals
  const int half = wave >> 1;         // NSW = 2: the row of a pair this thread's wave handles
  const bool cin = c < RW;
  const int ciA = BWD, ciB = 1 - BWD;
  const int pcA = 2 * lane + ciA, pcB = 2 * lane + ciB;
  auto validc = [&](int pc) { return pc >= 1 && pc <= RW - 2; };
  auto validr = [&](int pr) { return pr >= 1 && pr <= RH - 2; };
  // wrapped global column of region column lc (clamped into the region), and
  // wrapped global row of region row lr (wave-uniform)
  auto wcol = [&](int lc) {
    lc = lc < 0 ? 0 : (lc >= RW ? RW - 1 : lc);
    const int gi = mg.ox + lc;
    return gi < 0 ? gi + nx : (gi >= nx ? gi - nx : gi);
  };
  auto wrow = [&](int lr) {
    lr = lr < 0 ? 0 : (lr >= RH ? RH - 1 : lr);
    const int gj = mg.oy + lr;
    return (gj < 0 ? gj + ny : (gj >= ny ? gj - ny : gj));
  };
  // whether values of region row lr need a Bloch phase (lr is wave-uniform
  // at every use: one scalar test per row, so rows of a region inside the
  // grid's columns skip the per-lane test of mg.fix)
  const bool cwrap = mg.ox < 0 || mg.ox + RW > nx;
  auto needfix = [&](int lr) { return !inter && (cwrap || (unsigned)(mg.oy + lr) >= (unsigned)ny); };
  // Bloch phases of the thread's column c and of its lane's patch columns
  // 2 lane, 2 lane + 1 are formed once (colph) and applied per row (fixc):
  // the FP32 residual pass that reads x went from 4.7 to 4.45 ms.  The FP32
  // zero-start residual pass keeps the per-value form (fix): with the
  // precomputed phases its loop compiled slower (4.1 -> 4.5 ms)
  constexpr bool PRECOL = MODE != 10 || sizeof(T) == 8;
  const C<T> phc = mg.colph(c), ph0 = mg.colph(2 * lane), ph1 = mg.colph(2 * lane + 1);
  auto fixcol = [&](C<T> v, C<T> cph, int lc, int lr) { return PRECOL ? mg.fixc(v, cph, lr) : mg.fix(v, lc, lr); };
  const int gC = wcol(c);

  auto load_row = [&](int r, MarchRow<T>& R) {
    const int wj = wrow(r);
    const size_t k = gC + (size_t)wj * nx;
    R.e = eta[k];
    if (!ZERO) {
      R.x = gx[k];
      R.y = gy[k];
    }
  };
  // coarse row lr of the region (global row oy / 2 + lr, wrapped; rows past
  // the region clamped like wrow's), column tid (< NCW) of it
  const int ocx = mg.ox >> 1, ocy = mg.oy >> 1;  // the region's origin is even
  // (a true modulus: where nx is close to the halo, column RW / 2 of a
  // region lies a whole period past the grid, and the columns tid > RW / 2
  // that are staged but never read lie further)
  auto ccol = [&](int li) {
    const int gi = (ocx + li) % cnx;
    return gi < 0 ? gi + cnx : gi;
  };
  auto crow = [&](int lr) {
    const int last = (RH - 1) >> 1;
    lr = lr > last + 1 ? last + 1 : lr;
    const int gj = (ocy + lr) % cny;
    return gj < 0 ? gj + cny : gj;
  };
  const int gcc = PROL && tid < NCW ? ccol(tid) : 0;
  auto store_row = [&](int r, const MarchRow<T>& R) {
    C<T> vx = czero<T>(), vy = czero<T>();
    if (!ZERO) {
      vx = R.x;
      vy = R.y;
      if (PROL) {
        const int wj = wrow(r), I = gC >> 1, J = wj >> 1;
        // the raw coarse values a direct load reads: coarse column c / 2 (+1)
        // and row r / 2 (+1) of the region, wrapped
        const int lI = c >> 1, s0 = ((r >> 1) & 3) * NCW, s1 = (((r >> 1) + 1) & 3) * NCW;
        const C<T> c0 = scx[s0 + lI], d0 = scy[s0 + lI];
        if (wj & 1) {
          const C<T> c1 = scx[s1 + lI];
          vx += T(0.25) * (c0 + (J + 1 == cny ? g.yf * c1 : c1));
        } else {
          vx += T(0.5) * c0;
        }
        if (gC & 1) {
          const C<T> d1 = scy[s0 + lI + 1];
          vy += T(0.25) * (d0 + (I + 1 == cnx ? g.xf * d1 : d1));
        } else {
          vy += T(0.5) * d0;
        }
      }
      if (needfix(r)) {
        vx = fixcol(vx, phc, c, r);
        vy = fixcol(vy, phc, c, r);
      }
    }
    const int k = msl<RING>(r, c);
    sx[k] = vx;
    sy[k] = vy;
    se[k] = axy * R.e;
  };

  // row pairs are loaded one step ahead: NSW = 1: rows r0, r0 + 1 by every
  // thread; NSW = 2: row r0 + half
  MarchRow<T> R0, R1;
  C<T> cpx = czero<T>(), cpy = czero<T>();  // coarse row staged at the end of the step
  if (PROL) {
    if (tid < NCW) {
      scx[tid] = cxp[gcc + (size_t)crow(0) * cnx];
      scy[tid] = cyp[gcc + (size_t)crow(0) * cnx];
      scx[NCW + tid] = cxp[gcc + (size_t)crow(1) * cnx];
      scy[NCW + tid] = cyp[gcc + (size_t)crow(1) * cnx];
      cpx = cxp[gcc + (size_t)crow(2) * cnx];
      cpy = cyp[gcc + (size_t)crow(2) * cnx];
    }
    __syncthreads();
  }
  if (BWD) {
    if (NSW == 1 || half == 0) {
      load_row(0, R0);
      if (cin) store_row(0, R0);
    }
  }
  if (NSW == 1) {
    load_row(BWD, R0);
    load_row(BWD + 1, R1);
  } else {
    load_row(BWD + half, R0);
  }
  const int NS = RH / 2 + 2 + 3 * (NSW - 1);
  // Raw right-hand sides.  Lane l of a wave holds, for its wave's patch row
  // pr, bx and by at columns 2l and 2l + 1 and by at those columns of row
  // pr - 1 (each value loaded once, coalesced): a patch at column pc = 2l
  // takes bx(pc - 1) from lane l - 1 by a cross-lane shuffle, a patch at
  // pc = 2l + 1 has all four values.  Groups are reloaded for the next step
  // right after their last use (the by pair of phase A's parity after phase
  // A, the rest after phase B), so a load has most of a step to land and no
  // register is copied while a load into it is in flight.  Loads are
  // unconditional (clamped addresses): straight-line code lets the compiler
  // count outstanding loads exactly.  nb[8..11]: residual rows.
  C<T> nb[12];  // 0 bx(2l) 1 bx(2l+1) 2 by(2l) 3 by(2l+1) 4 by(2l,pr-1) 5 by(2l+1,pr-1)
  auto prow = [&](int s2) { return 2 * s2 + BWD - 6 * sweep - (lead ? 0 : 3); };
  auto frow = [&](int s2) { return 2 * s2 - 4 + BWD - 6 * (NSW - 1); };
  const int gE = wcol(2 * lane), gO = wcol(2 * lane + 1);
  auto load_by = [&](int s2, int parity) {  // the by pair of one column parity
    const size_t r0o = (size_t)wrow(prow(s2)) * nx, rmo = (size_t)wrow(prow(s2) - 1) * nx;
    const int gc = parity ? gO : gE;
    nb[2 + parity] = gby[gc + r0o];
    nb[4 + parity] = gby[gc + rmo];
  };
  auto load_bx = [&](int s2) {
    const size_t r0o = (size_t)wrow(prow(s2)) * nx;
    nb[0] = gbx[gE + r0o];
    nb[1] = gbx[gO + r0o];
  };
  // (NSW = 2: the wave's one row, f0 + half, in nb[8..9])
  auto load_bR = [&](int s2) {
    if (RES) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (NSW == 2 && q != half) continue;
        const int sl = NSW == 2 ? 0 : q;
        const size_t ro = (size_t)wrow(frow(s2) + q) * nx;
        nb[8 + 2 * sl] = gbx[gC + ro];
        nb[9 + 2 * sl] = gby[gC + ro];
      }
    }
  };
  auto shfl_up1 = [](C<T> v) {
    C<T> r;
    r.r = __shfl_up(v.r, 1, 64);
    r.i = __shfl_up(v.i, 1, 64);
    return r;
  };
  load_bx(0);
  load_by(0, 0);
  load_by(0, 1);
  load_bR(0);
  // restriction of the residual rows of one step (coarse row f0 / 2); run
  // at the top of the next step, after its barrier, so that it needs no
  // barrier of its own and overlaps the other waves' row loads.  FP32: by
  // the first wave; FP64: the x-edges by the first wave, the y-edges by the
  // second (8.9 -> 8.4-8.6 ms for mode 8, 9.1 -> 8.9 ms for mode 10; the
  // same split measured slower in FP32, 4.15 -> 4.6 ms for mode 10)
  auto restrict_row = [&](int f0) {
    if constexpr (sizeof(T) == 4) {
      const int J = f0 / 2, I = tid;  // region-local coarse indices (f0 even)
      const int fi = 2 * I, fj = 2 * J;
      if (f0 >= 0 && tid < MARCH_W / 2 && I >= H / 2 && I < (H + W) / 2 && J >= H / 2 && J < (H + Hgt) / 2) {
        const int GI = mg.ox / 2 + I, GJ = mg.oy / 2 + J;
        if (GI < cnx && GJ < cny) {
          C<T> cxv = T(0.5) * (rxs[rsl<RRES>(fj, fi)] + rxs[rsl<RRES>(fj, fi + 1)]) +
                     T(0.25) * (rxs[rsl<RRES>(fj - 1, fi)] + rxs[rsl<RRES>(fj - 1, fi + 1)] +
                                rxs[rsl<RRES>(fj + 1, fi)] + rxs[rsl<RRES>(fj + 1, fi + 1)]);
          C<T> cyv = T(0.5) * (rys[rsl<RRES>(fj, fi)] + rys[rsl<RRES>(fj + 1, fi)]) +
                     T(0.25) * (rys[rsl<RRES>(fj, fi - 1)] + rys[rsl<RRES>(fj + 1, fi - 1)] +
                                rys[rsl<RRES>(fj, fi + 1)] + rys[rsl<RRES>(fj + 1, fi + 1)]);
          C<T>* bc = BC + (size_t)v * cs;
          C<T>* xc = XC + (size_t)v * cs;
          const size_t k = GI + (size_t)GJ * cnx;
          bc[k] = cxv;
          bc[NC + k] = cyv;
          xc[k] = czero<T>();
          xc[NC + k] = czero<T>();
        }
      }
    } else {
      const int J = f0 / 2, I = lane;
      const int fi = 2 * I, fj = 2 * J;
      if (f0 >= 0 && wave < 2 && I >= H / 2 && I < (H + W) / 2 && J >= H / 2 && J < (H + Hgt) / 2) {
        const int GI = mg.ox / 2 + I, GJ = mg.oy / 2 + J;
        if (GI < cnx && GJ < cny) {
          const size_t k = GI + (size_t)GJ * cnx + (wave ? NC : 0);
          C<T> cv;
          if (wave == 0)
            cv = T(0.5) * (rxs[rsl<RRES>(fj, fi)] + rxs[rsl<RRES>(fj, fi + 1)]) +
                 T(0.25) * (rxs[rsl<RRES>(fj - 1, fi)] + rxs[rsl<RRES>(fj - 1, fi + 1)] +
                            rxs[rsl<RRES>(fj + 1, fi)] + rxs[rsl<RRES>(fj + 1, fi + 1)]);
          else
            cv = T(0.5) * (rys[rsl<RRES>(fj, fi)] + rys[rsl<RRES>(fj + 1, fi)]) +
                 T(0.25) * (rys[rsl<RRES>(fj, fi - 1)] + rys[rsl<RRES>(fj + 1, fi - 1)] +
                            rys[rsl<RRES>(fj, fi + 1)] + rys[rsl<RRES>(fj + 1, fi + 1)]);
          BC[(size_t)v * cs + k] = cv;
          XC[(size_t)v * cs + k] = czero<T>();
        }
      }
    }
  };
  for (int s = 0; s < NS; s++) {
    const int r0 = 2 * s + BWD;
    const int pr = prow(s);
    const bool rv = validr(pr);
    const int f0 = frow(s);
    __syncthreads();
    if (RES && s > 0) restrict_row(frow(s - 1));
    if (NSW == 1) {
      if (cin && r0 < RH) store_row(r0, R0);
      if (cin && r0 + 1 < RH) store_row(r0 + 1, R1);
      load_row(r0 + 2, R0);
      load_row(r0 + 3, R1);
    } else {
      if (cin && r0 + half < RH) store_row(r0 + half, R0);
      load_row(r0 + 2 + half, R0);
    }
    __syncthreads();
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
      const int pc = ph == 0 ? pcA : pcB;
      const int ci = ph == 0 ? ciA : ciB;
      // Bloch phases of this phase's right-hand sides, applied in place once
      // they are needed (the bx pair in phase A, the phase's by pair); the
      // shuffle below then moves a value already phased for its column
      if (needfix(pr) || needfix(pr - 1)) {
        if (ph == 0) {
          nb[0] = fixcol(nb[0], ph0, 2 * lane, pr);
          nb[1] = fixcol(nb[1], ph1, 2 * lane + 1, pr);
        }
        nb[2 + ci] = fixcol(nb[2 + ci], ci ? ph1 : ph0, pc, pr);
        nb[4 + ci] = fixcol(nb[4 + ci], ci ? ph1 : ph0, pc, pr - 1);
      }
      // bx(pc - 1): for an even column the odd value of lane l - 1 (all lanes
      // take part in the shuffle, outside the patch guard)
      const C<T> bl = ci ? nb[0] : shfl_up1(nb[1]);
      if (rv && validc(pc)) {
        const C<T> b0 = bl, b1 = ci ? nb[1] : nb[0];
        const C<T> b2 = nb[4 + ci], b3 = nb[2 + ci];
        const int rm = rrow<RING>(pr - 1) * MARCH_W, r0_ = rrow<RING>(pr) * MARCH_W;
        const int rp = rrow<RING>(pr + 1) * MARCH_W;
        const int cm = mcol(pc - 1), c0 = mcol(pc), cp = mcol(pc + 1);
        const T eLL = se[rm + cm], eRL = se[rm + c0];
        const T eLU = se[r0_ + cm], eRU = se[r0_ + c0];
        const int iL0 = r0_ + cm, iR0 = r0_ + c0, iCm = rm + c0, iC0 = iR0;
        C<T> xLm = sx[rm + cm], xL0 = sx[iL0], xLp = sx[rp + cm];
        C<T> xRm = sx[iCm], xR0 = sx[iR0], xRp = sx[rp + c0];
        C<T> yLm = sy[rm + cm], yCm = sy[iCm], yRm = sy[rm + cp];
        C<T> yL0 = sy[iL0], yC0 = sy[iC0], yR0 = sy[r0_ + cp];
        C<T> qLL = eLL * (xLm - xL0 - yLm + yCm);
        C<T> qRL = eRL * (xRm - xR0 - yCm + yRm);
        C<T> qLU = eLU * (xL0 - xLp - yL0 + yC0);
        C<T> qRU = eRU * (xR0 - xRp - yC0 + yR0);
        C<T> s0 = b0 - (qLU - qLL) - mx * (T(4) * xL0 + xLm + xLp);
        C<T> s1 = b1 - (qRU - qRL) - mx * (T(4) * xR0 + xRm + xRp);
        C<T> s2 = b2 - (qLL - qRL) - my * (T(4) * yCm + yLm + yRm);
        C<T> s3 = b3 - (qLU - qRU) - my * (T(4) * yC0 + yL0 + yR0);
        C<T> z0, z1, z2, z3;
        afw_local_solve(eLL, eRL, eLU, eRU, mx, my, s0, s1, s2, s3, z0, z1, z2, z3);
        sx[iL0] = xL0 + z0;
        sx[iR0] = xR0 + z1;
        sy[iCm] = yCm + z2;
        sy[iC0] = yC0 + z3;
      }
      if (s + 1 < NS) {
        if (ph == 0) {
          load_by(s + 1, ciA);
        } else {
          load_bx(s + 1);
          load_by(s + 1, ciB);
        }
      }
      // No barrier between the two phases: phase B reads what phase A wrote
      // in the same wave (LDS operations of a wave complete in order), and
      // within a step the four waves' patch rows pr, pr - 3, pr - 6, pr - 9
      // write rows {pr - 1, pr} and read rows pr - 1 .. pr + 1 of their own,
      // which no other wave of the step writes.  The barrier after phase B
      // orders the last sweep's rows before the final rows read them.
      if (ph == 1) __syncthreads();
    }
    // final rows: write the owned part, form residuals
    if (cin) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (NSW == 2 && q != half) continue;
        const int f = f0 + q;
        if (f < 0 || f >= RH) continue;
        const int gi = mg.ox + c, gj = mg.oy + f;
        const int k0 = msl<RING>(f, c);
        if (c >= H && c < H + W && f >= H && f < H + Hgt && gi < nx && gj < ny) {
          outx[gi + (size_t)gj * nx] = sx[k0];
          outy[gi + (size_t)gj * nx] = sy[k0];
        }
        if (RES) {
          C<T> rx = czero<T>(), ry = czero<T>();
          if (validc(c) && validr(f)) {
            const int fm = rrow<RING>(f - 1) * MARCH_W, fr = rrow<RING>(f) * MARCH_W;
            const int fp = rrow<RING>(f + 1) * MARCH_W;
            const int cm = mcol(c - 1), c0 = mcol(c), cp = mcol(c + 1);
            C<T> x00 = sx[k0], x0m = sx[fm + c0], x0p = sx[fp + c0];
            C<T> xm0 = sx[fr + cm], xmp = sx[fp + cm];
            C<T> y00 = sy[k0], ym0 = sy[fr + cm], yp0 = sy[fr + cp];
            C<T> y0m = sy[fm + c0], ypm = sy[fm + cp];
            C<T> c00 = x00 - x0p - y00 + yp0, c0m = x0m - x00 - y0m + ypm, cm0 = xm0 - xmp - ym0 + y00;
            const T e00 = se[k0], e0m = se[fm + c0], em0 = se[fr + cm];
            C<T> kx = e00 * c00 - e0m * c0m;
            C<T> ky = em0 * cm0 - e00 * c00;
            const int sl = NSW == 2 ? 0 : q;
            C<T> bxr = nb[8 + 2 * sl], byr = nb[9 + 2 * sl];
            if (needfix(f)) {
              bxr = fixcol(bxr, phc, c, f);
              byr = fixcol(byr, phc, c, f);
            }
            rx = bxr - kx - mx * (T(4) * x00 + x0m + x0p);
            ry = byr - ky - my * (T(4) * y00 + ym0 + yp0);
          }
          rxs[rsl<RRES>(f, c)] = rx;
          rys[rsl<RRES>(f, c)] = ry;
        }
      }
    }
    if (s + 1 < NS) load_bR(s + 1);
    if (PROL && tid < NCW) {
      // coarse row s + 2 (needed from the next step on; its slot was last
      // read two steps ago), then row s + 3 in flight for the step after
      scx[((s + 2) & 3) * NCW + tid] = cpx;
      scy[((s + 2) & 3) * NCW + tid] = cpy;
      const size_t ro = (size_t)crow(s + 3) * cnx;
      cpx = cxp[gcc + ro];
      cpy = cyp[gcc + ro];
    }
  }
  if (RES) {
    __syncthreads();
    restrict_row(frow(NS - 1));
  }
}

// The nodal counterpart of k_afw_march: pointwise Gauss-Seidel for the Q1
// Laplacian in four colours, the same row pipeline (a node of colour 2/3 on
// odd row r needs colours 0,1 done on rows r-1 and r+1, and colour 0/1 on
// row r+1 reads row r before colour 2/3 touches it) and the same trailing of
// further sweeps, with a ring of x.
template <typename T, int MODE, int NSW>
__global__ __launch_bounds__(MarchCfg<NSW>::THREADS) void k_node_march(mxh_level lv, mxh_bloch bl, const C<T>* B,
                                                                      long long bs, const C<T>* Xin, C<T>* Xout,
                                                                      long long xs, const C<T>* XCin, C<T>* BC,
                                                                      C<T>* XC, long long cs, int ntx, int W, int Hgt)
{
  // two sweeps: a 16-row ring in both precisions (FP64 too: its 42 KB of
  // LDS at three workgroups per CU with two barriers per step beat 36 KB at
  // four with three, 3.35 -> 3.13 ms for the residual pass)
  constexpr int H = MarchCfg<NSW>::H, RING = NSW == 2 ? 16 : MarchCfg<NSW>::template ring<T>();
  constexpr int RRES = MarchCfg<NSW>::template rres<T>();
  constexpr int BWD = MODE & 1;
  constexpr bool ZERO = MODE & 2, PROL = MODE & 4, RES = MODE & 8;
  constexpr int NRR = RES ? RRES * MARCH_W : 1;
  // FP32: the right-hand side rides in the ring with x (loaded once per pass
  // for both sweeps and the residual); FP64 keeps it in registers (LDS)
  constexpr bool BRING = sizeof(T) == 4;
  __shared__ C<T> sx[RING * MARCH_W];
  __shared__ C<T> sb[BRING ? RING * MARCH_W : 1];
  __shared__ C<T> rs[NRR];
  // MODE 4: coarse rows of the prolongation, 4-row ring of MARCH_W / 2 + 1
  // coarse columns: region row r needs coarse rows r / 2 and r / 2 + 1
  // (region-local; the region's origin is even), staged one step ahead so
  // that a fine node reads its four coarse values from LDS
  constexpr int NCW = MARCH_W / 2 + 1;
  __shared__ C<T> scx[PROL ? 4 * NCW : 1];
  // Two barriers per step instead of three where the ring is long enough:
  // the rows a step stores (r0 + half, r0 + 1 + half) then reuse slots of
  // rows RING behind, below every row the previous step's final rows read
  // (f0 - 1 = r0 - 13 for two sweeps), so the step's stores need not wait
  // for the previous step to end.  What the barrier at the top also ordered
  // moves behind the others: the restriction of the previous step's
  // residual rows runs after the stores' barrier, and the next coarse row
  // of the prolongation is staged before the final rows' barrier (the
  // stores of the next step read it; this step's read rows s and s + 1).
  constexpr bool NOTOP = NSW == 2 && RING >= 14;
  const int v = blockIdx.y;
  const int nx = lv.nx, ny = lv.ny, cnx = nx >> 1, cny = ny >> 1;
  const int tileX = blockIdx.x % ntx, tileY = blockIdx.x / ntx;
  const int RW = W + 2 * H, RH = Hgt + 2 * H;
  const Geo<T> g = make_geo<T>(lv, bl, v);
  const MarchGeo<T> mg{nx, ny, tileX * W - H, tileY * Hgt - H, g.xf, g.yf};
  const bool inter = mg.ox >= 0 && mg.ox + RW <= nx && mg.oy >= 0 && mg.oy + RH <= ny;
  const C<T>* gb = B + (size_t)v * bs;
  const C<T>* gx = Xin + (size_t)v * xs;
  const C<T>* xcp = PROL ? XCin + (size_t)v * cs : nullptr;
  C<T>* out = Xout + (size_t)v * xs;
  const NodeCoef<T> nc = node_coef<T>(lv, 1.0);
  const T iac = T(1) / nc.ac;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int sweep = wave >> 1, lead = !(wave & 1);
  const int c = tid & (MARCH_W - 1);
  const int half = wave >> 1;
  const bool cin = c < RW;
  const int ciA = BWD, ciB = 1 - BWD;
  const int pcA = 2 * lane + ciA, pcB = 2 * lane + ciB;
  auto wcol = [&](int lc) {
    lc = lc < 0 ? 0 : (lc >= RW ? RW - 1 : lc);
    const int gi = mg.ox + lc;
    return gi < 0 ? gi + nx : (gi >= nx ? gi - nx : gi);
  };
  auto wrow = [&](int lr) {
    lr = lr < 0 ? 0 : (lr >= RH ? RH - 1 : lr);
    const int gj = mg.oy + lr;
    return (gj < 0 ? gj + ny : (gj >= ny ? gj - ny : gj));
  };
  // whether values of region row lr need a Bloch phase (lr is wave-uniform
  // at every use: one scalar test per row, so rows of a region inside the
  // grid's columns skip the per-lane test of mg.fix)
  const bool cwrap = mg.ox < 0 || mg.ox + RW > nx;
  auto needfix = [&](int lr) { return !inter && (cwrap || (unsigned)(mg.oy + lr) >= (unsigned)ny); };
  const int gA = wcol(pcA), gB = wcol(pcB), gC = wcol(c);

  // raw loads of a row: x and (FP32) b
  struct NRow {
    C<T> x, b;
  };
  auto load_row = [&](int r, NRow& R) {
    const int wj = wrow(r);
    if (BRING) R.b = gb[gC + (size_t)wj * nx];
    if (!ZERO) R.x = gx[gC + (size_t)wj * nx];
  };
  // coarse row lr of the region (global row oy / 2 + lr, wrapped; rows past
  // the region are clamped like wrow's), column tid (< NCW) of it
  const int ocx = mg.ox >> 1, ocy = mg.oy >> 1;  // the region's origin is even
  // (a true modulus: where nx is close to the halo, column RW / 2 of a
  // region lies a whole period past the grid, and the columns tid > RW / 2
  // that are staged but never read lie further)
  auto ccol = [&](int li) {
    const int gi = (ocx + li) % cnx;
    return gi < 0 ? gi + cnx : gi;
  };
  auto crow = [&](int lr) {
    const int last = (RH - 1) >> 1;
    lr = lr > last + 1 ? last + 1 : lr;
    const int gj = (ocy + lr) % cny;
    return gj < 0 ? gj + cny : gj;
  };
  const int gcc = PROL && tid < NCW ? ccol(tid) : 0;
  auto load_crow = [&](int lr) { return xcp[gcc + (size_t)crow(lr) * cnx]; };
  auto finish_row = [&](int r, const NRow& R) {
    C<T> vx = czero<T>();
    if (!ZERO) {
      vx = R.x;
      if (PROL) {
        const int wj = wrow(r), wi = gC;
        const int I = wi >> 1, J = wj >> 1;
        const bool ox1 = I + 1 == cnx, oy1 = J + 1 == cny;
        const T wx1 = (wi & 1) ? T(0.5) : T(0), wx0 = (wi & 1) ? T(0.5) : T(1);
        const T wy1 = (wj & 1) ? T(0.5) : T(0), wy0 = (wj & 1) ? T(0.5) : T(1);
        // the same raw coarse values a direct load reads: coarse column c / 2
        // (+1) and row r / 2 (+1) of the region, wrapped
        const int lI = c >> 1, s0 = ((r >> 1) & 3) * NCW, s1 = (((r >> 1) + 1) & 3) * NCW;
        const C<T> p00 = scx[s0 + lI], p10 = scx[s0 + lI + 1];
        const C<T> p01 = (wj & 1) ? scx[s1 + lI] : czero<T>(), p11 = (wj & 1) ? scx[s1 + lI + 1] : czero<T>();
        C<T> p = (wx0 * wy0) * p00;
        if (wi & 1) p += (wx1 * wy0) * (ox1 ? g.xf * p10 : p10);
        if (wj & 1) p += (wx0 * wy1) * (oy1 ? g.yf * p01 : p01);
        if ((wi & 1) && (wj & 1)) {
          C<T> q = p11;
          if (ox1) q = g.xf * q;
          if (oy1) q = g.yf * q;
          p += (wx1 * wy1) * q;
        }
        vx += p;
      }
      if (needfix(r)) vx = mg.fix(vx, c, r);
    }
    sx[msl<RING>(r, c)] = vx;
    if (BRING) sb[msl<RING>(r, c)] = needfix(r) ? mg.fix(R.b, c, r) : R.b;
  };
  auto validc = [&](int pc) { return pc >= 1 && pc <= RW - 2; };
  auto validr = [&](int pr) { return pr >= 1 && pr <= RH - 2; };
  auto offd = [&](int pc, int pr) {
    const int rm = rrow<RING>(pr - 1) * MARCH_W, r0_ = rrow<RING>(pr) * MARCH_W;
    const int rp = rrow<RING>(pr + 1) * MARCH_W;
    const int cm = mcol(pc - 1), c0 = mcol(pc), cp = mcol(pc + 1);
    return nc.ax * (sx[r0_ + cm] + sx[r0_ + cp]) + nc.ay * (sx[rm + c0] + sx[rp + c0]) +
           nc.ad * (sx[rm + cm] + sx[rm + cp] + sx[rp + cm] + sx[rp + cp]);
  };

  NRow R0, R1;
  C<T> cpre = czero<T>();  // coarse row staged at the end of the step
  if (PROL) {
    if (tid < NCW) {
      scx[tid] = load_crow(0);
      scx[NCW + tid] = load_crow(1);
      cpre = load_crow(2);
    }
    __syncthreads();
  }
  if (BWD) {
    if (NSW == 1 || half == 0) {
      load_row(0, R0);
      if (cin) finish_row(0, R0);
    }
  }
  // FP32, NSW = 2: the wave's row of a step is loaded two steps ahead, into
  // R0 for even steps and R1 for odd ones (the loop below is unrolled by two,
  // so each register set is named statically and the load counts stay
  // exact): the FP32 passes 4-7% faster.  FP64 loads one step ahead (two
  // measured 15% slower there).
  constexpr bool PF2 = NSW == 2 && sizeof(T) == 4;
  if (NSW == 1) {
    load_row(BWD, R0);
    load_row(BWD + 1, R1);
  } else {
    load_row(BWD + half, R0);
    if (PF2) load_row(BWD + 2 + half, R1);
  }
  const int NS = RH / 2 + 2 + 3 * (NSW - 1);
  // raw right-hand sides (patch A, patch B, residual rows), each reloaded for
  // the next step right after its last use
  C<T> nb[4];
  auto prow = [&](int s2) { return 2 * s2 + BWD - 6 * sweep - (lead ? 0 : 3); };
  auto frow = [&](int s2) { return 2 * s2 - 4 + BWD - 6 * (NSW - 1); };
  auto load_bA = [&](int s2) {
    if (!BRING) nb[0] = gb[gA + (size_t)wrow(prow(s2)) * nx];
  };
  auto load_bB = [&](int s2) {
    if (!BRING) nb[1] = gb[gB + (size_t)wrow(prow(s2)) * nx];
  };
  auto bval = [&](int slot, int pc, int prr) {
    return BRING ? sb[msl<RING>(prr, pc)] : (needfix(prr) ? mg.fix(nb[slot], pc, prr) : nb[slot]);
  };
  auto load_bR = [&](int s2) {
    if (RES && !BRING) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (NSW == 2 && q != half) continue;
        nb[NSW == 2 ? 2 : 2 + q] = gb[gC + (size_t)wrow(frow(s2) + q) * nx];
      }
    }
  };
  load_bA(0);
  load_bB(0);
  load_bR(0);
  // restriction of the residual rows of one step (coarse row f0 / 2), run at
  // the top of the next step after its barrier (see k_afw_march)
  auto restrict_row = [&](int f0) {
    const int J = f0 / 2, I = tid;
    const int fi = 2 * I, fj = 2 * J;
    if (f0 >= 0 && tid < MARCH_W / 2 && I >= H / 2 && I < (H + W) / 2 && J >= H / 2 && J < (H + Hgt) / 2) {
      const int GI = mg.ox / 2 + I, GJ = mg.oy / 2 + J;
      if (GI < cnx && GJ < cny) {
        C<T> acc = czero<T>();
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
          for (int di = -1; di <= 1; di++) {
            const T w = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5));
            acc += w * rs[rsl<RRES>(fj + dj, fi + di)];
          }
        const size_t k = GI + (size_t)GJ * cnx;
        BC[(size_t)v * cs + k] = acc;
        XC[(size_t)v * cs + k] = czero<T>();
      }
    }
  };
  auto step = [&](int s, NRow& Rc) {
    const int r0 = 2 * s + BWD;
    const int pr = prow(s);
    const bool rv = validr(pr);
    const bool vA = rv && validc(pcA), vB = rv && validc(pcB);
    const int f0 = frow(s);
    if (!NOTOP) {
      __syncthreads();
      if (RES && s > 0) restrict_row(frow(s - 1));
    }
    if (NSW == 1) {
      if (cin && r0 < RH) finish_row(r0, R0);
      if (cin && r0 + 1 < RH) finish_row(r0 + 1, R1);
      load_row(r0 + 2, R0);
      load_row(r0 + 3, R1);
    } else {
      if (cin && r0 + half < RH) finish_row(r0 + half, Rc);
      load_row(r0 + (PF2 ? 4 : 2) + half, Rc);
    }
    __syncthreads();
    if (NOTOP && RES && s > 0) restrict_row(frow(s - 1));
    // colours A and B of a row in one wave with no barrier between them (as
    // in k_afw_march: a wave's rows pr - 1 .. pr + 1 are written by no other
    // wave within the step)
    if (vA) sx[msl<RING>(pr, pcA)] = iac * (bval(0, pcA, pr) - offd(pcA, pr));
    if (s + 1 < NS) load_bA(s + 1);
    if (vB) sx[msl<RING>(pr, pcB)] = iac * (bval(1, pcB, pr) - offd(pcB, pr));
    if (s + 1 < NS) load_bB(s + 1);
    if (NOTOP && PROL && tid < NCW) {
      scx[((s + 2) & 3) * NCW + tid] = cpre;
      cpre = load_crow(s + 3);
    }
    __syncthreads();
    if (cin) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (NSW == 2 && q != half) continue;
        const int f = f0 + q;
        if (f < 0 || f >= RH) continue;
        const int gi = mg.ox + c, gj = mg.oy + f;
        if (c >= H && c < H + W && f >= H && f < H + Hgt && gi < nx && gj < ny)
          out[gi + (size_t)gj * nx] = sx[msl<RING>(f, c)];
        if (RES) {
          C<T> r = czero<T>();
          if (validc(c) && validr(f)) r = bval(NSW == 2 ? 2 : 2 + q, c, f) - nc.ac * sx[msl<RING>(f, c)] - offd(c, f);
          rs[rsl<RRES>(f, c)] = r;
        }
      }
    }
    if (s + 1 < NS) load_bR(s + 1);
    if (!NOTOP && PROL && tid < NCW) {
      // coarse row s + 2 (needed from the next step on; its slot was last
      // read two steps ago), then row s + 3 in flight for the step after
      scx[((s + 2) & 3) * NCW + tid] = cpre;
      cpre = load_crow(s + 3);
    }
  };
  if (PF2) {
    int s = 0;
    for (; s + 1 < NS; s += 2) {
      step(s, R0);
      step(s + 1, R1);
    }
    if (s < NS) step(s, R0);
  } else {
    for (int s = 0; s < NS; s++) step(s, R0);
  }
  if (RES) {
    __syncthreads();
    restrict_row(frow(NS - 1));
  }
}

// ================================================================= launchers
#define LAUNCH_CHECK() do { hipError_t _e = hipGetLastError(); if (_e != hipSuccess) return (int)_e; } while (0)

extern "C" {

int mxh_init(void)
{
  if (g_stream) return 0;
  int n = 0, dev = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return (int)e;
  if (n < 1) return (int)hipErrorNoDevice;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= MXH_MAX_DEVICES) return (int)hipErrorInvalidDevice;
  // the device address of this device's copy of the static scratch array
  // (the array's name is a host-side stand-in, not a device address)
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = p;
  }
  e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_stream = nullptr;
    return (int)e;
  }
  g_device = dev;
  return 0;
}

void* mxh_stream(void) { return (void*)g_stream; }

int mxh_sync(void) { return (int)hipStreamSynchronize(g_stream); }

const char* mxh_error_string(int err)
{
  if (err == MXH_ESCRATCH) return "request exceeds the static scratch of mxh_gram_static";
  return hipGetErrorString((hipError_t)err);
}

void* mxh_malloc(size_t bytes)
{
  void* p = nullptr;
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void mxh_free(void* p)
{
  if (p) (void)hipFree(p);
}

int mxh_h2d(void* dst, const void* src, size_t bytes)
{
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream);
}

int mxh_d2h(void* dst, const void* src, size_t bytes)
{
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream);
  if (e != hipSuccess) return (int)e;
  return (int)hipStreamSynchronize(g_stream);
}

int mxh_d2h_async(void* dst, const void* src, size_t bytes)
{
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream);
}

void* mxh_host_alloc(size_t bytes)
{
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 256, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void mxh_host_free(void* p)
{
  if (p) (void)hipHostFree(p);
}

int mxh_event_sync(void* ev) { return (int)hipEventSynchronize((hipEvent_t)ev); }

int mxh_d2d(void* dst, const void* src, size_t bytes)
{
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream);
}

int mxh_memset(void* dst, int value, size_t bytes)
{
  return (int)hipMemsetAsync(dst, value, bytes, g_stream);
}



int mxh_addeval(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs x, mxh_vecs y,
                const void* eta, const mxh_coef* cf, int accumulate)
{
  const long long N = (long long)lv->nx * lv->ny;
  dim3 grid(nblocks(N), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_addeval<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)x.p,
                       x.stride, (C<double>*)y.p, y.stride, (const double*)eta, *cf, accumulate);
  else
    hipLaunchKernelGGL(k_addeval<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)x.p,
                       x.stride, (C<float>*)y.p, y.stride, (const float*)eta, *cf, accumulate);
  LAUNCH_CHECK();
  return 0;
}

int mxh_defect(int rprec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs e, mxh_vecs r,
               const double* eta64, const double* lam, double* out, double* work)
{
  const unsigned nb = row_blocks(lv->ny), nf = frame_blocks(lv);
  const dim3 grid(nb, nv), gridf(nf, nv);
  if (rprec == MXH_F64) {
    hipLaunchKernelGGL((k_defect<double, false>), grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (C<double>*)r.p, r.stride, eta64, lam, work, 0);
    hipLaunchKernelGGL((k_defect<double, true>), gridf, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (C<double>*)r.p, r.stride, eta64, lam, work, (int)nb);
  } else {
    hipLaunchKernelGGL((k_defect<float, false>), grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (C<float>*)r.p, r.stride, eta64, lam, work, 0);
    hipLaunchKernelGGL((k_defect<float, true>), gridf, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (C<float>*)r.p, r.stride, eta64, lam, work, (int)nb);
  }
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(2 * nv)), dim3(TPB), 0, g_stream, work, (int)(nb + nf),
                     2 * nv, out, 0);
  LAUNCH_CHECK();
  return 0;
}

int mxh_afw_phase(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x,
                  const void* eta, double alpha, double beta, int colour)
{
  if ((lv->nx & 1) || (lv->ny & 1)) return (int)hipErrorInvalidValue;
  const long long np = (long long)(lv->nx / 2) * (lv->ny / 2);
  dim3 grid(nblocks(np), nv);
  const int ci = colour & 1, cj = colour >> 1;
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_afw_phase<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)b.p,
                       b.stride, (C<double>*)x.p, x.stride, (const double*)eta, alpha, beta, ci, cj);
  else
    hipLaunchKernelGGL(k_afw_phase<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)b.p,
                       b.stride, (C<float>*)x.p, x.stride, (const float*)eta, alpha, beta, ci, cj);
  LAUNCH_CHECK();
  return 0;
}

int mxh_edge_gs_phase(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x,
                      const void* eta, const mxh_coef* cf, int colour)
{
  if ((lv->nx & 1) || (lv->ny & 1) || colour < 0 || colour > 3) return (int)hipErrorInvalidValue;
  const long long np = (long long)lv->nx * lv->ny / 2;
  dim3 grid(nblocks(np), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_edge_gs_phase<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)b.p,
                       b.stride, (C<double>*)x.p, x.stride, (const double*)eta, *cf, colour);
  else
    hipLaunchKernelGGL(k_edge_gs_phase<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)b.p,
                       b.stride, (C<float>*)x.p, x.stride, (const float*)eta, *cf, colour);
  LAUNCH_CHECK();
  return 0;
}

extern "C++" {
template <typename T, int MODE>
static void launch_tile(dim3 grid, dim3 block, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b, mxh_vecs x,
                        mxh_vecs xo, const void* eta, int nsweep, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc,
                        long long cs, int ntx)
{
  hipLaunchKernelGGL((k_afw_tile<T, MODE>), grid, block, 0, g_stream, *lv, *bl, (const C<T>*)b.p, b.stride,
                     (const C<T>*)x.p, (C<T>*)xo.p, x.stride, (const T*)eta, nsweep, (const C<T>*)xc_in.p,
                     (C<T>*)bc.p, (C<T>*)xc.p, cs, ntx);
}

template <typename T>
static int launch_tile_mode(int mode, dim3 grid, dim3 block, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b,
                            mxh_vecs x, mxh_vecs xo, const void* eta, int nsweep, mxh_vecs xc_in, mxh_vecs bc,
                            mxh_vecs xc, long long cs, int ntx)
{
  switch (mode) {
  case 0: launch_tile<T, 0>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 2: launch_tile<T, 2>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 8: launch_tile<T, 8>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 10: launch_tile<T, 10>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 1: launch_tile<T, 1>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 5: launch_tile<T, 5>(grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx); break;
  default: return (int)hipErrorInvalidValue;
  }
  return 0;
}
}  // extern "C++"

int mxh_afw_tile(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                 const void* eta, int nsweep, int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc)
{
  if (!xo.p || xo.p == x.p || xo.stride != x.stride) return (int)hipErrorInvalidValue;
  if ((lv->nx & 1) || (lv->ny & 1) || nsweep < 1) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (mode & 1)) return (int)hipErrorInvalidValue;
  // the halo must cover the sweeps (and the residual with its restriction)
  if (nsweep * ((mode & 1) ? TILE_SHRINK_BWD : TILE_SHRINK_FWD) + ((mode & 8) ? TILE_SHRINK_RES : 0) > TILE_H)
    return (int)hipErrorInvalidValue;
  if ((mode & 4) && !xc_in.p) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (!bc.p || !xc.p || bc.stride != xc.stride)) return (int)hipErrorInvalidValue;
  const int R = prec == MXH_F64 ? TileCfg<double>::R : TileCfg<float>::R, TT = R - 2 * TILE_H;
  const int ntx = (lv->nx + TT - 1) / TT, nty = (lv->ny + TT - 1) / TT;
  dim3 grid(ntx * nty, nv), block((R / 2) * (R / 2));
  long long cs = (mode & 8) ? bc.stride : (mode & 4) ? xc_in.stride : 0;
  int e = prec == MXH_F64
              ? launch_tile_mode<double>(mode, grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx)
              : launch_tile_mode<float>(mode, grid, block, lv, bl, b, x, xo, eta, nsweep, xc_in, bc, xc, cs, ntx);
  if (e) return e;
  LAUNCH_CHECK();
  return 0;
}

extern "C++" {
template <typename T, int MODE>
static void launch_ntile(dim3 grid, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                         int nsweep, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, long long cs, int ntx)
{
  hipLaunchKernelGGL((k_node_tile<T, MODE>), grid, dim3(1024), 0, g_stream, *lv, *bl, (const C<T>*)b.p, b.stride,
                     (const C<T>*)x.p, (C<T>*)xo.p, x.stride, nsweep, (const C<T>*)xc_in.p, (C<T>*)bc.p,
                     (C<T>*)xc.p, cs, ntx);
}

template <typename T>
static int launch_ntile_mode(int mode, dim3 grid, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b, mxh_vecs x,
                             mxh_vecs xo, int nsweep, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, long long cs, int ntx)
{
  switch (mode) {
  case 0: launch_ntile<T, 0>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 2: launch_ntile<T, 2>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 8: launch_ntile<T, 8>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 10: launch_ntile<T, 10>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 1: launch_ntile<T, 1>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  case 5: launch_ntile<T, 5>(grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx); break;
  default: return (int)hipErrorInvalidValue;
  }
  return 0;
}
}  // extern "C++"

int mxh_node_tile(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                  int nsweep, int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc)
{
  if (!xo.p || xo.p == x.p || xo.stride != x.stride) return (int)hipErrorInvalidValue;
  if ((lv->nx & 1) || (lv->ny & 1) || nsweep < 1) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (mode & 1)) return (int)hipErrorInvalidValue;
  // the halo must cover the sweeps (and the residual with its restriction);
  // backward sweeps are counted as the AFW's, which is conservative
  if (nsweep * ((mode & 1) ? TILE_SHRINK_BWD : TILE_SHRINK_FWD) + ((mode & 8) ? TILE_SHRINK_RES : 0) > TILE_H)
    return (int)hipErrorInvalidValue;
  if ((mode & 4) && !xc_in.p) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (!bc.p || !xc.p || bc.stride != xc.stride)) return (int)hipErrorInvalidValue;
  const int TT = 64 - 2 * TILE_H;
  const int ntx = (lv->nx + TT - 1) / TT, nty = (lv->ny + TT - 1) / TT;
  dim3 grid(ntx * nty, nv);
  long long cs = (mode & 8) ? bc.stride : (mode & 4) ? xc_in.stride : 0;
  int e = prec == MXH_F64 ? launch_ntile_mode<double>(mode, grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx)
                          : launch_ntile_mode<float>(mode, grid, lv, bl, b, x, xo, nsweep, xc_in, bc, xc, cs, ntx);
  if (e) return e;
  LAUNCH_CHECK();
  return 0;
}

extern "C++" {
// region geometry of the marching kernels: owned W x Hgt, even, W + 2H <= MARCH_W
static void march_geometry(const mxh_level* lv, int rows, int H, int& ntx, int& nty, int& W, int& Hgt)
{
  const int wmax = MARCH_W - 2 * H;
  ntx = (lv->nx + wmax - 1) / wmax;
  W = (lv->nx + ntx - 1) / ntx;
  W += W & 1;
  if (rows < 2) rows = 128;
  nty = (lv->ny + rows - 1) / rows;
  Hgt = (lv->ny + nty - 1) / nty;
  Hgt += Hgt & 1;
}

extern "C" long long mxh_march_workgroups(const mxh_level* lv, int rows, int nsweep, int nv)
{
  int ntx, nty, W, Hgt;
  march_geometry(lv, rows, nsweep == 2 ? MarchCfg<2>::H : MarchCfg<1>::H, ntx, nty, W, Hgt);
  return (long long)ntx * nty * nv;
}

extern "C" int mxh_cu_count(void)
{
  static int cus = 0;
  if (!cus) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      return 256;
  }
  return cus;
}

template <typename T, int MODE, int NSW, bool NODE>
static void launch_march(dim3 grid, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                         const void* eta, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, long long cs, int ntx, int W,
                         int Hgt)
{
  if (NODE)
    hipLaunchKernelGGL((k_node_march<T, MODE, NSW>), grid, dim3(MarchCfg<NSW>::THREADS), 0, g_stream, *lv, *bl,
                       (const C<T>*)b.p, b.stride, (const C<T>*)x.p, (C<T>*)xo.p, x.stride, (const C<T>*)xc_in.p,
                       (C<T>*)bc.p, (C<T>*)xc.p, cs, ntx, W, Hgt);
  else
    hipLaunchKernelGGL((k_afw_march<T, MODE, NSW>), grid, dim3(MarchCfg<NSW>::THREADS), 0, g_stream, *lv, *bl,
                       (const C<T>*)b.p, b.stride, b.map, (const C<T>*)x.p, (C<T>*)xo.p, x.stride, (const T*)eta,
                       (const C<T>*)xc_in.p, (C<T>*)bc.p, (C<T>*)xc.p, cs, ntx, W, Hgt);
}

template <typename T, int NSW, bool NODE>
static int launch_march_mode(int mode, dim3 grid, const mxh_level* lv, const mxh_bloch* bl, mxh_vecs b, mxh_vecs x,
                             mxh_vecs xo, const void* eta, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, long long cs,
                             int ntx, int W, int Hgt)
{
#define MARCH_CASE(M)                                                                                      \
  case M:                                                                                                  \
    launch_march<T, M, NSW, NODE>(grid, lv, bl, b, x, xo, eta, xc_in, bc, xc, cs, ntx, W, Hgt);            \
    break;
  switch (mode) {
    MARCH_CASE(0)
    MARCH_CASE(2)
    MARCH_CASE(8)
    MARCH_CASE(10)
    MARCH_CASE(1)
    MARCH_CASE(5)
  default: return (int)hipErrorInvalidValue;
  }
#undef MARCH_CASE
  return 0;
}

template <bool NODE>
static int march_common(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x,
                        mxh_vecs xo, const void* eta, int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, int rows,
                        int nsweep)
{
  if (!xo.p || xo.p == x.p || xo.stride != x.stride) return (int)hipErrorInvalidValue;
  if (nsweep != 1 && nsweep != 2) return (int)hipErrorInvalidValue;
  // the halo wraps at most once: nx, ny >= H
  const int H = nsweep == 1 ? MarchCfg<1>::H : MarchCfg<2>::H;
  if ((lv->nx & 1) || (lv->ny & 1) || lv->nx < H || lv->ny < H) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (mode & 1)) return (int)hipErrorInvalidValue;
  if ((mode & 4) && !xc_in.p) return (int)hipErrorInvalidValue;
  if ((mode & 8) && (!bc.p || !xc.p || bc.stride != xc.stride)) return (int)hipErrorInvalidValue;
  if (!NODE && !eta) return (int)hipErrorInvalidValue;
  int ntx, nty, W, Hgt;
  march_geometry(lv, rows, H, ntx, nty, W, Hgt);
  if (W + 2 * H > MARCH_W) return (int)hipErrorInvalidValue;
  // the last tile's region (tiles rounded up to even sizes) ends within one
  // period past the grid (very short tiles on a grid of about the halo's size
  // would not)
  if (ntx * W + H > 2 * lv->nx || nty * Hgt + H > 2 * lv->ny) return (int)hipErrorInvalidValue;
  dim3 grid(ntx * nty, nv);
  long long cs = (mode & 8) ? bc.stride : (mode & 4) ? xc_in.stride : 0;
  int e;
  if (nsweep == 1)
    e = prec == MXH_F64
            ? launch_march_mode<double, 1, NODE>(mode, grid, lv, bl, b, x, xo, eta, xc_in, bc, xc, cs, ntx, W, Hgt)
            : launch_march_mode<float, 1, NODE>(mode, grid, lv, bl, b, x, xo, eta, xc_in, bc, xc, cs, ntx, W, Hgt);
  else
    e = prec == MXH_F64
            ? launch_march_mode<double, 2, NODE>(mode, grid, lv, bl, b, x, xo, eta, xc_in, bc, xc, cs, ntx, W, Hgt)
            : launch_march_mode<float, 2, NODE>(mode, grid, lv, bl, b, x, xo, eta, xc_in, bc, xc, cs, ntx, W, Hgt);
  if (e) return e;
  LAUNCH_CHECK();
  return 0;
}
}  // extern "C++"

int mxh_afw_march(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                  const void* eta, int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, int rows, int nsweep)
{
  return march_common<false>(prec, lv, bl, nv, b, x, xo, eta, mode, xc_in, bc, xc, rows, nsweep);
}

int mxh_node_march(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x, mxh_vecs xo,
                   int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, int rows, int nsweep)
{
  return march_common<true>(prec, lv, bl, nv, b, x, xo, nullptr, mode, xc_in, bc, xc, rows, nsweep);
}

int mxh_resid_restrict(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs bf, mxh_vecs xf,
                       const void* eta_f, mxh_vecs bc, mxh_vecs xc)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_resid_restrict<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl,
                       (const C<double>*)bf.p, bf.stride, (const C<double>*)xf.p, xf.stride,
                       (const double*)eta_f, (C<double>*)bc.p, bc.stride, (C<double>*)xc.p, xc.stride);
  else
    hipLaunchKernelGGL(k_resid_restrict<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl,
                       (const C<float>*)bf.p, bf.stride, (const C<float>*)xf.p, xf.stride,
                       (const float*)eta_f, (C<float>*)bc.p, bc.stride, (C<float>*)xc.p, xc.stride);
  LAUNCH_CHECK();
  return 0;
}

int mxh_prolong(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs c, mxh_vecs f,
                double alpha)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_prolong<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<double>*)c.p,
                       c.stride, (C<double>*)f.p, f.stride, alpha);
  else
    hipLaunchKernelGGL(k_prolong<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<float>*)c.p,
                       c.stride, (C<float>*)f.p, f.stride, (float)alpha);
  LAUNCH_CHECK();
  return 0;
}

int mxh_restrict(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs f, mxh_vecs c,
                 double alpha)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_restrict<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<double>*)f.p,
                       f.stride, (C<double>*)c.p, c.stride, alpha);
  else
    hipLaunchKernelGGL(k_restrict<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<float>*)f.p,
                       f.stride, (C<float>*)c.p, c.stride, (float)alpha);
  LAUNCH_CHECK();
  return 0;
}

int mxh_coarse_setup(int n0, int nprob, const double* bloch, const double* eta0, double* Ae, double* An,
                     int* info)
{
  const int ne = 2 * n0 * n0, nn = n0 * n0;
  if (n0 < 1 || ne > MXH_COARSE_SETUP_MAX || nprob < 1) return (int)hipErrorInvalidValue;
  const size_t sh = sizeof(C<double>) * ((size_t)ne * ne + (size_t)nn * nn) + 8 * sizeof(double);
  hipLaunchKernelGGL(k_coarse_setup, dim3(nprob), dim3(TPB), sh, g_stream, n0, bloch, eta0, Ae, An, info);
  LAUNCH_CHECK();
  return 0;
}

int mxh_coarse_solve(int prec, int n, const double* L, const double* Lper, int vpp, const int* vmap, int nv,
                     mxh_vecs b, mxh_vecs x)
{
  if (n > TPB || n < 1) return (int)hipErrorInvalidValue;
  size_t sh = sizeof(double) * 2 * n;
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_coarse_solve<double>, dim3(nv), dim3(TPB), sh, g_stream, n, L, Lper, vpp, vmap,
                       (const C<double>*)b.p, b.stride, (C<double>*)x.p, x.stride);
  else
    hipLaunchKernelGGL(k_coarse_solve<float>, dim3(nv), dim3(TPB), sh, g_stream, n, L, Lper, vpp, vmap,
                       (const C<float>*)b.p, b.stride, (C<float>*)x.p, x.stride);
  LAUNCH_CHECK();
  return 0;
}

// Partial-sum sets of a Gram launch: bands * nprob <= GRAM_SETS, or nprob
// (one band per problem) when there are more problems than that.
#define GRAM_SETS 2048

size_t mxh_gram_work_bytes(int m, int nprob)
{
  const size_t sets = nprob > GRAM_SETS ? (size_t)nprob : GRAM_SETS;
  return sizeof(double) * 4 * (size_t)(m * (m + 1) / 2) * sets;
}

// The Gram launch with at most max(sets_cap, nprob) partial-sum sets in work.
static int gram_launch(const mxh_level* lv, const mxh_bloch* bl, int nprob, const int* pmap, const int* cmap, int m,
                       mxh_vecs u, const double* eta64, double* out, void* work, int sets_cap)
{
  if (m < 1 || m > GRAM_MMAX || nprob < 1 || sets_cap < 1) return (int)hipErrorInvalidValue;
  const int np = m * (m + 1) / 2;
  const int nbk = (m + GB - 1) / GB, nbp = nbk * (nbk + 1) / 2;
  if (nbp > TPB) return (int)hipErrorInvalidValue;
  // row segment staged per step: T cells of gram_cell_slots(mp) complex and
  // one double each within 64 KB, mp T <= GRAM_ITEMS TPB items.  The kernel's registers
  // allow two workgroups per CU, which the LDS of the widest segment still
  // fits; wider segments (fewer barriers) beat an even spread of cells over
  // the thread groups (block 12: T = 64 measured 9% faster than T = 50).
  const int mp = nbk * GB;
  const size_t cell = (size_t)gram_cell_slots(mp) * sizeof(C<double>) + sizeof(double);
  int T = GRAM_TILE_MAX;
  while (T > 1 && (T * cell > 65536 || mp * T > GRAM_ITEMS * TPB)) T--;
  // row bands: bands * nprob partial sets (sets_cap or nprob) fit the work
  // buffer; the kernel walks each band's segment columns down the rows, so
  // taller bands reload fewer x rows (8 rows per band for 32 problems at 512^2)
  const int sets = nprob > sets_cap ? nprob : sets_cap;
  int bands = sets / nprob;
  if (bands > lv->ny) bands = lv->ny;
  if (bands < 1) bands = 1;
  const int rpb = (lv->ny + bands - 1) / bands;
  bands = (lv->ny + rpb - 1) / rpb;
  size_t sh = T * cell;
  if (sh < TPB * 4 * sizeof(double)) sh = TPB * 4 * sizeof(double);
  hipLaunchKernelGGL(k_gram, dim3(bands, nprob), dim3(TPB), sh, g_stream, *lv, *bl, m, pmap, cmap,
                     (const C<double>*)u.p, u.stride, eta64, (double*)work, T, rpb);
  LAUNCH_CHECK();
  int width = 4 * np * nprob;
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(width)), dim3(TPB), 0, g_stream, (const double*)work,
                     bands, width, out, 0);
  LAUNCH_CHECK();
  return 0;
}

int mxh_gram(const mxh_level* lv, const mxh_bloch* bl, int nprob, const int* pmap, int m, mxh_vecs u,
             const double* eta64, double* out, void* work)
{
  return gram_launch(lv, bl, nprob, pmap, nullptr, m, u, eta64, out, work, GRAM_SETS);
}

int mxh_gram_cols(const mxh_level* lv, const mxh_bloch* bl, int nprob, const int* pmap, const int* cmap, int m,
                  mxh_vecs u, const double* eta64, double* out, void* work)
{
  if (!cmap) return (int)hipErrorInvalidValue;
  return gram_launch(lv, bl, nprob, pmap, cmap, m, u, eta64, out, work, GRAM_SETS);
}

int mxh_gram_static(const mxh_level* lv, const mxh_bloch* bl, int nprob, const int* pmap, int m, mxh_vecs u,
                    const double* eta64, double* out)
{
  if (!g_stream || g_device < 0 || !g_scratch_ptr[g_device]) return (int)hipErrorNotInitialized;
  if (m < 1 || m > GRAM_MMAX || nprob < 1) return (int)hipErrorInvalidValue;
  // the partial-sum sets that fit the scratch array; every set the kernel
  // uses is written in full before the reduction reads it, so nothing needs
  // clearing between calls
  const size_t set_bytes = sizeof(double) * 4 * (size_t)(m * (m + 1) / 2);
  const size_t cap = MXH_GRAM_SCRATCH_BYTES / set_bytes;
  if ((size_t)nprob > cap) return MXH_ESCRATCH;
  const int sets_cap = cap < (size_t)GRAM_SETS ? (int)cap : GRAM_SETS;
  return gram_launch(lv, bl, nprob, pmap, nullptr, m, u, eta64, out, g_scratch_ptr[g_device], sets_cap);
}

int mxh_block_combine(long long n, int nprob, const int* pmap, int m, mxh_vecs u, mxh_vecs e, const double* Cm)
{
  if (m < 1 || m > COMB_MMAX || nprob < 1) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_block_combine, dim3(nblocks(n), nprob), dim3(TPB), 0, g_stream, n, m, pmap,
                     (const C<double>*)u.p, u.stride, (C<double>*)e.p, e.stride, Cm);
  LAUNCH_CHECK();
  return 0;
}

int mxh_combine_x(long long n, int nprob, const int* pmap, int nin, int nout, int nsplit, int gap, mxh_vecs x,
                  int xvpp, mxh_vecs y, int yvpp, const double* Cm)
{
  if (nin < 1 || nin > COMBX_NIN_MAX || nout < 1 || nout > 32 || nprob < 1 || nsplit < 0 || gap < 0)
    return (int)hipErrorInvalidValue;
  dim3 grid(nblocks(n), nprob);
  if (nout <= 16)
    hipLaunchKernelGGL(k_combine_x<16>, grid, dim3(TPB), 0, g_stream, n, nin, nout, nsplit, gap, pmap,
                       (const C<double>*)x.p, x.stride, xvpp, (C<double>*)y.p, y.stride, yvpp, Cm);
  else
    hipLaunchKernelGGL(k_combine_x<32>, grid, dim3(TPB), 0, g_stream, n, nin, nout, nsplit, gap, pmap,
                       (const C<double>*)x.p, x.stride, xvpp, (C<double>*)y.p, y.stride, yvpp, Cm);
  LAUNCH_CHECK();
  return 0;
}

int mxh_lobpcg_combine(long long n, int nprob, const int* pmap, int m, mxh_vecs x, mxh_vecs y, const double* Cm)
{
  if (m < 2 || m > 16 || (m & 1) || nprob < 1) return (int)hipErrorInvalidValue;
  dim3 grid(nblocks(n), nprob);
  if (m <= 8)
    hipLaunchKernelGGL(k_lobpcg_combine<8>, grid, dim3(TPB), 0, g_stream, n, m, pmap, (const C<double>*)x.p,
                       x.stride, (C<double>*)y.p, y.stride, Cm);
  else if (m <= 12)
    hipLaunchKernelGGL(k_lobpcg_combine<12>, grid, dim3(TPB), 0, g_stream, n, m, pmap, (const C<double>*)x.p,
                       x.stride, (C<double>*)y.p, y.stride, Cm);
  else
    hipLaunchKernelGGL(k_lobpcg_combine<16>, grid, dim3(TPB), 0, g_stream, n, m, pmap, (const C<double>*)x.p,
                       x.stride, (C<double>*)y.p, y.stride, Cm);
  LAUNCH_CHECK();
  return 0;
}

size_t mxh_orthonormalize_work_bytes(int k) { return sizeof(double) * (512 * 4 * (size_t)k + 8 * (size_t)k + 64); }

int mxh_orthonormalize(const mxh_level* lv, double xr, double xi, double yr, double yi, const double* eta64, int k,
                       double* X, double* V, double* tmp, double* tau, double* work)
{
  const int nx = lv->nx, ny = lv->ny;
  if (k < 1 || ny < 2 || (long long)k > (long long)nx * (ny / 2)) return (int)hipErrorInvalidValue;
  const long long n = 2LL * nx * ny;
  mxh_bloch bl;
  bl.xr = xr;
  bl.xi = xi;
  bl.yr = yr;
  bl.yi = yi;
  bl.per_problem = nullptr;
  bl.vpp = 1;
  bl.vmap = nullptr;
  C<double> yf;
  yf.r = yr;
  yf.i = yi;
  double* dots = work + 512 * 4 * (size_t)k;
  double* part = work;
  C<double>* Xc = (C<double>*)X;
  C<double>* Vc = (C<double>*)V;
  C<double>* Tc = (C<double>*)tmp;
  const mxh_coef mass = {0.0, 0.0, 1.0, 0.0, nullptr};
  const unsigned nb = nblocks(n) < 1024 ? nblocks(n) : 1024;
  int e;
  // forward pass: Householder vectors in X (edge2d.c:6273-6285)
  for (int i = 0; i < k; i++) {
    mxh_vecs xi_v = {(void*)(X + 2 * (size_t)i * n), n, nullptr};
    if ((e = mxh_addeval(MXH_F64, lv, &bl, 1, xi_v, mxh_vecs{tmp, n, nullptr}, eta64, &mass, 0))) return e;
    if ((e = mxh_dot(MXH_F64, n, 1, mxh_vecs{tmp, 0, nullptr}, xi_v, dots, part))) return e;
    hipLaunchKernelGGL(k_hh_build, dim3(1), dim3(1), 0, g_stream, *lv, yf, i, Xc + (size_t)i * n, Tc, dots, tau);
    LAUNCH_CHECK();
    const int nr = k - i - 1;
    if (nr < 1) continue;
    mxh_vecs rest = {(void*)(X + 2 * (size_t)(i + 1) * n), n, nullptr};
    // d_j = v_i^H M x_j = (M v_i)^H x_j
    if ((e = mxh_dot(MXH_F64, n, nr, mxh_vecs{tmp, 0, nullptr}, rest, dots + 2, part))) return e;
    hipLaunchKernelGGL(k_hh_axpy, dim3(nb, nr), dim3(TPB), 0, g_stream, n, nr, (const C<double>*)(Xc + (size_t)i * n),
                       Xc + (size_t)(i + 1) * n, n, tau, i, 0, dots + 2);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hh_unitfix, dim3((nr + 63) / 64), dim3(64), 0, g_stream, *lv, yf, i, nr,
                       Xc + (size_t)(i + 1) * n, n);
    LAUNCH_CHECK();
  }
  // backward pass: V_i = H_0 ... H_i unit_i (edge2d.c:6287-6296), all i at
  // once: reflection j applied to V_i for i >= j, j from k - 1 down to 0
  if ((e = (int)hipMemsetAsync(V, 0, sizeof(double) * 2 * (size_t)n * k, g_stream))) return e;
  hipLaunchKernelGGL(k_hh_units, dim3((k + 63) / 64), dim3(64), 0, g_stream, *lv, k, Vc, n);
  LAUNCH_CHECK();
  for (int j = k - 1; j >= 0; j--) {
    mxh_vecs xj = {(void*)(X + 2 * (size_t)j * n), n, nullptr};
    if ((e = mxh_addeval(MXH_F64, lv, &bl, 1, xj, mxh_vecs{tmp, n, nullptr}, eta64, &mass, 0))) return e;
    const int nr = k - j;
    mxh_vecs vs = {(void*)(V + 2 * (size_t)j * n), n, nullptr};
    if ((e = mxh_dot(MXH_F64, n, nr, mxh_vecs{tmp, 0, nullptr}, vs, dots, part))) return e;
    hipLaunchKernelGGL(k_hh_axpy, dim3(nb, nr), dim3(TPB), 0, g_stream, n, nr, (const C<double>*)(Xc + (size_t)j * n),
                       Vc + (size_t)j * n, n, tau, j, 1, dots);
    LAUNCH_CHECK();
  }
  return 0;
}

int mxh_gather(int prec, long long n, int nv, mxh_vecs dst, mxh_vecs src)
{
  if (!src.map || nv < 1) return (int)hipErrorInvalidValue;
  dim3 grid(nblocks(n), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_gather<double>, grid, dim3(TPB), 0, g_stream, n, (C<double>*)dst.p, dst.stride,
                       (const C<double>*)src.p, src.stride, src.map);
  else
    hipLaunchKernelGGL(k_gather<float>, grid, dim3(TPB), 0, g_stream, n, (C<float>*)dst.p, dst.stride,
                       (const C<float>*)src.p, src.stride, src.map);
  LAUNCH_CHECK();
  return 0;
}

int mxh_node_addeval(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs x, mxh_vecs y,
                     double ar, double ai, int accumulate)
{
  const long long N = (long long)lv->nx * lv->ny;
  dim3 grid(nblocks(N), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_node_addeval<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)x.p,
                       x.stride, (C<double>*)y.p, y.stride, ar, ai, accumulate);
  else
    hipLaunchKernelGGL(k_node_addeval<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)x.p,
                       x.stride, (C<float>*)y.p, y.stride, ar, ai, accumulate);
  LAUNCH_CHECK();
  return 0;
}

int mxh_node_gs_phase(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs b, mxh_vecs x,
                      double alpha, int colour)
{
  if ((lv->nx & 1) || (lv->ny & 1)) return (int)hipErrorInvalidValue;
  const long long np = (long long)(lv->nx / 2) * (lv->ny / 2);
  dim3 grid(nblocks(np), nv);
  const int ci = colour & 1, cj = colour >> 1;
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_node_gs_phase<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)b.p,
                       b.stride, (C<double>*)x.p, x.stride, alpha, ci, cj);
  else
    hipLaunchKernelGGL(k_node_gs_phase<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)b.p,
                       b.stride, (C<float>*)x.p, x.stride, alpha, ci, cj);
  LAUNCH_CHECK();
  return 0;
}

int mxh_node_resid_restrict(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs bf,
                            mxh_vecs xf, mxh_vecs bc, mxh_vecs xc)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_node_resid_restrict<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl,
                       (const C<double>*)bf.p, bf.stride, (const C<double>*)xf.p, xf.stride, (C<double>*)bc.p,
                       bc.stride, (C<double>*)xc.p, xc.stride);
  else
    hipLaunchKernelGGL(k_node_resid_restrict<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl,
                       (const C<float>*)bf.p, bf.stride, (const C<float>*)xf.p, xf.stride, (C<float>*)bc.p,
                       bc.stride, (C<float>*)xc.p, xc.stride);
  LAUNCH_CHECK();
  return 0;
}

int mxh_node_prolong(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs c, mxh_vecs f,
                     double ar, double ai)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_node_prolong<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<double>*)c.p,
                       c.stride, (C<double>*)f.p, f.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_node_prolong<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<float>*)c.p,
                       c.stride, (C<float>*)f.p, f.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_node_restrict(int prec, const mxh_level* fine, const mxh_bloch* bl, int nv, mxh_vecs f, mxh_vecs c,
                      double ar, double ai)
{
  if ((fine->nx & 1) || (fine->ny & 1)) return (int)hipErrorInvalidValue;
  const long long NC = (long long)(fine->nx / 2) * (fine->ny / 2);
  dim3 grid(nblocks(NC), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_node_restrict<double>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<double>*)f.p,
                       f.stride, (C<double>*)c.p, c.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_node_restrict<float>, grid, dim3(TPB), 0, g_stream, *fine, *bl, (const C<float>*)f.p,
                       f.stride, (C<float>*)c.p, c.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_lift_rhs(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs e, mxh_vecs x,
                 mxh_vecs u, mxh_vecs bg, mxh_vecs xg, double* coef, double* work)
{
  if (x.p && u.p && u.p == e.p) return (int)hipErrorInvalidValue;
  if (coef && !work) return (int)hipErrorInvalidValue;
  const unsigned nb = lift_blocks(lv->ny), nf = frame_blocks(lv);
  const dim3 grid(nb, nv), gridf(nf, nv);
  double* ph = coef ? work : nullptr;
  if (prec == MXH_F64) {
    hipLaunchKernelGGL((k_lift_rhs<double, false>), grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (const C<double>*)x.p, x.stride, (C<double>*)u.p, u.stride, u.map,
                       (C<double>*)bg.p, bg.stride, (C<double>*)xg.p, xg.stride, ph, 0);
    hipLaunchKernelGGL((k_lift_rhs<double, true>), gridf, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (const C<double>*)x.p, x.stride, (C<double>*)u.p, u.stride, u.map,
                       (C<double>*)bg.p, bg.stride, (C<double>*)xg.p, xg.stride, ph, (int)nb);
  } else {
    hipLaunchKernelGGL((k_lift_rhs<float, false>), grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (const C<float>*)x.p, x.stride, (C<double>*)u.p, u.stride, u.map,
                       (C<float>*)bg.p, bg.stride, (C<float>*)xg.p, xg.stride, ph, 0);
    hipLaunchKernelGGL((k_lift_rhs<float, true>), gridf, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, e.map, (const C<float>*)x.p, x.stride, (C<double>*)u.p, u.stride, u.map,
                       (C<float>*)bg.p, bg.stride, (C<float>*)xg.p, xg.stride, ph, (int)nb);
  }
  LAUNCH_CHECK();
  if (coef) {
    double* csum = work + (size_t)(nb + nf) * 2 * nv;
    hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(2 * nv)), dim3(TPB), 0, g_stream, work, (int)(nb + nf),
                       2 * nv, csum, 0);
    LAUNCH_CHECK();
    dim3 g2(row_blocks(lv->ny), nv);
    if (prec == MXH_F64)
      hipLaunchKernelGGL(k_lift_deflate<double>, g2, dim3(TPB), 0, g_stream, *lv, *bl, (C<double>*)bg.p, bg.stride,
                         csum, coef);
    else
      hipLaunchKernelGGL(k_lift_deflate<float>, g2, dim3(TPB), 0, g_stream, *lv, *bl, (C<float>*)bg.p, bg.stride,
                         csum, coef);
    LAUNCH_CHECK();
  }
  return 0;
}

int mxh_lift_apply(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs xg, mxh_vecs u,
                   mxh_vecs x, mxh_vecs e, double alpha, const double* coef, double* work)
{
  if (coef && !work) return (int)hipErrorInvalidValue;
  const unsigned nb = row_blocks(lv->ny);
  double* pdot = nullptr;
  if (coef) {
    pdot = work + (size_t)nb * 2 * nv;
    if (prec == MXH_F64)
      hipLaunchKernelGGL(k_phi_dot<double>, dim3(nb, nv), dim3(TPB), 0, g_stream, *lv, *bl,
                         (const C<double>*)xg.p, xg.stride, work);
    else
      hipLaunchKernelGGL(k_phi_dot<float>, dim3(nb, nv), dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)xg.p,
                         xg.stride, work);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(2 * nv)), dim3(TPB), 0, g_stream, work, (int)nb, 2 * nv,
                       pdot, 0);
    LAUNCH_CHECK();
  }
  dim3 grid(nb, nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_lift_apply<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)xg.p,
                       xg.stride, (const C<double>*)u.p, u.stride, u.map, (const C<double>*)x.p, x.stride,
                       (C<double>*)e.p, e.stride, e.map, alpha, coef, pdot);
  else
    hipLaunchKernelGGL(k_lift_apply<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)xg.p,
                       xg.stride, (const C<double>*)u.p, u.stride, u.map, (const C<float>*)x.p, x.stride,
                       (C<double>*)e.p, e.stride, e.map, alpha, coef, pdot);
  LAUNCH_CHECK();
  return 0;
}

int mxh_gradient(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs v, mxh_vecs e, double ar,
                 double ai)
{
  const long long N = (long long)lv->nx * lv->ny;
  dim3 grid(nblocks(N), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_gradient<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)v.p,
                       v.stride, (C<double>*)e.p, e.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_gradient<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)v.p, v.stride,
                       (C<float>*)e.p, e.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_adjgradient(int prec, const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs e, mxh_vecs v,
                    double ar, double ai)
{
  const long long N = (long long)lv->nx * lv->ny;
  dim3 grid(nblocks(N), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_adjgradient<double>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<double>*)e.p,
                       e.stride, (C<double>*)v.p, v.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_adjgradient<float>, grid, dim3(TPB), 0, g_stream, *lv, *bl, (const C<float>*)e.p,
                       e.stride, (C<float>*)v.p, v.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_scale(int prec, long long n, int nv, mxh_vecs x, double ar, double ai)
{
  dim3 grid(nblocks(n), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_scale<double>, grid, dim3(TPB), 0, g_stream, n, (C<double>*)x.p, x.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_scale<float>, grid, dim3(TPB), 0, g_stream, n, (C<float>*)x.p, x.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_axpy(int prec, long long n, int nv, mxh_vecs x, mxh_vecs y, double ar, double ai)
{
  dim3 grid(nblocks(n), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_axpy<double>, grid, dim3(TPB), 0, g_stream, n, (const C<double>*)x.p, x.stride,
                       (C<double>*)y.p, y.stride, ar, ai);
  else
    hipLaunchKernelGGL(k_axpy<float>, grid, dim3(TPB), 0, g_stream, n, (const C<float>*)x.p, x.stride,
                       (C<float>*)y.p, y.stride, ar, ai);
  LAUNCH_CHECK();
  return 0;
}

int mxh_dot(int prec, long long n, int nv, mxh_vecs x, mxh_vecs y, double* out, double* work)
{
  unsigned nb = red_blocks(n);
  dim3 grid(nb, nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_dot<double>, grid, dim3(TPB), 0, g_stream, n, (const C<double>*)x.p, x.stride,
                       (const C<double>*)y.p, y.stride, work);
  else
    hipLaunchKernelGGL(k_dot<float>, grid, dim3(TPB), 0, g_stream, n, (const C<float>*)x.p, x.stride,
                       (const C<float>*)y.p, y.stride, work);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(2 * nv)), dim3(TPB), 0, g_stream, work, (int)nb, 2 * nv, out,
                     0);
  LAUNCH_CHECK();
  return 0;
}

int mxh_norm(int prec, long long n, int nv, mxh_vecs x, int kind, double* out, double* work)
{
  unsigned nb = red_blocks(n);
  dim3 grid(nb, nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_norm<double>, grid, dim3(TPB), 0, g_stream, n, (const C<double>*)x.p, x.stride, kind,
                       work);
  else
    hipLaunchKernelGGL(k_norm<float>, grid, dim3(TPB), 0, g_stream, n, (const C<float>*)x.p, x.stride, kind,
                       work);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(nv)), dim3(TPB), 0, g_stream, work, (int)nb, nv, out,
                     kind == 0 ? 0 : 1);
  LAUNCH_CHECK();
  return 0;
}

int mxh_center(int prec, long long half, int nv, mxh_vecs x, double* work)
{
  unsigned nb = red_blocks(half);
  dim3 grid(nb, nv);
  double* sums = work + (size_t)nb * nv * 4;
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_halfsums<double>, grid, dim3(TPB), 0, g_stream, half, (const C<double>*)x.p, x.stride,
                       work);
  else
    hipLaunchKernelGGL(k_halfsums<float>, grid, dim3(TPB), 0, g_stream, half, (const C<float>*)x.p, x.stride,
                       work);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_blocks(4 * nv)), dim3(TPB), 0, g_stream, work, (int)nb, 4 * nv,
                     sums, 0);
  LAUNCH_CHECK();
  dim3 g2(nblocks(half), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_center_apply<double>, g2, dim3(TPB), 0, g_stream, half, (C<double>*)x.p, x.stride, sums);
  else
    hipLaunchKernelGGL(k_center_apply<float>, g2, dim3(TPB), 0, g_stream, half, (C<float>*)x.p, x.stride, sums);
  LAUNCH_CHECK();
  return 0;
}

int mxh_convert(int dprec, int sprec, long long n, int nv, mxh_vecs dst, mxh_vecs src)
{
  dim3 grid(nblocks(n), nv);
  if (dprec == MXH_F64 && sprec == MXH_F32)
    hipLaunchKernelGGL((k_convert<double, float>), grid, dim3(TPB), 0, g_stream, n, (C<double>*)dst.p, dst.stride,
                       (const C<float>*)src.p, src.stride);
  else if (dprec == MXH_F32 && sprec == MXH_F64)
    hipLaunchKernelGGL((k_convert<float, double>), grid, dim3(TPB), 0, g_stream, n, (C<float>*)dst.p, dst.stride,
                       (const C<double>*)src.p, src.stride);
  else if (dprec == MXH_F64)
    hipLaunchKernelGGL((k_convert<double, double>), grid, dim3(TPB), 0, g_stream, n, (C<double>*)dst.p,
                       dst.stride, (const C<double>*)src.p, src.stride);
  else
    hipLaunchKernelGGL((k_convert<float, float>), grid, dim3(TPB), 0, g_stream, n, (C<float>*)dst.p, dst.stride,
                       (const C<float>*)src.p, src.stride);
  LAUNCH_CHECK();
  return 0;
}

int mxh_random(int prec, long long n, int nv, mxh_vecs x, unsigned long long seed, int first_vector)
{
  dim3 grid(nblocks(n), nv);
  if (prec == MXH_F64)
    hipLaunchKernelGGL(k_random<double>, grid, dim3(TPB), 0, g_stream, n, (C<double>*)x.p, x.stride, seed,
                       first_vector);
  else
    hipLaunchKernelGGL(k_random<float>, grid, dim3(TPB), 0, g_stream, n, (C<float>*)x.p, x.stride, seed,
                       first_vector);
  LAUNCH_CHECK();
  return 0;
}

int mxh_curl2(const mxh_level* lv, const mxh_bloch* bl, int nv, mxh_vecs e, const double* w, double* grad)
{
  const long long N = (long long)lv->nx * lv->ny;
  hipLaunchKernelGGL(k_curl2, dim3(nblocks(N)), dim3(TPB), 0, g_stream, *lv, *bl, nv, (const C<double>*)e.p,
                     e.stride, w, grad);
  LAUNCH_CHECK();
  return 0;
}

int mxh_event_create(void** ev)
{
  hipEvent_t e;
  hipError_t r = hipEventCreate(&e);
  *ev = (void*)e;
  return (int)r;
}

int mxh_event_record(void* ev) { return (int)hipEventRecord((hipEvent_t)ev, g_stream); }

int mxh_event_elapsed_ms(void* start, void* stop, float* ms)
{
  hipError_t r = hipEventSynchronize((hipEvent_t)stop);
  if (r != hipSuccess) return (int)r;
  return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}

int mxh_event_destroy(void* ev) { return (int)hipEventDestroy((hipEvent_t)ev); }

// Stream capture into a reusable graph.  Capture is thread-local, so other
// host threads keep launching eagerly on their own streams; everything the
// library enqueues on its stream between begin and end becomes a node.
int mxh_capture_begin(void) { return (int)hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal); }

void mxh_capture_abort(void)
{
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(g_stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
}

int mxh_capture_end(void** exec)
{
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(g_stream, &g);
  if (e != hipSuccess) return (int)e;
  if (*exec) {
    // same topology (one node per launch, same kernels), other launch
    // shapes or arguments: update the instantiated graph in place
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    if (hipGraphExecUpdate((hipGraphExec_t)*exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
      (void)hipGraphDestroy(g);
      return 0;
    }
    (void)hipGetLastError();
    (void)hipGraphExecDestroy((hipGraphExec_t)*exec);
    *exec = nullptr;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return (int)e;
  *exec = (void*)x;
  return 0;
}

int mxh_graph_launch(void* exec) { return (int)hipGraphLaunch((hipGraphExec_t)exec, g_stream); }

void mxh_graph_destroy(void* exec)
{
  if (exec) (void)hipGraphExecDestroy((hipGraphExec_t)exec);
}

}  // extern "C"
