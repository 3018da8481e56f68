// prep.hip -- gfx950 kernels of the projection -> sinogram preprocessing
// (tomography/projections_to_sinogram_hpc.F90): 3x3x3 median filter,
// rotation / centre-of-rotation correction by clamped bilinear resampling
// written straight into the sinogram layout, and the mirror error that the
// centre-of-rotation search minimises.
//
// Layouts follow the reference (first index fastest): a projection is
// img(x, y), n x n (x = detector column, y = detector row = slice); the
// corrected projection row jy of projection `job` is sinogram row `job` of
// slice jy, Nw = n + add_Nx values long.
#include "acu_prep.hpp"
#include "median_net.hpp"

#include <algorithm>

namespace acu {
namespace {

constexpr int MT = 16;  // median tile side (16 x 16 threads)

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// The 14th smallest of 27 values in registers: the compile-time network of
// median_net.hpp, fully unrolled (153 compare-exchanges).
__device__ __forceinline__ float median27(float (&v)[27]) {
#pragma unroll
  for (int i = 0; i < MEDIAN27_NET.n; ++i) {
    const float a = v[MEDIAN27_NET.lo[i]], b = v[MEDIAN27_NET.hi[i]];
    v[MEDIAN27_NET.lo[i]] = fminf(a, b);
    v[MEDIAN27_NET.hi[i]] = fmaxf(a, b);
  }
  return v[MEDIAN27_NET.out];
}

// 3x3x3 median (:420-447): projections l, c, r; neighbours clamped into the
// image; the 14th of 27 in descending order = the 14th in ascending order.
// This is the median the reference's loop nest describes: its register
// slots 3*jxs+jys+2 (+9, +18) run from -2 to 24 in an array of 27 (:433-435
// against :134), so its own code writes 3 neighbours out of bounds and sorts
// 3 stale slots in their place; that undefined result is not reproduced.
// Values that compare equal but differ in bits (+0 and -0) or NaN are the
// exception: the reference's exchange sort is not stable, so which of them
// it leaves in the middle depends on its order of exchanges.
__global__ __launch_bounds__(MT * MT) void median27_kernel(int n, const float* __restrict__ l,
                                                           const float* __restrict__ c,
                                                           const float* __restrict__ r, float* __restrict__ out,
                                                           int tiles_x) {
  constexpr int H = MT + 2, HH = H * H, NS = (3 * HH + MT * MT - 1) / (MT * MT);
  __shared__ float t[3 * HH];  // [p][yy][xx]
  const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
  const int x0 = bx * MT, y0 = by * MT;  // 0-based tile origin
  // halo tiles of the three projections: every load issued before the first
  // LDS store (element i = p*HH + yy*H + xx; the divisions are by constants)
  float st[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int i = (int)threadIdx.x + k * MT * MT;
    const int ic = i < 3 * HH ? i : 0;
    const int p = ic / HH, q = ic - p * HH;
    const int yy = q / H, xx = q - yy * H;
    const int gx = clampi(x0 + xx - 1, 0, n - 1), gy = clampi(y0 + yy - 1, 0, n - 1);
    const float* sp = p == 0 ? l : p == 1 ? c : r;  // no private-memory pointer table
    st[k] = sp[(size_t)gy * n + gx];
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int i = (int)threadIdx.x + k * MT * MT;
    if (i < 3 * HH) t[i] = st[k];
  }
  __syncthreads();
  const int tx = threadIdx.x % MT, ty = threadIdx.x / MT;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= n || y >= n) return;  // after the only barrier
  float v[27];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) v[p * 9 + dy * 3 + dx] = t[p * HH + (ty + dy) * H + tx + dx];
  out[(size_t)y * n + x] = median27(v);
}

// The resampling of :449-464 at output (jx, jy), 1-based, in the
// reference's FP32 order: u = (jx - mx) - S*shift, xr = clamp((c*u - s*(jy-my)) + mx),
// yr = clamp((s*u + c*(jy-my)) + my), taps clamped into the image; a tap
// pair clamped onto one pixel has both weights zero.
__device__ __forceinline__ float prep_sample(const PrepGeom& g, const float* __restrict__ img, int jx, int jy) {
#pragma clang fp contract(off)
  const int n = g.n;
  const float u = ((float)jx - g.mx) - g.S * g.shift;
  const float w = (float)jy - g.my;
  const float xr = clampf((g.c * u - g.s * w) + g.mx, 1.0f, (float)n);
  const float yr = clampf((g.s * u + g.c * w) + g.my, 1.0f, (float)n);
  const int y1 = clampi((int)floorf(yr), 1, n), y2 = clampi(y1 + 1, 1, n);
  const int x1 = clampi((int)floorf(xr), 1, n), x2 = clampi(x1 + 1, 1, n);
  const float ax = (float)x2 - xr, bx = xr - (float)x1, ay = (float)y2 - yr, by = yr - (float)y1;
  const float* r1 = img + (size_t)(y1 - 1) * n;
  const float* r2 = img + (size_t)(y2 - 1) * n;
  float v = (r1[x1 - 1] * ax) * ay;
  v = v + (r1[x2 - 1] * bx) * ay;
  v = v + (r2[x1 - 1] * ax) * by;
  v = v + (r2[x2 - 1] * bx) * by;
  return v;
}

// Corrected projections of a batch -> sinogram rows: projection k of the
// batch (global index job0 + k) fills row job0 + k of every slice's
// sinogram; sino is nslices(= n) x nfiles x Nw, slice-major.  Grid: x over
// the Nw output columns, y = the detector row jy - 1 - row0, z = the projection.
__global__ __launch_bounds__(256) void correct_kernel(PrepGeom g, const float* __restrict__ imgs, int job0,
                                                      int nfiles, int row0, float* __restrict__ sino) {
  const int nw = g.n + g.add_nx;
  const int jx = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  if (jx > nw) return;
  const int jy = row0 + (int)blockIdx.y + 1, k = (int)blockIdx.z;
  const float v = prep_sample(g, imgs + (size_t)k * g.n * g.n, jx, jy);
  sino[((size_t)(jy - 1) * nfiles + (size_t)(job0 + k)) * nw + (jx - 1)] = v;
}

// |V0(jx,jy) - V180(Nw-jx+1,jy)| summed per block in FP64 (:278-283).
__global__ __launch_bounds__(256) void mirror_kernel(PrepGeom g, const float* __restrict__ p0,
                                                     const float* __restrict__ p180, double* __restrict__ part) {
  __shared__ double ws[4];
  const int nw = g.n + g.add_nx;
  const size_t total = (size_t)nw * g.n;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < total) {
    const int jy = (int)(i / nw) + 1, jx = (int)(i - (size_t)(jy - 1) * nw) + 1;
    e = (double)fabsf(prep_sample(g, p0, jx, jy) - prep_sample(g, p180, nw - jx + 1, jy));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__global__ __launch_bounds__(256) void sum_kernel(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double ws[4];
  double e = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) e += part[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) *out = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

}  // namespace

void launch_median27(int n, const float* l, const float* c, const float* r, float* out, hipStream_t s) {
  const int tx = (n + MT - 1) / MT;
  hipLaunchKernelGGL(median27_kernel, dim3(tx * tx), dim3(MT * MT), 0, s, n, l, c, r, out, tx);
}

void launch_correct(const PrepGeom& g, const float* imgs, int count, int job0, int nfiles, float* sino,
                    hipStream_t s) {
  const unsigned gx = (unsigned)((g.n + g.add_nx + 255) / 256);
  for (int k0 = 0; k0 < count; k0 += 65535)  // the grid's y and z extents are at most 65535
    for (int r0 = 0; r0 < g.n; r0 += 65535) {
      const int cnt = std::min(65535, count - k0), rows = std::min(65535, g.n - r0);
      hipLaunchKernelGGL(correct_kernel, dim3(gx, (unsigned)rows, (unsigned)cnt), dim3(256), 0, s, g,
                         imgs + (size_t)k0 * g.n * g.n, job0 + k0, nfiles, r0, sino);
    }
}

int mirror_blocks(const PrepGeom& g) { return (int)(((size_t)(g.n + g.add_nx) * g.n + 255) / 256); }

void launch_mirror(const PrepGeom& g, const float* p0, const float* p180, double* part, double* out,
                   hipStream_t s) {
  const int nb = mirror_blocks(g);
  hipLaunchKernelGGL(mirror_kernel, dim3(nb), dim3(256), 0, s, g, p0, p180, part);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, part, nb, out);
}

PrepGeom prep_geom(int n, int add_nx, float theta, float shift) {
#pragma clang fp contract(off)
  PrepGeom g;
  g.n = n;
  g.add_nx = add_nx;
  const int nw = n + add_nx;
  g.mx = (float)(nw - 1) / 2.0f;
  g.my = (float)(n - 1) / 2.0f;
  g.S = (float)n;
  g.shift = shift;
  g.c = cosf(theta);
  g.s = sinf(theta);
  return g;
}

}  // namespace acu

// ---- FITS ingest normalisation (tomography/fits_to_hpc.F90:196-248) ----
//
// mean: one thread per pixel (VEC: four adjacent pixels, one 16-byte
// access) loops over the images in file order.
//
// normalize: -log(I / b), b = max(OB, 1).  The open beam is fixed per pixel
// across the projections, so its reciprocal y = RN(1/b) is formed once per
// pixel with the IEEE division and each quotient then costs one product and
// two FMAs:  q = RN(I y),  q' = RN(q + RN(I - b q) y).  With y correctly
// rounded and q within one ulp of I/b, the remainder I - b q is exact and q'
// is the correctly rounded quotient (P. Markstein, IBM J. Res. Dev. 34(1),
// 1990; Cornea-Hasegan, Golliver, Markstein, ARITH-14, 1999, Theorem 1),
// i.e. bit-identical to I / b, provided nothing underflows or overflows: for
// |I| < 2^-100, |q| < 2^-100, b = inf, or I not finite the kernel takes the
// IEEE division instead.  Each thread handles one pixel group of
// ING_CHUNK consecutive projections (grid: pixel groups x projection
// chunks); longer per-thread runs across images measured slower (the
// write streams lose DRAM locality).  VEC needs npix % 4 == 0 and 16-byte
// aligned pointers (checked on the host).
namespace acu {
namespace {

constexpr int ING_CHUNK = 2;

template <bool VEC> struct Px { typedef float type; };
template <> struct Px<true> { typedef float4 type; };

__device__ __forceinline__ float add_(float a, float b) { return a + b; }
__device__ __forceinline__ float4 add_(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float div_(float a, float c) { return a / c; }
__device__ __forceinline__ float4 div_(float4 a, float c) { return make_float4(a.x / c, a.y / c, a.z / c, a.w / c); }

struct BY { float b, y; };
__device__ __forceinline__ BY by_(float ob) {
  const float b = fmaxf(ob, 1.0f);
  return {b, 1.0f / b};
}
__device__ __forceinline__ float nlog_(float a, BY p) {
  float q = a * p.y;
  if (fabsf(a) >= 0x1p-100f && fabsf(a) <= 3.40282347e38f && fabsf(q) >= 0x1p-100f && p.b <= 3.40282347e38f)
    q = fmaf(fmaf(-p.b, q, a), p.y, q);
  else
    q = a / p.b;
  return -logf(q);
}

// out = (sum over images, in order, FP32) / count: the open-beam average of
// :203-215, accumulated one file at a time from bg = 0
template <bool VEC>
__global__ __launch_bounds__(256) void mean_kernel(const float* __restrict__ imgs, int count, size_t npix,
                                                   float* __restrict__ out) {
#pragma clang fp contract(off)
  typedef typename Px<VEC>::type P;
  const size_t n = VEC ? npix / 4 : npix;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P* src = reinterpret_cast<const P*>(imgs);
  P s;
  if constexpr (VEC) s = make_float4(0.0f, 0.0f, 0.0f, 0.0f); else s = 0.0f;  // -0 + ... gives +0, as bg = 0
  for (int k = 0; k < count; ++k) s = add_(s, src[(size_t)k * n + i]);
  reinterpret_cast<P*>(out)[i] = div_(s, (float)count);
}

// projection = -log(image / max(open_beam, 1)) (:243); blockIdx.y = chunk
template <bool VEC>
__global__ __launch_bounds__(256) void normalize_kernel(const float* __restrict__ imgs, int count, size_t npix,
                                                        const float* __restrict__ ob, float* __restrict__ out) {
#pragma clang fp contract(off)
  const size_t n = VEC ? npix / 4 : npix;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k0 = blockIdx.y * ING_CHUNK, k1 = min(count, k0 + ING_CHUNK);
  if constexpr (VEC) {
    const float4 o = reinterpret_cast<const float4*>(ob)[i];
    const BY p0 = by_(o.x), p1 = by_(o.y), p2 = by_(o.z), p3 = by_(o.w);
    const float4* src = reinterpret_cast<const float4*>(imgs);
    float4* dst = reinterpret_cast<float4*>(out);
    for (int k = k0; k < k1; ++k) {
      const float4 v = src[(size_t)k * n + i];
      dst[(size_t)k * n + i] = make_float4(nlog_(v.x, p0), nlog_(v.y, p1), nlog_(v.z, p2), nlog_(v.w, p3));
    }
  } else {
    const BY p = by_(ob[i]);
    for (int k = k0; k < k1; ++k) out[(size_t)k * n + i] = nlog_(imgs[(size_t)k * n + i], p);
  }
}

bool vec_ok(size_t npix, const void* a, const void* b, const void* c) {
  const uintptr_t m = (uintptr_t)a | (uintptr_t)b | (uintptr_t)c;
  return npix % 4 == 0 && (m & 15) == 0;
}

}  // namespace

void launch_mean(const float* imgs, int count, size_t npix, float* out, hipStream_t s) {
  if (vec_ok(npix, imgs, out, out))
    hipLaunchKernelGGL(mean_kernel<true>, dim3((unsigned)((npix / 4 + 255) / 256)), dim3(256), 0, s, imgs, count,
                       npix, out);
  else
    hipLaunchKernelGGL(mean_kernel<false>, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, imgs, count,
                       npix, out);
}

void launch_normalize(const float* imgs, int count, size_t npix, const float* ob, float* out, hipStream_t s) {
  const bool vec = vec_ok(npix, imgs, ob, out);
  const unsigned gx = (unsigned)(((vec ? npix / 4 : npix) + 255) / 256);
  constexpr int kMaxImages = 65535 * ING_CHUNK;  // gridDim.y limit
  for (int k = 0; k < count; k += kMaxImages) {
    const int c = std::min(kMaxImages, count - k);
    const dim3 grid(gx, (unsigned)((c + ING_CHUNK - 1) / ING_CHUNK));
    const float* in = imgs + (size_t)k * npix;
    float* o = out + (size_t)k * npix;
    if (vec)
      hipLaunchKernelGGL(normalize_kernel<true>, grid, dim3(256), 0, s, in, c, npix, ob, o);
    else
      hipLaunchKernelGGL(normalize_kernel<false>, grid, dim3(256), 0, s, in, c, npix, ob, o);
  }
}

}  // namespace acu
