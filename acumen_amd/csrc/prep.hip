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

namespace acu {
namespace {

constexpr int MT = 16;  // median tile side (16 x 16 threads)

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// Ascending bitonic sort of 32 values in registers (fully unrolled).
__device__ __forceinline__ void bitonic32(float (&v)[32]) {
#pragma unroll
  for (int k = 2; k <= 32; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const float a = v[i], b = v[l];
          const bool up = (i & k) == 0;
          v[i] = up ? fminf(a, b) : fmaxf(a, b);
          v[l] = up ? fmaxf(a, b) : fminf(a, b);
        }
      }
}

// 3x3x3 median (:420-447): projections l, c, r; neighbours clamped into the
// image; the 14th of 27 in descending order = the 14th in ascending order.
__global__ __launch_bounds__(MT * MT) void median27_kernel(int n, const float* __restrict__ l,
                                                           const float* __restrict__ c,
                                                           const float* __restrict__ r, float* __restrict__ out,
                                                           int tiles_x) {
  __shared__ float t[3][MT + 2][MT + 2];
  const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
  const int x0 = bx * MT, y0 = by * MT;  // 0-based tile origin
  const float* src[3] = {l, c, r};
  for (int i = threadIdx.x; i < 3 * (MT + 2) * (MT + 2); i += MT * MT) {
    const int p = i / ((MT + 2) * (MT + 2)), q = i - p * (MT + 2) * (MT + 2);
    const int yy = q / (MT + 2), xx = q - yy * (MT + 2);
    const int gx = clampi(x0 + xx - 1, 0, n - 1), gy = clampi(y0 + yy - 1, 0, n - 1);
    t[p][yy][xx] = src[p][(size_t)gy * n + gx];
  }
  __syncthreads();
  const int tx = threadIdx.x % MT, ty = threadIdx.x / MT;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= n || y >= n) return;  // after the only barrier
  float v[32];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) v[p * 9 + dy * 3 + dx] = t[p][ty + dy][tx + dx];
#pragma unroll
  for (int i = 27; i < 32; ++i) v[i] = __builtin_inff();
  bitonic32(v);
  out[(size_t)y * n + x] = v[13];
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
// sinogram; sino is nslices(= n) x nfiles x Nw, slice-major.
__global__ __launch_bounds__(256) void correct_kernel(PrepGeom g, const float* __restrict__ imgs, int count,
                                                      int job0, int nfiles, float* __restrict__ sino) {
  const int nw = g.n + g.add_nx;
  const size_t per = (size_t)nw * g.n;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * count) return;
  const int k = (int)(i / per);
  const size_t r = i - (size_t)k * per;
  const int jy = (int)(r / nw) + 1, jx = (int)(r - (size_t)(jy - 1) * nw) + 1;
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
  const size_t total = (size_t)(g.n + g.add_nx) * g.n * count;
  hipLaunchKernelGGL(correct_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, imgs, count, job0,
                     nfiles, sino);
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
