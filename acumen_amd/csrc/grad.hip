// grad.hip -- gfx950 kernels of the projector's adjoint and of the gradients
// of L(x, s) = 1/2 |shift_s(radon(x)) - b|^2 + lambda TV_eps(x).
//
// adjoint_kernel is pixel-driven and deterministic: each thread owns one
// pixel and, per angle, tests the samples whose bilinear stencil can cover
// it.  Sample (jx, jy) of angle t sits at
//   xr = (c (jx-m) - s (jy-m)) + m,  yr = (s (jx-m) + c (jy-m)) + m
// (tomography/tv_hpc.F90:143-144) and covers the pixels x1..x1+1, y1..y1+1,
// x1 = floor(xr), y1 = floor(yr), when 2 <= x1,y1 <= N-1 (:145-150).  The
// samples covering pixel (px,py) therefore lie in the rotated 2x2 square
// around the inverse rotation q of (px,py); its bounding box holds at most
// 3 x 3 lattice points, and each is tested with the projector's own FP32
// expressions (contraction off), so the weights are exactly the projector's
// and the kernel is its transpose.
#include "acu_grad.hpp"

namespace acu {
namespace {

template <class Task, int Task::*Begin>
__device__ __forceinline__ int find_by(const Task* tasks, int ntasks, int gb) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].*Begin <= gb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum of v over the workgroup (GR_THREADS), valid in thread 0.
__device__ double block_sum(double v, double* wsum) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < GR_THREADS / 64; ++w) r += wsum[w];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(ADJ_T * ADJ_T) void adjoint_kernel(const AdjTask* __restrict__ tasks, int ntasks,
                                                                int base, int nblocks) {
#pragma clang fp contract(off)
  const int lbk = xcd_remap(blockIdx.x, gridDim.x);
  if (lbk >= nblocks) return;
  const int gb = base + lbk;
  const AdjTask& tk = tasks[find_by<AdjTask, &AdjTask::block_begin>(tasks, ntasks, gb)];
  const int N = tk.N, T = tk.T;
  const float m = tk.mid;
  const int local = gb - tk.block_begin;
  const int ty = local / tk.ntx, tx = local - ty * tk.ntx;
  const int px = tx * ADJ_T + (int)(threadIdx.x % ADJ_T) + 1;  // 1-based pixel
  const int py = ty * ADJ_T + (int)(threadIdx.x / ADJ_T) + 1;
  if (px > N || py > N) return;  // no barrier below
  const float X = (float)px - m, Y = (float)py - m;
  const ACU_GLOBAL float* __restrict__ sino = as_global(tk.sino);
  // Pixels with 3 <= px, py <= N-1: the cells x1 in {px-1, px} that can
  // cover them always pass 2 <= x1 <= N-1, and the projector's weights
  // (x1+1) - xr and xr - x1 are differences of floats within a factor of two
  // of each other, hence exact (Sterbenz), and equal to 1 - |xr - px|, which
  // is exact too (xr - px is, and 1 - d is a multiple of xr's ulp below 1).
  // So the test and weights of those pixels reduce to the two weights being
  // positive (a weight of 0 adds nothing) -- the same bits with no floor,
  // conversion or select.  The rim keeps the projector's own test.
  const bool inner = px >= 3 && px <= N - 1 && py >= 3 && py <= N - 1;
  const float pxf = (float)px, pyf = (float)py;
  float acc = 0.0f;
  for (int t0 = 0; t0 < T; t0 += ADJ_AB) {
    const int t1 = min(T, t0 + ADJ_AB);
    float blk = 0.0f;  // FP32 sum over one block of angles, then into acc
    for (int t = t0; t < t1; ++t) {
      const float2 v = const_f2(tk.sc, t);
      const float s = v.x, c = v.y;
      // inverse rotation of the pixel centre, lattice coordinates; the margin
      // covers the rounding of both this estimate and the exact positions
      const float qx = (c * X + s * Y) + m;
      const float qy = (c * Y - s * X) + m;
      const float hw = fabsf(c) + fabsf(s) + 1e-3f;
      const int jx0 = (int)ceilf(qx - hw), jy0 = (int)ceilf(qy - hw);
      const ACU_GLOBAL float* row = sino + (size_t)t * N;
      if (inner) {
        float sw[3], cw[3], rv[3];
        bool yok[3];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
          const int jy = jy0 + bb;
          const float w = (float)jy - m;
          sw[bb] = s * w;
          cw[bb] = c * w;
          yok[bb] = jy >= 1 && jy <= N;
          rv[bb] = yok[bb] ? row[jy - 1] : 0.0f;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int jx = jx0 + a;
          const float u = (float)jx - m;
          const float cu = c * u, su = s * u;
          const bool xok = jx >= 1 && jx <= N;
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            const float xr = (cu - sw[bb]) + m;
            const float yr = (su + cw[bb]) + m;
            const float wx = 1.0f - fabsf(xr - pxf);
            const float wy = 1.0f - fabsf(yr - pyf);
            if (xok && yok[bb] && wx > 0.0f && wy > 0.0f) blk += (rv[bb] * wx) * wy;
          }
        }
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int jx = jx0 + a;
          const float u = (float)jx - m;
          const float cu = c * u, su = s * u;
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            const int jy = jy0 + bb;
            const float w = (float)jy - m;
            const float xr = (cu - s * w) + m;
            const float yr = (su + c * w) + m;
            const float fx = floorf(xr), fy = floorf(yr);
            const int x1 = (int)fx, y1 = (int)fy;
            const int dx = px - x1, dy = py - y1;
            const bool hit = jx >= 1 && jx <= N && jy >= 1 && jy <= N && (unsigned)dx <= 1u &&
                             (unsigned)dy <= 1u && x1 >= 2 && x1 <= N - 1 && y1 >= 2 && y1 <= N - 1;
            if (hit) {
              // the projector's weights: (x2 - xr) for x1, (xr - x1) for x2
              const float wx = dx == 0 ? (fx + 1.0f) - xr : xr - fx;
              const float wy = dy == 0 ? (fy + 1.0f) - yr : yr - fy;
              blk += (row[jy - 1] * wx) * wy;
            }
          }
        }
      }
    }
    acc += blk;
  }
  tk.out[(size_t)(py - 1) * N + (px - 1)] = acc / (float)N;
}

__device__ __forceinline__ float at(const ACU_GLOBAL float* row, int N, int i) {
  return (i >= 1 && i <= N) ? row[i - 1] : 0.0f;
}

// detector shift of one row: out(d) = (1 - w) y(d + o) + w y(d + o + 1),
// o = floor(-s), w = -s - o (exact in FP32)
__device__ __forceinline__ void shift_weights(float s, int& o, float& w) {
  const float sg = -s;
  const float fl = floorf(sg);
  o = (int)fl;
  w = sg - fl;
}

__global__ __launch_bounds__(GR_THREADS) void shift_resid_kernel(const ResTask* __restrict__ tasks, int ntasks,
                                                                 int base, int nrows) {
#pragma clang fp contract(off)
  __shared__ double wsum[GR_THREADS / 64];
  const int lrow = xcd_remap(blockIdx.x, gridDim.x);
  if (lrow >= nrows) return;  // uniform per workgroup
  const int gr = base + lrow;
  const ResTask& tk = tasks[find_by<ResTask, &ResTask::row_begin>(tasks, ntasks, gr)];
  const int N = tk.N;
  const int t = gr - tk.row_begin;
  int o = 0;
  float w = 0.0f;
  if (tk.shift) shift_weights(tk.shift[t], o, w);
  const ACU_GLOBAL float* y = as_global(tk.y) + (size_t)t * N;
  const ACU_GLOBAL float* bm = as_global(tk.b) + (size_t)t * N;
  const float w1 = 1.0f - w;
  auto resid = [&](int d) -> float {  // r(d), zero off the detector
    if (d < 1 || d > N) return 0.0f;
    return (w1 * at(y, N, d + o) + w * at(y, N, d + o + 1)) - bm[d - 1];
  };
  double data = 0.0, g = 0.0, h = 0.0;
  for (int d = 1 + (int)threadIdx.x; d <= N; d += GR_THREADS) {
    const float ya = at(y, N, d + o), yb = at(y, N, d + o + 1);
    const float r = (w1 * ya + w * yb) - bm[d - 1];
    const double dy = (double)ya - (double)yb;
    data += 0.5 * (double)r * (double)r;
    g += (double)r * dy;
    h += dy * dy;
    // rs(k) = (1 - w) r(k - o) + w r(k - o - 1) at k = d
    tk.rs[(size_t)t * N + d - 1] = w1 * resid(d - o) + w * resid(d - o - 1);
  }
  data = block_sum(data, wsum);
  g = block_sum(g, wsum);
  h = block_sum(h, wsum);
  if (threadIdx.x == 0) {
    tk.part[t] = data;
    if (tk.gs) tk.gs[t] = (float)g;
    if (tk.hs) tk.hs[t] = (float)h;
  }
}

__global__ __launch_bounds__(GR_THREADS) void tvgrad_kernel(const TVGTask* __restrict__ tasks, int ntasks,
                                                            int base, int nblocks) {
  __shared__ double wsum[GR_THREADS / 64];
  const int lbk = xcd_remap(blockIdx.x, gridDim.x);
  if (lbk >= nblocks) return;  // uniform per workgroup
  const int gb = base + lbk;
  const TVGTask& tk = tasks[find_by<TVGTask, &TVGTask::block_begin>(tasks, ntasks, gb)];
  const int N = tk.N;
  const int lb = gb - tk.block_begin;
  const size_t i = (size_t)lb * GR_THREADS + threadIdx.x;
  double tv = 0.0;
  if (i < (size_t)N * N) {
    const int y = (int)(i / (size_t)N), x = (int)(i - (size_t)y * N);  // 0-based
    const int xp = x == N - 1 ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
    const int yp = y == N - 1 ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
    const ACU_GLOBAL float* u = as_global(tk.x);
    auto U = [&](int a, int b) { return u[(size_t)b * N + a]; };
    const float e2 = tk.eps * tk.eps;
    const float u0 = U(x, y);
    // difference vector and its norm at (x, y), at (x+1, y) and at (x, y+1)
    const float dx0 = U(xm, y) - u0, dy0 = U(x, ym) - u0;
    const float n0 = sqrtf(dx0 * dx0 + dy0 * dy0 + e2);
    tv = (double)n0;
    if (tk.gx) {
      const float u1 = U(xp, y);
      const float dx1 = u0 - u1, dy1 = U(xp, ym) - u1;
      const float n1 = sqrtf(dx1 * dx1 + dy1 * dy1 + e2);
      const float u2 = U(x, yp);
      const float dx2 = U(xm, yp) - u2, dy2 = u0 - u2;
      const float n2 = sqrtf(dx2 * dx2 + dy2 * dy2 + e2);
      const float gtv = (dx1 / n1 - dx0 / n0) + (dy2 / n2 - dy0 / n0);
      tk.gx[i] = tk.gx[i] + tk.lambda * gtv;
    }
  }
  tv = block_sum(tv, wsum);
  if (threadIdx.x == 0) tk.part[lb] = tv;
}

__global__ __launch_bounds__(GR_THREADS) void loss_finalize_kernel(const LossTask* __restrict__ tasks) {
  __shared__ double wsum[GR_THREADS / 64];
  const LossTask& tk = tasks[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < tk.nd; i += GR_THREADS) a += tk.pd[i];
  for (int i = threadIdx.x; i < tk.nt; i += GR_THREADS) b += tk.pt[i];
  a = block_sum(a, wsum);
  b = block_sum(b, wsum);
  if (threadIdx.x == 0) {
    tk.loss[0] = a + tk.lambda * b;
    tk.loss[1] = a;
    tk.loss[2] = b;
  }
}

}  // namespace

void launch_adjoint(const AdjTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(adjoint_kernel, dim3(xcd_grid(blocks)), dim3(ADJ_T * ADJ_T), 0, s, d_tasks, ntasks, base,
                     blocks);
}

void launch_shift_resid(const ResTask* d_tasks, int ntasks, int base, int rows, hipStream_t s) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(shift_resid_kernel, dim3(xcd_grid(rows)), dim3(GR_THREADS), 0, s, d_tasks, ntasks, base,
                     rows);
}

void launch_tvgrad(const TVGTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(tvgrad_kernel, dim3(xcd_grid(blocks)), dim3(GR_THREADS), 0, s, d_tasks, ntasks, base,
                     blocks);
}

void launch_loss_finalize(const LossTask* d_tasks, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(loss_finalize_kernel, dim3(ntasks), dim3(GR_THREADS), 0, s, d_tasks);
}

}  // namespace acu
