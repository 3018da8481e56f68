// fp.hip -- ray-driven bilinear forward projection (radon) on gfx950.
//
// Restates radon (tomography/tv_hpc.F90:129-159):
//   xr = cos(c)*(x-m) - sin(c)*(y-m) + m,  yr = sin(c)*(x-m) + cos(c)*(y-m) + m,
//   c = pi*(theta-90)/180, x = 1..N steps along the ray, y = detector bin,
//   sino(y,t) = (1/N) * sum_x bilinear(img; xr, yr)
// where a sample is used only if x1 = floor(xr), y1 = floor(yr) satisfy
// min(x1,y1) > 1 and max(x1+1,y1+1) < N+1.  Steps are summed in order.
//
// gfx950 mapping:
//  * sample positions are bit-identical to the reference's FP32 ones: the
//    products cos(c)*(x-m), sin(c)*(x-m), sin(c)*(y-m), cos(c)*(y-m) are each
//    rounded once and the two sums are evaluated in the reference's order
//    with FMA contraction off.  So every floor decision -- including the
//    bounds test on the image border, which flips on rounding at angles such
//    as 0 and 90 degrees -- is the reference's own;
//  * cells pass: every image is first rewritten as (N+1)^2 bilinear cells
//    Q(cx,cy) = (v11, v21-v11, v12-v11, v22-v21-v12+v11) indexed by the
//    1-based floor pair (cx, cy) in [0, N]^2, and Q = 0 for every cell the
//    reference's bounds test rejects.  A sample is then one 16-byte load and
//    three FMAs, the bounds test costs nothing, and clamping the cell index
//    into [0, N] keeps every access in bounds without changing any result;
//  * the cells are written twice, x-major and y-major; each angle reads the
//    copy whose contiguous axis is closer to the detector direction, and the
//    lanes are skewed along their rays (see ray_sum) so that one gather of a
//    wave stays within about one row of that copy;
//  * each wave skips the steps before the first and after the last cell
//    any of its rays can use (the skipped samples are all rejected ones);
//  * workgroups are ordered (chunk, ray block, angle) and remapped so that
//    the ones an XCD runs together are neighbouring angles of one ray block:
//    their rays sweep nearly the same cells, so the gathers hit that XCD's
//    L2 (measured: 11.7 ms -> 6.5 ms per 2049^2 x 1500 launch against the
//    hardware's round-robin order).  A ray's sweep may be cut into chunks of
//    clen steps (tuning knob ACUMEN_FP_CHUNK, default: one chunk); each
//    chunk writes its own partial sum and sum_kernel adds a ray's chunks in
//    step order.
#include "acu_internal.hpp"

namespace acu {
namespace {

constexpr int CT = 32;  // cell tile side of the cells pass

template <typename Task>
__device__ __forceinline__ int find_task(const Task* tasks, int ntasks, int gblock) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block_begin <= gblock) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void cells_kernel(const CellTask* __restrict__ tasks, int ntasks, int base) {
  __shared__ float4 tile[CT][CT + 1];
  const int gb = base + blockIdx.x;
  const CellTask& tk = tasks[find_task(tasks, ntasks, gb)];
  const int N = tk.N, P = N + 1;
  const int local = gb - tk.block_begin;
  const int nb = (P + CT - 1) / CT;
  const int bx = local % nb, by = local / nb;
  const int tid = threadIdx.x;
  const float* __restrict__ img = tk.img;
  // x-major cells, one row of 32 cells per 32 lanes
  for (int e = tid; e < CT * CT; e += 256) {
    const int lx = e % CT, ly = e / CT;
    const int cx = bx * CT + lx, cy = by * CT + ly;  // 1-based floor indices
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cx >= 2 && cx <= N - 1 && cy >= 2 && cy <= N - 1) {
      const size_t o = (size_t)(cy - 1) * N + (cx - 1);
      const float v11 = img[o], v21 = img[o + 1];
      const float v12 = img[o + N], v22 = img[o + N + 1];
      q = make_float4(v11, v21 - v11, v12 - v11, (v22 - v21) - (v12 - v11));
    }
    tile[ly][lx] = q;
    if (cx < P && cy < P) tk.cells[(size_t)cy * P + cx] = q;
  }
  __syncthreads();
  for (int e = tid; e < CT * CT; e += 256) {
    const int ly = e % CT, lx = e / CT;
    const int cx = bx * CT + lx, cy = by * CT + ly;
    if (cx < P && cy < P) tk.cellsT[(size_t)cx * P + cy] = tile[ly][lx];
  }
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Range of step offsets u = x - m over which a + b*u stays in [lo, hi].
__device__ __forceinline__ void step_range(float a, float b, float lo, float hi, float& u0, float& u1) {
  if (fabsf(b) < 1e-6f) {
    if (a >= lo && a <= hi) { u0 = -1e30f; u1 = 1e30f; } else { u0 = 1e30f; u1 = -1e30f; }
  } else {
    const float p = (lo - a) / b, q = (hi - a) / b;
    u0 = fminf(p, q);
    u1 = fmaxf(p, q);
  }
}

// One wave = 64 neighbouring rays of one angle, one lane per ray, each lane
// summing its own samples in step order (the reference's order).  Lane l is
// skewed by off_l steps (x_l = k - off_l at iteration k) with off_l chosen so
// that the wave's 64 samples of one iteration stay within about one row of
// cells of the copy being read: neighbouring rays are offset by (-s, c) per
// bin, and the skew cancels the drift across the contiguous axis.  A gather
// then touches ~8-12 cache lines at every angle instead of up to 64.
//
// Positions are the reference's FP32 values: u = x - m exact, c*u and s*u
// rounded once, then (c*u - s*(y-m)) + m and (s*u + c*(y-m)) + m.
template <bool XMAJOR>
__device__ __forceinline__ float ray_sum(const ACU_GLOBAL float4* __restrict__ cells, int xs, int xe,
                                         int off, int k0, int k1, float s, float c, float sy,
                                         float cy, float m, int N) {
  // lanes sum steps x in [xs, xe]; iteration k visits x = k - off
#pragma clang fp contract(off)
  const int P = N + 1;
  const float fN = (float)N;
  const unsigned span = (unsigned)(xe - xs);
  float acc = 0.0f;
  int k = k0;
  constexpr int U = 4;
  for (; k + U - 1 <= k1; k += U) {
    float4 q[U];
    float wx[U], wy[U];
    bool on[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int x = k + j - off;
      on[j] = (unsigned)(x - xs) <= span;
      const float u = (float)x - m;
      const float xr = (c * u - sy) + m;  // reference order, one rounding each
      const float yr = (s * u + cy) + m;
      const int cx = (int)__builtin_amdgcn_fmed3f(xr, 0.0f, fN);  // floor, clamped to [0, N]
      const int cyi = (int)__builtin_amdgcn_fmed3f(yr, 0.0f, fN);
      wx[j] = __builtin_amdgcn_fractf(xr);
      wy[j] = __builtin_amdgcn_fractf(yr);
      const unsigned idx = XMAJOR ? (unsigned)(__mul24(cyi, P) + cx) : (unsigned)(__mul24(cx, P) + cyi);
      q[j] = global_f4(cells, idx);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const float t1 = fmaf(wx[j], q[j].w, q[j].z);
      const float t2 = fmaf(wx[j], q[j].y, q[j].x);
      const float v = fmaf(wy[j], t1, t2);
      acc += on[j] ? v : 0.0f;
    }
  }
  for (; k <= k1; ++k) {
    const int x = k - off;
    const bool on = (unsigned)(x - xs) <= span;
    const float u = (float)x - m;
    const float xr = (c * u - sy) + m;
    const float yr = (s * u + cy) + m;
    const int cx = (int)__builtin_amdgcn_fmed3f(xr, 0.0f, fN);
    const int cyi = (int)__builtin_amdgcn_fmed3f(yr, 0.0f, fN);
    const float wx = __builtin_amdgcn_fractf(xr), wy = __builtin_amdgcn_fractf(yr);
    const unsigned idx = XMAJOR ? (unsigned)(__mul24(cyi, P) + cx) : (unsigned)(__mul24(cx, P) + cyi);
    const float4 q = global_f4(cells, idx);
    const float t1 = fmaf(wx, q.w, q.z);
    const float t2 = fmaf(wx, q.y, q.x);
    const float v = fmaf(wy, t1, t2);
    acc += on ? v : 0.0f;
  }
  return acc;
}

__global__ __launch_bounds__(FP_THREADS) void fp_kernel(const FPTask* __restrict__ tasks, int ntasks,
                                                        int base, int nblocks) {
#pragma clang fp contract(off)
  // XCD-aware order: the workgroups an XCD runs together are neighbouring
  // angles of the same ray block, whose rays sweep nearly the same cells in
  // step, so the cell gathers hit that XCD's L2.
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  if (lb >= nblocks) return;
  const int gb = base + lb;
  const FPTask& tk = tasks[find_task(tasks, ntasks, gb)];
  const int N = tk.N;
  const float m = tk.mid;
  const int local = gb - tk.block_begin;
  // chunk-major, then ray block, then angle: the workgroups one XCD runs
  // together are neighbouring angles of one ray block (and one chunk)
  const int chunk = local / (tk.nrb * tk.T);
  const int rem = local - chunk * (tk.nrb * tk.T);
  const int rb = rem / tk.T;
  const int ang = rem - rb * tk.T;
  const int cx0 = chunk * tk.clen + 1, cx1 = min(N, cx0 + tk.clen - 1);  // 1-based steps
  const int ray0 = rb * FP_THREADS + (threadIdx.x & ~63);
  if (ray0 >= N) return;  // whole wave beyond the detector
  const int lane = threadIdx.x & 63;
  const int ray = ray0 + lane;  // 0-based detector bin
  const float2 v = const_f2(tk.sc, ang);
  const float s = v.x, c = v.y;
  const float ym = (float)(min(ray, N - 1) + 1) - m;
  // per-ray products, one rounding each (tv_hpc.F90:143-144)
  const float sy = s * ym;
  const float cy = c * ym;
  // conservative step window: xr, yr in [2, N) (+ margin) -- estimate only
  float ua, ub, va, vb;
  step_range(m - sy, c, 1.99f, (float)N + 0.01f, ua, ub);
  step_range(m + cy, s, 1.99f, (float)N + 0.01f, va, vb);
  float u0 = fmaxf(ua, va), u1 = fminf(ub, vb);
  if (ray >= N) { u0 = 1e30f; u1 = -1e30f; }
  u0 = wave_min(u0);
  u1 = wave_max(u1);
  // 1-based step x = u + m, clipped to [1, N] with one step of margin
  int xs = 1, xe = 0;
  if (u0 <= u1) {
    xs = max(cx0, (int)floorf(fmaxf(u0 + m, -1e9f)) - 1);
    xe = min(cx1, (int)ceilf(fminf(u1 + m, 1e9f)) + 1);
  }
  xs = __builtin_amdgcn_readfirstlane(xs);
  xe = __builtin_amdgcn_readfirstlane(xe);
  const bool xmajor = fabsf(s) >= fabsf(c);
  // skew: keep the minor coordinate of the copy being read constant across lanes
  const float ratio = xmajor ? c / s : -s / c;
  const int off = (int)rintf((float)lane * ratio);
  const int off_end = (int)rintf(63.0f * ratio);
  const int k0 = xs + min(0, off_end), k1 = xe + max(0, off_end);
  float acc = 0.0f;
  if (xs <= xe)
    acc = xmajor ? ray_sum<true>(as_global(tk.cells), xs, xe, off, k0, k1, s, c, sy, cy, m, N)
                 : ray_sum<false>(as_global(tk.cellsT), xs, xe, off, k0, k1, s, c, sy, cy, m, N);
  if (ray < N) tk.partial[((size_t)chunk * tk.T + ang) * N + ray] = acc;
}

// sino(y, t) = (sum over chunks, in step order, of the partial sums) / N
__global__ __launch_bounds__(256) void sum_kernel(const FPTask* __restrict__ tasks, int ntasks, int base) {
  const int gb = base + blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].sum_begin <= gb) lo = mid; else hi = mid - 1;
  }
  const FPTask& tk = tasks[lo];
  const size_t NT = (size_t)tk.N * tk.T;
  const size_t e = (size_t)(gb - tk.sum_begin) * 256 + threadIdx.x;
  if (e >= NT) return;
  const float* p = tk.partial + e;
  float acc = 0.0f;
  for (int c = 0; c < tk.nchunk; ++c) acc += p[(size_t)c * NT];
  tk.sino[e] = acc / (float)tk.N;
}

}  // namespace

int fp_ray_blocks(int N) { return (N + FP_THREADS - 1) / FP_THREADS; }
int fp_chunks(int N, int clen) { return (N + clen - 1) / clen; }
int fp_sum_blocks(int N, int T) { return (int)(((size_t)N * T + 255) / 256); }
int cell_blocks(int N) { const int nb = (N + 1 + CT - 1) / CT; return nb * nb; }

void launch_cells(const CellTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(cells_kernel, dim3(blocks), dim3(256), 0, s, d_tasks, ntasks, base);
}

void launch_forward(const FPTask* d_tasks, int ntasks, int base, int blocks, int sum_base,
                    int sum_blocks, hipStream_t s) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(fp_kernel, dim3(xcd_grid(blocks)), dim3(FP_THREADS), 0, s, d_tasks, ntasks, base,
                     blocks);
  hipLaunchKernelGGL(sum_kernel, dim3(sum_blocks), dim3(256), 0, s, d_tasks, ntasks, sum_base);
}

}  // namespace acu
