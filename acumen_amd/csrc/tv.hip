// tv.hip -- fused elementwise updates and FP64 reductions of the
// split-Bregman TV solver (tomography/tv_hpc.F90:162-363) on gfx950.
//
// The reference runs every loop nest as its own kernel and returns each FP32
// reduction to the host (12 host round trips per inner iteration at ncg=5).
// Here each kernel fuses an update with the reduction that consumes it, the
// reductions are FP64 (wavefront shuffles, then LDS, then one workgroup per
// slice sums the block partials in a fixed order), and every scalar
// (rsold, p.Ap, rsnew, alpha, beta, Err_diff) lives in device memory, so an
// outer iteration needs no host synchronisation and can be replayed from a
// HIP graph.  All kernels are batched: blocks are flattened over slices.
//
// Scalar slots per slice (doubles): rs[j] at j (j = 0..ncg), pAp[j] at
// 64+j, outer-iteration counter at 126, Err_diff at 127.  Every reduction
// kernel writes block partials and is followed by finalize_kernel.
#include "acu_tv.hpp"

namespace acu {
namespace {

constexpr int NT_ = EW_THREADS;
constexpr int IT = EW_ITEMS;
constexpr int PER_BLOCK = NT_ * IT;

__device__ __forceinline__ int find_task_e(const TVTask* tasks, int ntasks, int gb) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].eb_begin <= gb) lo = mid; else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ int find_task_s(const TVTask* tasks, int ntasks, int gb) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].sb_begin <= gb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-reduce v and write this block's partial; finalize_kernel (launched
// next on the same stream) sums a slice's partials in index order.  No
// device-scope fence per block: on a multi-XCD part each such fence writes
// back the L2, which made these kernels ~10x slower than their traffic.
__device__ void reduce_to_partial(double v, double* partials, int lblock) {
  __shared__ double wsum[NT_ / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
#pragma unroll
    for (int w = 0; w < NT_ / 64; ++w) b += wsum[w];
    partials[lblock] = b;
  }
}

__device__ __forceinline__ float lap(const float* __restrict__ u, int x, int y, int N) {
  const int xl = x == 0 ? N - 1 : x - 1, xr = x == N - 1 ? 0 : x + 1;
  const int yl = y == 0 ? N - 1 : y - 1, yr = y == N - 1 ? 0 : y + 1;
  const size_t row = (size_t)y * N;
  float l = 4.0f * u[row + x] - u[row + xl];
  l = l - u[row + xr];
  l = l - u[(size_t)yl * N + x];
  l = l - u[(size_t)yr * N + x];
  return l;
}

// K9 + K10 + K11 (tv_hpc.F90:207-249): ukp = uk; rhs from d, b; r = rhs -
// mu*iukt - lambda*L(uk); p = r; rs[0] = sum r^2.  LIN: ifkt and iukt are
// the linear schedule's FP64 carried terms, rounded to FP32 here.
template <bool LIN>
__global__ __launch_bounds__(NT_) void residual_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int N = tk.N;
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)N * N;
  const float mu = tk.mu, lam = tk.lambda;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      const int y = (int)(e / N), x = (int)(e - (size_t)y * N);
      const int xn = x == N - 1 ? 0 : x + 1, yn = y == N - 1 ? 0 : y + 1;
      const size_t ex = (size_t)y * N + xn, ey = (size_t)yn * N + x;
      float t = tk.dkx[ex] - tk.bkx[ex];
      t = t - tk.dkx[e];
      t = t + tk.bkx[e];
      t = t + tk.dky[ey];
      t = t - tk.bky[ey];
      t = t - tk.dky[e];
      t = t + tk.bky[e];
      const float ifkt = LIN ? (float)tk.ifkt64[e] : tk.ifkt[e];
      const float iukt = LIN ? (float)tk.iukt64[e] : tk.iukt[e];
      const float rhs = mu * ifkt + lam * t;
      const float rv = (rhs - mu * iukt) - lam * lap(tk.uk, x, y, N);
      tk.r[e] = rv;
      tk.p[e] = rv;
      tk.ukp[e] = tk.uk[e];
      acc += (double)rv * (double)rv;
    }
  }
  reduce_to_partial(acc, tk.partials, lb);
}

// K12 + K13 (tv_hpc.F90:257-274): Ap = mu*irpt + lambda*L(p); pAp[j] = p.Ap
__global__ __launch_bounds__(NT_) void apdot_kernel(const TVTask* __restrict__ tasks, int ntasks, int base, int j) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int N = tk.N;
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)N * N;
  const float mu = tk.mu, lam = tk.lambda;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      const int y = (int)(e / N), x = (int)(e - (size_t)y * N);
      const float a = mu * tk.irpt[e] + lam * lap(tk.p, x, y, N);
      tk.Ap[e] = a;
      acc += (double)tk.p[e] * (double)a;
    }
  }
  reduce_to_partial(acc, tk.partials, lb);
}

// K14 + K15 (tv_hpc.F90:277-292): alpha = rs[j]/pAp[j] (FP64);
// uk += alpha p; r -= alpha Ap; rs[j+1] = sum r^2.
// A zero p.Ap (all-zero residual) gives alpha = 0 instead of the
// reference's 0/0 = NaN.
// LIN (linear schedule): iukt = F(R(uk)) is carried along by linearity in
// FP64, iukt64 += alpha irpt with irpt = F(R(p)), instead of being
// recomputed from uk by a projection and a backprojection.
__device__ __forceinline__ float cg_alpha(const double* scal, int j) {
  const double num = scal[64 + j];
  return num != 0.0 ? (float)(scal[j] / num) : 0.0f;
}

template <bool LIN>
__global__ __launch_bounds__(NT_) void axpy_kernel(const TVTask* __restrict__ tasks, int ntasks, int base, int j) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int N = tk.N;
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)N * N;
  const float alpha = cg_alpha(tk.scal, j);
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      tk.uk[e] = tk.uk[e] + alpha * tk.p[e];
      const float rv = tk.r[e] - alpha * tk.Ap[e];
      tk.r[e] = rv;
      if (LIN) tk.iukt64[e] += (double)alpha * (double)tk.irpt[e];
      acc += (double)rv * (double)rv;
    }
  }
  reduce_to_partial(acc, tk.partials, lb);
}

// Linear schedule, sinogram part of K14: ruk = R(uk) is carried along by
// linearity in FP64, ruk64 += alpha rp with rp = R(p), so the add-back of
// tv_hpc.F90:339-346 needs no projection of the new uk.
__global__ __launch_bounds__(NT_) void sino_axpy_kernel(const TVTask* __restrict__ tasks, int ntasks, int base,
                                                        int j) {
  const TVTask& tk = tasks[find_task_s(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.sb_begin;
  const size_t NT = (size_t)tk.N * tk.T;
  const float alpha = cg_alpha(tk.scal, j);
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NT) tk.ruk64[e] += (double)alpha * (double)tk.rp[e];
  }
}

// K16 (tv_hpc.F90:294-299): p = r + (rs[j+1]/rs[j]) p
__global__ __launch_bounds__(NT_) void pupdate_kernel(const TVTask* __restrict__ tasks, int ntasks, int base, int j) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int N = tk.N;
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)N * N;
  const double rs = tk.scal[j];
  const float beta = rs != 0.0 ? (float)(tk.scal[j + 1] / rs) : 0.0f;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) tk.p[e] = tk.r[e] + beta * tk.p[e];
  }
}

// One workgroup per slice: scal[slot] = sum of the slice's block partials,
// in a fixed order (thread t sums a contiguous range, then a tree).
__global__ __launch_bounds__(NT_) void finalize_kernel(const TVTask* __restrict__ tasks, int first,
                                                       int slot) {
  const TVTask& tk = tasks[first + blockIdx.x];
  __shared__ double wsum[NT_ / 64];
  const int nb = tk.nb;
  const int chunk = (nb + NT_ - 1) / NT_;
  const int b0 = threadIdx.x * chunk;
  double s = 0.0;
  for (int b = b0; b < min(nb, b0 + chunk); ++b) s += tk.partials[b];
  s = wave_sum(s);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < NT_ / 64; ++w) tot += wsum[w];
    tk.scal[slot] = tot;
  }
}

__device__ __forceinline__ void shrink(float t, float inv_lambda, float& d, float& b) {
  if (t == 0.0f) {
    d = 0.0f;
  } else {
    const float at = fabsf(t);
    const float a = fmaxf(at - inv_lambda, 0.0f);
    d = (a * t) / at;
  }
  b = t - d;
}

// K17 + K18 (tv_hpc.F90:305-333): anisotropic shrink of D u + b at 1/lambda,
// Err_diff = sum |uk - ukp| (also appended to the per-slice delta log).
// UPD (linear schedule, last inner iteration): the next outer iteration's
// ifkt = F(fk + f - ruk) = ifkt + F(f) - F(ruk) = (ifkt + fbp) - iukt by
// linearity of F (FP64 carried terms), instead of the backprojection of
// tv_hpc.F90:215.
template <bool UPD>
__global__ __launch_bounds__(NT_) void shrink_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int N = tk.N;
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)N * N;
  const float inv_lambda = 1.0f / tk.lambda;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      const int y = (int)(e / N), x = (int)(e - (size_t)y * N);
      const int xp = x == 0 ? N - 1 : x - 1, yp = y == 0 ? N - 1 : y - 1;
      const float u = tk.uk[e];
      float d, b;
      shrink((tk.uk[(size_t)y * N + xp] - u) + tk.bkx[e], inv_lambda, d, b);
      tk.dkx[e] = d;
      tk.bkx[e] = b;
      shrink((tk.bky[e] + tk.uk[(size_t)yp * N + x]) - u, inv_lambda, d, b);
      tk.dky[e] = d;
      tk.bky[e] = b;
      if (UPD) tk.ifkt64[e] = (tk.ifkt64[e] + (double)tk.fbp[e]) - tk.iukt64[e];
      acc += (double)fabsf(u - tk.ukp[e]);
    }
  }
  reduce_to_partial(acc, tk.partials, lb);
}

// K19 (tv_hpc.F90:341-346): fk = (fk + f) - ruk, plus the delta log entry.
// LIN: ruk is the linear schedule's FP64 carried term.
template <bool LIN>
__global__ __launch_bounds__(NT_) void addback_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_s(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.sb_begin;
  const size_t NT = (size_t)tk.N * tk.T;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NT) {
      if (LIN) tk.fk[e] = (float)(((double)tk.fk[e] + (double)tk.f[e]) - tk.ruk64[e]);
      else tk.fk[e] = (tk.fk[e] + tk.f[e]) - tk.ruk[e];
    }
  }
  if (lb == 0 && threadIdx.x == 0) {
    const int j = (int)tk.scal[126];
    if (tk.delta_log != nullptr && j < tk.delta_cap) tk.delta_log[j] = (float)tk.scal[127];
    tk.scal[126] = (double)(j + 1);
  }
}

// Linear schedule: the FP64 carried terms from exact FP32 operator results
// (image part: ifkt, iukt; sinogram part: ruk).
__global__ __launch_bounds__(NT_) void linload_img_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)tk.N * tk.N;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      tk.ifkt64[e] = (double)tk.ifkt[e];
      tk.iukt64[e] = (double)tk.iukt[e];
    }
  }
}
__global__ __launch_bounds__(NT_) void linload_sino_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_s(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.sb_begin;
  const size_t NT = (size_t)tk.N * tk.T;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NT) tk.ruk64[e] = (double)tk.ruk[e];
  }
}

// tvopt init (tv_hpc.F90:187-202), image part: d = b = 0.
__global__ __launch_bounds__(NT_) void tvinit_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_e(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.eb_begin;
  const size_t NN = (size_t)tk.N * tk.N;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NN) {
      tk.dkx[e] = 0.0f; tk.dky[e] = 0.0f; tk.bkx[e] = 0.0f; tk.bky[e] = 0.0f;
    }
  }
  if (lb == 0 && threadIdx.x == 0) tk.scal[126] = 0.0;  // outer-iteration counter
}

// tvopt init, sinogram part: fk = f.
__global__ __launch_bounds__(NT_) void fkinit_kernel(const TVTask* __restrict__ tasks, int ntasks, int base) {
  const TVTask& tk = tasks[find_task_s(tasks, ntasks, base + blockIdx.x)];
  const int lb = base + blockIdx.x - tk.sb_begin;
  const size_t NT = (size_t)tk.N * tk.T;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t e = (size_t)lb * PER_BLOCK + it * NT_ + threadIdx.x;
    if (e < NT) tk.fk[e] = tk.f[e];
  }
}

}  // namespace

int ew_blocks(size_t n) { return (int)((n + PER_BLOCK - 1) / PER_BLOCK); }

void launch_tv_finalize(const TVTask* t, int first, int count, int slot, hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(count), dim3(NT_), 0, s, t, first, slot);
}
void launch_tv_residual(const TVTask* t, int n, int base, int blocks, bool lin, hipStream_t s) {
  if (lin) hipLaunchKernelGGL(residual_kernel<true>, dim3(blocks), dim3(NT_), 0, s, t, n, base);
  else hipLaunchKernelGGL(residual_kernel<false>, dim3(blocks), dim3(NT_), 0, s, t, n, base);
}
void launch_tv_apdot(const TVTask* t, int n, int base, int blocks, int j, hipStream_t s) {
  hipLaunchKernelGGL(apdot_kernel, dim3(blocks), dim3(NT_), 0, s, t, n, base, j);
}
void launch_tv_axpy(const TVTask* t, int n, int base, int blocks, int j, bool lin, hipStream_t s) {
  if (lin) hipLaunchKernelGGL(axpy_kernel<true>, dim3(blocks), dim3(NT_), 0, s, t, n, base, j);
  else hipLaunchKernelGGL(axpy_kernel<false>, dim3(blocks), dim3(NT_), 0, s, t, n, base, j);
}
void launch_tv_sino_axpy(const TVTask* t, int n, int base, int sblocks, int j, hipStream_t s) {
  hipLaunchKernelGGL(sino_axpy_kernel, dim3(sblocks), dim3(NT_), 0, s, t, n, base, j);
}
void launch_tv_pupdate(const TVTask* t, int n, int base, int blocks, int j, hipStream_t s) {
  hipLaunchKernelGGL(pupdate_kernel, dim3(blocks), dim3(NT_), 0, s, t, n, base, j);
}
void launch_tv_shrink(const TVTask* t, int n, int base, int blocks, bool upd_ifkt, hipStream_t s) {
  if (upd_ifkt) hipLaunchKernelGGL(shrink_kernel<true>, dim3(blocks), dim3(NT_), 0, s, t, n, base);
  else hipLaunchKernelGGL(shrink_kernel<false>, dim3(blocks), dim3(NT_), 0, s, t, n, base);
}
void launch_tv_addback(const TVTask* t, int n, int base, int sblocks, bool lin, hipStream_t s) {
  if (lin) hipLaunchKernelGGL(addback_kernel<true>, dim3(sblocks), dim3(NT_), 0, s, t, n, base);
  else hipLaunchKernelGGL(addback_kernel<false>, dim3(sblocks), dim3(NT_), 0, s, t, n, base);
}
void launch_tv_linload(const TVTask* t, int n, int base, int blocks, int sbase, int sblocks, hipStream_t s) {
  hipLaunchKernelGGL(linload_img_kernel, dim3(blocks), dim3(NT_), 0, s, t, n, base);
  hipLaunchKernelGGL(linload_sino_kernel, dim3(sblocks), dim3(NT_), 0, s, t, n, sbase);
}
void launch_tv_init(const TVTask* t, int n, int base, int blocks, int sbase, int sblocks,
                    hipStream_t s) {
  hipLaunchKernelGGL(tvinit_kernel, dim3(blocks), dim3(NT_), 0, s, t, n, base);
  hipLaunchKernelGGL(fkinit_kernel, dim3(sblocks), dim3(NT_), 0, s, t, n, sbase);
}

}  // namespace acu
