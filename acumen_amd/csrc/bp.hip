// bp.hip -- pixel-driven linear-interpolation backprojection on gfx950.
//
// Restates the second half of iradon (tomography/tv_hpc.F90:107-123):
//   yr  = sin(c)*(x-m) + cos(c)*(y-m) + m,   c = pi*(90-theta)/180
//   img(x,y) = cmask(x,y) * sum_t [S(y1,t)*(y2-yr) + S(y2,t)*(yr-y1)]
// with y1 = floor(yr), y2 = y1+1 and the taps used only when y1 > 1 and
// y2 < N+1.  Angles are summed in order, one pixel per accumulator.
//
// gfx950 mapping:
//  * one workgroup = a 64 (x) x 32 (y) pixel tile, one lane per x column,
//    each wave owns 8 rows, so one (sin, cos) pair and one LDS row serve
//    8 updates per lane;
//  * angles are processed in blocks of 32; for each angle the tile only
//    touches a window of <= 73 detector bins, which is staged in LDS as
//    pairs P[t] = (S[lo+t], S[lo+t+1]-S[lo+t]) so an update is one 64-bit
//    LDS read and one FMA:  S(y1) + frac*(S(y2)-S(y1));
//  * a pair whose lower tap y1 violates the reference's bounds test is
//    staged as (0, 0), which reproduces the test exactly with no branch;
//  * positions are taken relative to the tile origin: yr - 1 - lo =
//    s*xl + c*yl + K with xl, yl < 64 and K = s*(x0+1-m) + c*(y0+1-m) + m-1-lo
//    formed in FP64 once per angle and tile, so the FP32 rounding of a
//    position is that of a number below ~140 instead of one near N (the
//    reference rounds s*(x-m) + c*(y-m) + m at magnitude N, ~1e-4 px at
//    N = 2049); each 32-angle block is summed separately and then added;
//  * tiles whose mask is identically zero only store zeros.
#include "acu_internal.hpp"

namespace acu {
namespace {

constexpr int TX = BP_TX, TY = BP_TY, PY = BP_PY, AB = BP_AB, W = BP_W;
constexpr int NWAVES = BP_THREADS / 64;

__device__ __forceinline__ int find_task(const BPTask* tasks, int ntasks, int gtile) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_begin <= gtile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(BP_THREADS) void bp_kernel(const BPTask* __restrict__ tasks, int ntasks, int base) {
  __shared__ float2 P[AB * W];
  __shared__ int lo_s[AB];
  __shared__ float k_s[AB];    // K of the tile-local position (see above)

  const int gtile = base + blockIdx.x;
  const BPTask& tk = tasks[find_task(tasks, ntasks, gtile)];
  const int N = tk.N, T = tk.T;
  const float m = tk.mid;
  const int local = gtile - tk.tile_begin;
  const int tx = local % tk.ntx, ty = local / tk.ntx;
  const int x0 = tx * TX, y0 = ty * TY;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int x = x0 + lane;
  const int yb = y0 + wave * PY;

  if (!tk.tile_active[local]) {
    if (x < N) {
#pragma unroll
      for (int r = 0; r < PY; ++r)
        if (yb + r < N) tk.img[(size_t)(yb + r) * N + x] = 0.0f;
    }
    return;
  }

  // tile-local coordinates, exact in fp32
  const float xl = (float)lane;
  float yl[PY];
#pragma unroll
  for (int r = 0; r < PY; ++r) yl[r] = (float)(wave * PY + r);
  // tile corners (1-based) for the staging windows
  const float cx0 = (float)(x0 + 1) - m, cx1 = (float)(x0 + TX) - m;
  const float cy0 = (float)(y0 + 1) - m, cy1 = (float)(y0 + TY) - m;

  float acc[PY];
#pragma unroll
  for (int r = 0; r < PY; ++r) acc[r] = 0.0f;

  const ACU_GLOBAL float* __restrict__ S = as_global(tk.sino);
  const float2* __restrict__ sc = tk.sc;

  for (int a0 = 0; a0 < T; a0 += AB) {
    const int na = min(AB, T - a0);
    // window start per angle: lowest 0-based detector position over the tile, minus one bin
    if (tid < na) {
      const float2 v = const_f2(sc, a0 + tid);
      const float ex = fminf(v.x * cx0, v.x * cx1);
      const float ey = fminf(v.y * cy0, v.y * cy1);
      const int lo = (int)floorf(ex + ey + m - 1.0f) - 1;
      lo_s[tid] = lo;
      k_s[tid] = (float)((double)v.x * (double)cx0 + (double)v.y * (double)cy0 + (double)m - 1.0 -
                         (double)lo);
    }
    __syncthreads();
    for (int idx = tid; idx < na * W; idx += BP_THREADS) {
      const int a = idx / W, t = idx - a * W;
      const int d = lo_s[a] + t;
      const ACU_GLOBAL float* row = S + (size_t)(a0 + a) * N;
      float2 pv = make_float2(0.0f, 0.0f);
      if (d >= 1 && d <= N - 2) {
        const float s0 = row[d], s1 = row[d + 1];
        pv = make_float2(s0, s1 - s0);
      }
      P[idx] = pv;
    }
    __syncthreads();
    float blk[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) blk[r] = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float2 v = const_f2(sc, a0 + a);
      const float base = fmaf(v.x, xl, k_s[a]);
      const float2* Pa = P + a * W;
#pragma unroll
      for (int r = 0; r < PY; ++r) {
        const float yr = fmaf(v.y, yl[r], base);  // detector position - 1 - lo, in [1, 73)
        const int t = (int)yr;
        const float fr = __builtin_amdgcn_fractf(yr);
        const float2 pv = Pa[t];
        blk[r] += fmaf(fr, pv.y, pv.x);
      }
    }
#pragma unroll
    for (int r = 0; r < PY; ++r) acc[r] += blk[r];
    __syncthreads();
  }

  if (x < N) {
#pragma unroll
    for (int r = 0; r < PY; ++r) {
      const int y = yb + r;
      if (y < N) {
        const size_t o = (size_t)y * N + x;
        tk.img[o] = tk.cmask[o] * acc[r];
      }
    }
  }
}

}  // namespace

int bp_tiles_x(int N) { return (N + TX - 1) / TX; }
int bp_tiles_y(int N) { return (N + TY - 1) / TY; }

void launch_backproject(const BPTask* d_tasks, int ntasks, int base, int tiles, hipStream_t s) {
  if (tiles <= 0) return;
  hipLaunchKernelGGL(bp_kernel, dim3(tiles), dim3(BP_THREADS), 0, s, d_tasks, ntasks, base);
}

}  // namespace acu
