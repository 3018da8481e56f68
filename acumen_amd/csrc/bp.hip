// bp.hip -- pixel-driven linear-interpolation backprojection on gfx950.
//
// Restates the second half of iradon (tomography/tv_hpc.F90:107-123):
//   yr  = sin(c)*(x-m) + cos(c)*(y-m) + m,   c = pi*(90-theta)/180
//   img(x,y) = cmask(x,y) * sum_t [S(y1,t)*(y2-yr) + S(y2,t)*(yr-y1)]
// with y1 = floor(yr), y2 = y1+1 and the taps used only when y1 > 1 and
// y2 < N+1.  Angles are summed in order, one pixel per accumulator.
//
// gfx950 mapping:
//  * one workgroup = a 64 (x) x TY (y) pixel tile, TY = 16, 32 or 64 (a
//    template parameter the autotuner picks per batch), one lane per x
//    column; 4 waves (8 for TY = 64) own 8 rows each (4 for TY = 16), so
//    one (sin, cos) pair and one LDS row serve that many updates per lane;
//  * angles are processed in blocks of 32; for each angle the tile only
//    touches a window of W detector bins (72 / 80 / 96), which is staged in
//    LDS as pairs P[t] = (S[lo+t], S[lo+t+1]-S[lo+t]) so an update is one
//    64-bit LDS read and one FMA:  S(y1) + frac*(S(y2)-S(y1));
//  * staging: 8 (16) threads per angle, each loading W/8 (W/16) pairs at
//    clamped (always in-image) positions with every load issued before the
//    first is used;
//  * a pair whose lower tap y1 violates the reference's bounds test is
//    staged as (0, 0), which reproduces the test exactly with no branch;
//  * positions (REFPOS, the default): the reference's own FP32 value,
//    yr = (s*(x-m) + c*(y-m)) + m with each product and sum rounded once and
//    no contraction, so every floor, weight and bounds decision is the
//    reference's.  Its rounding at magnitude N (~1e-4 px at N = 2049) is
//    what dominates the reference's distance from FP64 on noisy sinograms:
//    with these positions the FBP is ~1e-6 from the reference's FP32 FBP,
//    with exact positions ~2.7e-5 (noisy 2049^2 x 1500);
//  * positions (!REFPOS, ACUMEN_BP_POSITIONS=tile): relative to the tile
//    origin, yr - 1 - lo = s*xl + c*yl + K with xl < 64, yl < TY and
//    K = s*(x0+1-m) + c*(y0+1-m) + m-1-lo formed in FP64 once per angle and
//    tile, so a position is rounded at magnitude below ~100: closer to FP64,
//    farther from the reference, two VALU fewer per update;
//  * each 32-angle block is summed separately and then added;
//  * tiles whose mask is identically zero (by 64 x 16 sub-tiles) only store
//    zeros.
#include "acu_internal.hpp"

#include <algorithm>

namespace acu {
namespace {

constexpr int TX = BP_TX, AB = BP_AB;
constexpr int bp_waves(int ty) { return ty == 64 ? 8 : 4; }

// Detector bins staged per angle at tile height ty.  Over a 64 x ty tile the
// positions span at most hypot(63, ty-1) bins; the window starts at least
// one bin (less FP32 rounding) below the lowest position, and a sample at
// local position t reads pair t, so W >= hypot(63, ty-1) + 3.01 suffices
// (67.8, 73.2, 92.1); W is a multiple of the staging threads per angle.
constexpr int bp_window(int ty) { return ty == 16 ? 72 : ty == 32 ? 80 : 96; }

__device__ __forceinline__ int find_task(const BPTask* tasks, int ntasks, int gtile) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_begin <= gtile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int TY, bool REFPOS>
__global__ __launch_bounds__(64 * bp_waves(TY)) void bp_kernel(const BPTask* __restrict__ tasks, int ntasks, int base,
                                                               BPShares sh) {
  constexpr int NW = bp_waves(TY), NT = 64 * NW, PY = TY / NW, W = bp_window(TY), SUBS = TY / BP_SUB;
  constexpr int SPA = NT / AB, NST = W / SPA;  // staging threads per angle, pairs per thread
  static_assert(TY % BP_SUB == 0 && TY % NW == 0 && NT % AB == 0 && W % SPA == 0, "tile shape");
  __shared__ float2 P[AB * W];
  __shared__ int lo_s[AB];
  __shared__ float k_s[AB];    // K of the tile-local position (see above)

  // XCD k (xcd_remap) runs the tiles of its share in order
  const int l = xcd_remap(blockIdx.x, gridDim.x);
  const int k = l / sh.C;
  const int lt = sh.b[k] + (l - k * sh.C);
  if (lt >= sh.b[k + 1]) return;
  const int gtile = base + lt;
  const BPTask& tk = tasks[find_task(tasks, ntasks, gtile)];
  const int N = tk.N, T = tk.T;
  const float m = tk.mid;
  const int local = gtile - tk.tile_begin;
  const int tx = local % tk.ntx, tyi = local / tk.ntx;
  const int x0 = tx * TX, y0 = tyi * TY;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int x = x0 + lane;
  const int yb = y0 + wave * PY;

  bool active = false;
#pragma unroll
  for (int j = 0; j < SUBS; ++j) {
    const int r16 = tyi * SUBS + j;
    if (r16 < tk.nsub) active = active || tk.tile_active[r16 * tk.ntx + tx] != 0;
  }
  if (!active) {
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
  // the reference's jx - midpoint (1-based), exact in fp32
  const float ux = (float)(x + 1) - m;
  // tile corners (1-based) for the staging windows
  const float cx0 = (float)(x0 + 1) - m, cx1 = (float)(x0 + TX) - m;
  const float cy0 = (float)(y0 + 1) - m, cy1 = (float)(y0 + TY) - m;

  float acc[PY];
#pragma unroll
  for (int r = 0; r < PY; ++r) acc[r] = 0.0f;

  const ACU_GLOBAL float* __restrict__ S = as_global(tk.sino);
  const float2* __restrict__ sc = tk.sc;
  const int sa = tid / SPA, sj = tid % SPA;  // staging: angle of the block, first bin

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
    if (sa < na) {
      const int lo = lo_s[sa];
      const ACU_GLOBAL float* row = S + (size_t)(a0 + sa) * N;
      float2 v[NST];
#pragma unroll
      for (int k = 0; k < NST; ++k) {
        const int dc = min(max(lo + sj + SPA * k, 1), N - 2);  // N >= 3: row[dc + 1] is in the row
        v[k] = make_float2(row[dc], row[dc + 1]);
      }
#pragma unroll
      for (int k = 0; k < NST; ++k) {
        const int d = lo + sj + SPA * k;
        P[sa * W + sj + SPA * k] = (d >= 1 && d <= N - 2) ? make_float2(v[k].x, v[k].y - v[k].x)
                                                          : make_float2(0.0f, 0.0f);
      }
    }
    __syncthreads();
    float blk[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) blk[r] = 0.0f;
    if (REFPOS) {
      // the wave's rows of the row-term table: wave-uniform, scalar loads
      const int ybu = __builtin_amdgcn_readfirstlane(yb);
      const ACU_CONST float* rt = as_const(tk.rowterm) + (size_t)a0 * N + ybu;
      for (int a = 0; a < na; ++a, rt += N) {
#pragma clang fp contract(off)
        const float sa = const_f2(sc, a0 + a).x;
        const float A = sa * ux;  // sin(c)*(jx-midpoint), tv_hpc.F90:114
        // byte offset of P[a*W - 1 - lo] (wave-uniform): the pair of the
        // 1-based tap y1 is at Pb + 8*y1
        const char* Pb = (const char*)P + __builtin_amdgcn_readfirstlane(8 * (a * W - 1 - lo_s[a]));
#pragma unroll
        for (int r = 0; r < PY; ++r) {
          const float yr = (A + rt[r]) + m;  // (sin(c)*(jx-m) + cos(c)*(jy-m)) + m
          // y1 = floor(yr); truncation differs only for yr < 0, whose pairs
          // (taps below 2) are staged as zeros either way
          const int y1 = (int)yr;
          const float fr = __builtin_amdgcn_fractf(yr);  // yr - floor(yr) = yr - y1, exact
          const float2 pv = *(const float2*)(Pb + (y1 << 3));
          blk[r] += fmaf(fr, pv.y, pv.x);
        }
      }
    } else {
      for (int a = 0; a < na; ++a) {
        const float2 v = const_f2(sc, a0 + a);
        const float base = fmaf(v.x, xl, k_s[a]);
        const float2* Pa = P + a * W;
#pragma unroll
        for (int r = 0; r < PY; ++r) {
          const float yr = fmaf(v.y, yl[r], base);  // detector position - 1 - lo, in (0, W - 1)
          const int t = (int)yr;
          const float fr = __builtin_amdgcn_fractf(yr);
          const float2 pv = Pa[t];
          blk[r] += fmaf(fr, pv.y, pv.x);
        }
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

__global__ __launch_bounds__(256) void rowterm_kernel(const float2* __restrict__ sc, int N, int T, float mid,
                                                      float* __restrict__ out) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)N * T + BP_ROWTERM_PAD) return;
  const int t = (int)(i / N), y = (int)(i - (size_t)t * N);
  out[i] = t < T ? sc[t].y * ((float)(y + 1) - mid) : 0.0f;
}

void launch_bp_rowterm(const float2* sc, int N, int T, float mid, float* rowterm, hipStream_t s) {
  const size_t n = (size_t)N * T + BP_ROWTERM_PAD;
  hipLaunchKernelGGL(rowterm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sc, N, T, mid, rowterm);
}

int bp_tiles_x(int N) { return (N + TX - 1) / TX; }
int bp_tiles_y(int N, int ty) { return (N + ty - 1) / ty; }

template <bool REFPOS>
void launch_bp(const BPTask* d_tasks, int ntasks, int base, int ty, const BPShares& sh, hipStream_t s) {
  const dim3 g((unsigned)(XCDS * sh.C));
  switch (ty) {
    case 16: hipLaunchKernelGGL((bp_kernel<16, REFPOS>), g, dim3(64 * bp_waves(16)), 0, s, d_tasks, ntasks, base, sh); break;
    case 32: hipLaunchKernelGGL((bp_kernel<32, REFPOS>), g, dim3(64 * bp_waves(32)), 0, s, d_tasks, ntasks, base, sh); break;
    case 64: hipLaunchKernelGGL((bp_kernel<64, REFPOS>), g, dim3(64 * bp_waves(64)), 0, s, d_tasks, ntasks, base, sh); break;
    default: throw Error("backprojection tile height must be 16, 32 or 64");
  }
}

void launch_backproject(const BPTask* d_tasks, int ntasks, int base, int tiles, int ty, bool refpos,
                        const BPShares& sh, hipStream_t s) {
  if (tiles <= 0) return;
  if (sh.b[0] != 0 || sh.b[XCDS] != tiles || sh.C < 1) throw Error("backprojection shares do not match the launch");
  if (refpos) launch_bp<true>(d_tasks, ntasks, base, ty, sh, s);
  else launch_bp<false>(d_tasks, ntasks, base, ty, sh, s);
}

// Cut points at k/XCDS of the total estimated work (a tile starts the next
// share once the work before it reaches the share's target), so the shares
// are contiguous runs of the launch's tiles with about equal work; with 8
// equal slices each XCD gets one slice.
BPShares bp_xcd_shares(const std::vector<float>& w) {
  BPShares sh{};
  const int n = (int)w.size();
  double tot = 0.0;
  for (float v : w) tot += v;
  int k = 1;
  double acc = 0.0;
  for (int i = 0; i < n && k < XCDS; ++i) {
    while (k < XCDS && acc >= tot * k / XCDS - 1e-9 * tot) sh.b[k++] = i;
    acc += w[i];
  }
  while (k < XCDS) sh.b[k++] = n;
  sh.b[0] = 0;
  sh.b[XCDS] = n;
  sh.C = 1;
  for (int j = 0; j < XCDS; ++j) sh.C = std::max(sh.C, sh.b[j + 1] - sh.b[j]);
  return sh;
}

}  // namespace acu
