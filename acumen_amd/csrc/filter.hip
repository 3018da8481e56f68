// filter.hip -- ramp filtering of sinograms as a Toeplitz GEMM on the
// FP64 matrix cores.
//
// Restates the first half of iradon (tomography/tv_hpc.F90:93-105):
//   sinoCS(i, t) = sum_{j=0..N-1} h[i - j + N] * sino(j, t)
// i.e. C = H * S with H[i][j] = h[i-j+N] an N x N Toeplitz matrix generated
// from the 2N+1 taps and S the N x T sinogram (detector fastest).  The guard
// of :99 never fails for 0 <= i, j < N, so the product is the full one.
//
// Precision.  For a smooth sinogram the sum cancels heavily: h(0) = 2N/pi^2
// times the sample is balanced by the negative taps out to |k| ~ the width
// of the object, so the result is ~N times smaller than its largest terms.
// FP32 accumulation in any order leaves an error of ~eps * h(0) * s, which
// the backprojection then amplifies (measured at 2049^2 x 1500: FP32 ramp
// 2.9e-6 -> FBP 1.8e-5 relative to FP64).  Here the FP32 inputs are widened
// to FP64 exactly, products are exact and the accumulation is FP64 on
// v_mfma_f64_16x16x4_f64 (78.6 TF dense on MI355X), so the only rounding is
// the final conversion of each output to FP32.
//
// Variants (RAMP_*, for the precision study of the filter; FP64 is the
// default): FP32 MFMA (v_mfma_f32_16x16x4_f32, FP32 products and sums),
// BF16x3 (each FP32 operand split into bf16 hi + lo, products hi*hi + hi*lo
// + lo*hi on v_mfma_f32_16x16x32_bf16, FP32 sums) and BF16x6 (hi + mid + lo,
// the six products down to 2^-16 relative, FP32 sums).  The splits are
// formed in registers from the FP32 LDS tiles.
//
// gfx950 mapping: one workgroup = 4 waves = a 64 (detector) x 64 (angle)
// output tile; each wave owns a 32 x 32 block as 2 x 2 MFMA blocks of
// 16 x 16.  The A operand is never materialised: the tile of H needs only
// the BM+BK-1 taps of one diagonal window, staged in LDS once per K step.
// The epilogue goes through LDS so every global store is a contiguous run
// along the detector axis.
#include "acu_internal.hpp"

namespace acu {

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int BM = RF_BM, BN = RF_BN, BK = RF_BK;
constexpr int BKP = BK + 1;        // padded LDS row (angle-major B tile)
constexpr int HW = BM + BK - 1;    // taps per window
constexpr int CP = BM + 1;         // padded LDS row of the epilogue tile

__device__ __forceinline__ int find_task(const RFTask* tasks, int ntasks, int gtile) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_begin <= gtile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int V> struct RampAcc { typedef doublex4 type; };
template <> struct RampAcc<RAMP_FP32> { typedef floatx4 type; };
template <> struct RampAcc<RAMP_BF16X3> { typedef floatx4 type; };
template <> struct RampAcc<RAMP_BF16X6> { typedef floatx4 type; };

// bf16 parts of 8 FP32 values: p0 = bf16(x), p1 = bf16(x - p0), p2 = bf16(x - p0 - p1)
template <int P>
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8 (&out)[P]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    float r = x[e];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const __bf16 q = (__bf16)r;
      out[p][e] = q;
      r = r - (float)q;
    }
  }
}

template <int V>
__global__ __launch_bounds__(RF_THREADS) void ramp_kernel(const RFTask* __restrict__ tasks, int ntasks, int base) {
  typedef typename RampAcc<V>::type accx4;
  __shared__ float hwin[HW + 1];
  __shared__ float smem[BN * CP > BN * BKP ? BN * CP : BN * BKP];

  const int gtile = base + blockIdx.x;
  const RFTask& tk = tasks[find_task(tasks, ntasks, gtile)];
  const int N = tk.N, T = tk.T;
  const int local = gtile - tk.tile_begin;
  const int tm = local % tk.tm, tn = local / tk.tm;
  const int i0 = tm * BM, n0 = tn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lk = lane >> 4;

  accx4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;

  const float* __restrict__ S = tk.sino;
  const float* __restrict__ h = tk.h;
  float* Bs = smem;

  for (int k0 = 0; k0 < N; k0 += BK) {
    // taps for rows i0..i0+BM-1 and columns k0..k0+BK-1: h index i-j+N
    const int hbase = i0 - k0 - (BK - 1) + N;
    if (tid < HW) {
      const int hi = hbase + tid;
      hwin[tid] = (hi >= 0 && hi <= 2 * N) ? h[hi] : 0.0f;
    }
    // B tile: S[k0+k, n0+n] -> Bs[n][k]
#pragma unroll
    for (int q = 0; q < (BK * BN) / RF_THREADS; ++q) {
      const int idx = tid + q * RF_THREADS;
      const int k = idx % BK, n = idx / BK;
      const int gk = k0 + k, gn = n0 + n;
      Bs[n * BKP + k] = (gk < N && gn < T) ? S[(size_t)gn * N + gk] : 0.0f;
    }
    __syncthreads();
    if constexpr (V == RAMP_FP64 || V == RAMP_FP32) {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int kl = 4 * kk + lk;
        float a[2], b[2];
#pragma unroll
        for (int ra = 0; ra < 2; ++ra) a[ra] = hwin[wm * 32 + ra * 16 + l16 - kl + (BK - 1)];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) b[cb] = Bs[(wn * 32 + cb * 16 + l16) * BKP + kl];
#pragma unroll
        for (int ra = 0; ra < 2; ++ra)
#pragma unroll
          for (int cb = 0; cb < 2; ++cb) {
            if constexpr (V == RAMP_FP64)
              acc[ra][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[ra], (double)b[cb], acc[ra][cb], 0, 0, 0);
            else
              acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ra], b[cb], acc[ra][cb], 0, 0, 0);
          }
      }
    } else {
      // one K step of 32: lane supplies k = 8 * (lane / 16) + e, e = 0..7
      constexpr int P = V == RAMP_BF16X3 ? 2 : 3;
      bf16x8 ap[2][P], bp[2][P];
#pragma unroll
      for (int ra = 0; ra < 2; ++ra) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = hwin[wm * 32 + ra * 16 + l16 - (8 * lk + e) + (BK - 1)];
        split8<P>(x, ap[ra]);
      }
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = Bs[(wn * 32 + cb * 16 + l16) * BKP + 8 * lk + e];
        split8<P>(x, bp[cb]);
      }
#pragma unroll
      for (int ra = 0; ra < 2; ++ra)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          // smallest products first
          if constexpr (P == 3) {
            acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][1], bp[cb][1], acc[ra][cb], 0, 0, 0);
            acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][0], bp[cb][2], acc[ra][cb], 0, 0, 0);
            acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][2], bp[cb][0], acc[ra][cb], 0, 0, 0);
          }
          acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][0], bp[cb][1], acc[ra][cb], 0, 0, 0);
          acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][1], bp[cb][0], acc[ra][cb], 0, 0, 0);
          acc[ra][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ra][0], bp[cb][0], acc[ra][cb], 0, 0, 0);
        }
    }
    __syncthreads();
  }

  // epilogue: C block (rows = detector, cols = angle) -> LDS -> coalesced stores.
  // v_mfma_f64_16x16x4_f64 keeps element (i, j) of a 16x16 block in lane
  // j + 16*(i%4), register i/4; v_mfma_f32_16x16x4_f32 and
  // v_mfma_f32_16x16x32_bf16 keep it in lane j + 16*(i/4), register i%4
  // (both measured on gfx950).
  float* Cs = smem;  // [BN][CP], detector fastest
#pragma unroll
  for (int ra = 0; ra < 2; ++ra)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + ra * 16 + (V == RAMP_FP64 ? 4 * r + lk : 4 * lk + r);
        const int n = wn * 32 + cb * 16 + l16;
        Cs[n * CP + i] = (float)acc[ra][cb][r];
      }
  __syncthreads();
  for (int idx = tid; idx < BM * BN; idx += RF_THREADS) {
    const int i = idx % BM, n = idx / BM;
    const int gi = i0 + i, gn = n0 + n;
    if (gi < N && gn < T) tk.out[(size_t)gn * N + gi] = Cs[n * CP + i];
  }
}

}  // namespace

int rf_tiles_m(int N) { return (N + BM - 1) / BM; }
int rf_tiles_n(int T) { return (T + BN - 1) / BN; }

void launch_ramp(const RFTask* d_tasks, int ntasks, int base, int tiles, int variant, hipStream_t s) {
  if (tiles <= 0) return;
  const dim3 g(tiles), t(RF_THREADS);
  switch (variant) {
    case RAMP_FP32: hipLaunchKernelGGL(ramp_kernel<RAMP_FP32>, g, t, 0, s, d_tasks, ntasks, base); break;
    case RAMP_BF16X3: hipLaunchKernelGGL(ramp_kernel<RAMP_BF16X3>, g, t, 0, s, d_tasks, ntasks, base); break;
    case RAMP_BF16X6: hipLaunchKernelGGL(ramp_kernel<RAMP_BF16X6>, g, t, 0, s, d_tasks, ntasks, base); break;
    default: hipLaunchKernelGGL(ramp_kernel<RAMP_FP64>, g, t, 0, s, d_tasks, ntasks, base); break;
  }
}

}  // namespace acu
