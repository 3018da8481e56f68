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
// gfx950 mapping: one workgroup = 4 waves = a 64 (detector) x 64 (angle)
// output tile; each wave owns a 32 x 32 block as 2 x 2 MFMA blocks of
// 16 x 16.  The A operand is never materialised: the tile of H needs only
// the BM+BK-1 taps of one diagonal window, staged in LDS once per K step.
// The epilogue goes through LDS so every global store is a contiguous run
// along the detector axis.
#include "acu_internal.hpp"

namespace acu {

typedef double doublex4 __attribute__((ext_vector_type(4)));

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

__global__ __launch_bounds__(RF_THREADS) void ramp_kernel(const RFTask* __restrict__ tasks, int ntasks, int base) {
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

  doublex4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;

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
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kl = 4 * kk + lk;
      double a[2], b[2];
#pragma unroll
      for (int ra = 0; ra < 2; ++ra) a[ra] = (double)hwin[wm * 32 + ra * 16 + l16 - kl + (BK - 1)];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) b[cb] = (double)Bs[(wn * 32 + cb * 16 + l16) * BKP + kl];
#pragma unroll
      for (int ra = 0; ra < 2; ++ra)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
          acc[ra][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ra], b[cb], acc[ra][cb], 0, 0, 0);
    }
    __syncthreads();
  }

  // epilogue: C block (rows = detector, cols = angle) -> LDS -> coalesced stores.
  // v_mfma_f64_16x16x4_f64 keeps element (i, j) of a 16x16 block in lane
  // j + 16*(i%4), register i/4 (measured on gfx950; the FP32 16x16 variants
  // use rows 4*(lane/16) + r instead).
  float* Cs = smem;  // [BN][CP], detector fastest
#pragma unroll
  for (int ra = 0; ra < 2; ++ra)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + ra * 16 + 4 * r + lk;
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

void launch_ramp(const RFTask* d_tasks, int ntasks, int base, int tiles, hipStream_t s) {
  if (tiles <= 0) return;
  hipLaunchKernelGGL(ramp_kernel, dim3(tiles), dim3(RF_THREADS), 0, s, d_tasks, ntasks, base);
}

}  // namespace acu
