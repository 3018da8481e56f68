// fft.hip -- the ramp filter (tomography/tv_hpc.F90:93-105) by FFT
// convolution, the RAMP_FFT variant of the filter's precision study.
//
//   sinoCS(i, t) = sum_{j=0..N-1} g(i - j) sino(j, t),   g(m) = h[m + N],
//
// is a linear convolution with the symmetric kernel g on |m| <= N-1.  With
// L a power of two >= 2N-2 it equals the circular convolution of length L
// with gw(r) = sum_{m = r mod L} g(m), except when L = 2N-2, where m = N-1
// and m = -(N-1) share a residue: then outputs 0 and N-1 pick up one extra
// term each, g(N-1) s(N-1) and g(N-1) s(0), which are subtracted.  G =
// DFT(gw) is real and even (g is symmetric), so two real rows a, b ride in
// one complex FFT: IDFT(G * DFT(a + i b)) = (a * g) + i (b * g).
//
// gfx950 mapping: one workgroup per pair of angle rows; the L complex
// values live in LDS (L <= 8192, 64 KB).  The forward transform is a
// radix-2 decimation in time over bit-reversed input (natural-order
// output); after the multiplication by G the inverse is a radix-2
// decimation in frequency (bit-reversed output), read back through the bit
// reversal, so no permutation pass is needed.  G and the twiddles are
// computed on the host in FP64 from the reference's FP32 taps; the
// transforms run in FP32.
#include "acu_internal.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace acu {
namespace {

constexpr int FFT_THREADS = 256;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ int find_task(const FFTTask* tasks, int ntasks, int gb) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block_begin <= gb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int bitrev(int i, int lg) { return (int)(__brev((unsigned)i) >> (32 - lg)); }

__global__ __launch_bounds__(FFT_THREADS) void fft_ramp_kernel(const FFTTask* __restrict__ tasks, int ntasks,
                                                               int base, int nblocks) {
  extern __shared__ float2 X[];
  const int lbk = xcd_remap(blockIdx.x, gridDim.x);
  if (lbk >= nblocks) return;  // uniform per workgroup
  const int gb = base + lbk;
  const FFTTask& tk = tasks[find_task(tasks, ntasks, gb)];
  const int N = tk.N, T = tk.T, L = tk.L, lg = tk.lg;
  const int ra = 2 * (gb - tk.block_begin), rb = ra + 1;
  const bool hasb = rb < T;
  const float* sa = tk.sino + (size_t)ra * N;
  const float* sb = tk.sino + (size_t)(hasb ? rb : ra) * N;
  const float2* __restrict__ W = tk.tw;  // exp(-2 pi i k / L), k < L/2

  // bit-reversed load: X[bitrev(j)] = a(j) + i b(j), zero beyond N
  for (int j = threadIdx.x; j < L; j += FFT_THREADS) {
    float2 v = make_float2(0.0f, 0.0f);
    if (j < N) v = make_float2(sa[j], hasb ? sb[j] : 0.0f);
    X[bitrev(j, lg)] = v;
  }
  __syncthreads();
  // forward: decimation in time
  for (int m = 1, ts = L / 2; m < L; m <<= 1, ts >>= 1) {
    for (int k = threadIdx.x; k < L / 2; k += FFT_THREADS) {
      const int pos = k & (m - 1);
      const int i0 = ((k - pos) << 1) + pos, i1 = i0 + m;
      const float2 w = W[pos * ts];
      const float2 x1 = X[i1], x0 = X[i0];
      const float2 t = make_float2(w.x * x1.x - w.y * x1.y, w.x * x1.y + w.y * x1.x);
      X[i0] = make_float2(x0.x + t.x, x0.y + t.y);
      X[i1] = make_float2(x0.x - t.x, x0.y - t.y);
    }
    __syncthreads();
  }
  // filter: the spectrum of the kernel is real; 1/L folded in
  for (int k = threadIdx.x; k < L; k += FFT_THREADS) {
    const float g = tk.G[k];
    const float2 v = X[k];
    X[k] = make_float2(v.x * g, v.y * g);
  }
  __syncthreads();
  // inverse: decimation in frequency with conjugate twiddles
  for (int m = L / 2, ts = 1; m >= 1; m >>= 1, ts <<= 1) {
    for (int k = threadIdx.x; k < L / 2; k += FFT_THREADS) {
      const int pos = k & (m - 1);
      const int i0 = ((k - pos) << 1) + pos, i1 = i0 + m;
      const float2 w = W[pos * ts];
      const float2 x0 = X[i0], x1 = X[i1];
      const float2 d = make_float2(x0.x - x1.x, x0.y - x1.y);
      X[i0] = make_float2(x0.x + x1.x, x0.y + x1.y);
      X[i1] = make_float2(w.x * d.x + w.y * d.y, w.x * d.y - w.y * d.x);
    }
    __syncthreads();
  }
  float* oa = tk.out + (size_t)ra * N;
  float* ob = tk.out + (size_t)rb * N;
  for (int i = threadIdx.x; i < N; i += FFT_THREADS) {
    float2 c = X[bitrev(i, lg)];
    if (tk.alias) {  // L = 2N-2: remove the wrapped term of the shared residue
      if (i == N - 1) { c.x -= tk.g_edge * sa[0]; c.y -= hasb ? tk.g_edge * sb[0] : 0.0f; }
      if (i == 0) { c.x -= tk.g_edge * sa[N - 1]; c.y -= hasb ? tk.g_edge * sb[N - 1] : 0.0f; }
    }
    oa[i] = c.x;
    if (hasb) ob[i] = c.y;
  }
}

void fft_host(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * kPi / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

}  // namespace

int fft_len(int N) {
  int L = 1;
  while (L < 2 * N - 2) L <<= 1;
  return L;
}

bool fft_supported(int N) { return N >= 3 && fft_len(N) <= FFT_MAX_L; }

void fft_tables(int N, const float* taps, float* G, float2* tw) {
  const int L = fft_len(N);
  std::vector<std::complex<double>> a(L, 0.0);
  for (int m = -(N - 1); m <= N - 1; ++m) a[((m % L) + L) % L] += (double)taps[m + N];
  fft_host(a);
  for (int k = 0; k < L; ++k) G[k] = (float)(a[k].real() / (double)L);
  for (int k = 0; k < L / 2; ++k) {
    const double ang = -2.0 * kPi * (double)k / (double)L;
    tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
}

void launch_ramp_fft(const FFTTask* d_tasks, int ntasks, int base, int blocks, int maxL, hipStream_t s) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(fft_ramp_kernel, dim3(xcd_grid(blocks)), dim3(FFT_THREADS), (size_t)maxL * sizeof(float2), s,
                     d_tasks, ntasks, base, blocks);
}

}  // namespace acu
