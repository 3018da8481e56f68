// acu_grad.hpp -- differentiation of the projector (grad.hip): the exact
// adjoint of radon and the objective
//   L(x, s) = 1/2 |shift_s(radon(x)) - b|^2 + lambda TV_eps(x)
// with its gradients with respect to the image x and the per-angle
// centre-of-rotation shifts s.  There is no counterpart in the reference;
// the operators build on its projector (tomography/tv_hpc.F90:129-159) and
// the periodic differences of its tvopt (:219-221).
#pragma once

#include "acu_internal.hpp"

namespace acu {

constexpr int ADJ_T = 16;          // pixel tile side of the adjoint (16 x 16 threads)
constexpr int ADJ_AB = 32;         // angles per FP32 block sum
constexpr int GR_THREADS = 256;

// Exact transpose of radon: img(px,py) = (1/N) sum over angles and samples
// whose bilinear stencil covers (px,py) of sino * weight, with the
// projector's own FP32 positions and bounds test.
struct AdjTask {
  int N, T;
  float mid;
  int block_begin;   // first global block of this slice
  int ntx;           // tiles per side: ceil(N / ADJ_T)
  const float* sino; // (T, N)
  const float2* sc;  // (sin, cos) of the projector angles
  float* out;        // (N, N)
};

// One angle row per workgroup: r = shift_s(y) - b, rs = shift_s^T(r),
// gs[t] = sum_d r(d) (y(d+o) - y(d+o+1)), hs[t] = sum_d (y(d+o) - y(d+o+1))^2
// (Gauss-Newton curvature in s_t), part[t] = 1/2 sum_d r(d)^2.
struct ResTask {
  int N, T;
  int row_begin;     // first global row (block) of this slice
  const float* y;    // radon(x), (T, N)
  const float* b;    // measured sinogram, (T, N)
  const float* shift;  // T shifts in detector bins, or null (zero)
  float* rs;         // (T, N) out
  float* gs;         // T out, or null
  float* hs;         // T out, or null
  double* part;      // T data-term partials
};

// gx += lambda (Dx^T qx + Dy^T qy) (when gx is set), part[block] = partial of
// TV_eps = sum sqrt(Dx^2 + Dy^2 + eps^2); Dx u(x,y) = u(x-1,y) - u(x,y), periodic.
struct TVGTask {
  int N;
  int block_begin, nb;
  float lambda, eps;
  const float* x;
  float* gx;
  double* part;
};

// loss[0..2] = (data + lambda tv, data, tv) from the partials of a slice.
struct LossTask {
  int nd, nt;
  double lambda;
  const double* pd;
  const double* pt;
  double* loss;
};

inline int adj_tiles(int N) { return (N + ADJ_T - 1) / ADJ_T; }
inline int tvg_blocks(int N) { return (int)(((size_t)N * N + GR_THREADS - 1) / GR_THREADS); }

void launch_adjoint(const AdjTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s);
void launch_shift_resid(const ResTask* d_tasks, int ntasks, int base, int rows, hipStream_t s);
void launch_tvgrad(const TVGTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s);
void launch_loss_finalize(const LossTask* d_tasks, int ntasks, hipStream_t s);

}  // namespace acu
