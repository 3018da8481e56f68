// acu_tv.hpp -- per-slice state of the batched split-Bregman TV kernels.
#pragma once

#include "acu_internal.hpp"

namespace acu {

// Per-slice operands of the TV kernels (tv.hip).  N*N arrays share the
// elementwise block space [eb_begin, eb_begin+nb), N*T arrays the sinogram
// block space [sb_begin, sb_begin+nsb).
struct TVTask {
  int N, T;
  int eb_begin, nb;
  int sb_begin, nsb;
  float mu, lambda;
  float *uk, *ukp, *dkx, *dky, *bkx, *bky, *ifkt, *iukt, *r, *p, *irpt, *Ap;
  float *f, *fk, *ruk, *rp;
  const float* fbp;      // the FBP warm start, iradon(f)
  // linear schedule: ifkt, iukt (N*N) and ruk (N*T) carried in FP64
  double *ifkt64, *iukt64, *ruk64;
  double* scal;          // 128 doubles: rs[0..ncg], pAp at 64.., Err_diff at 127
  double* partials;      // nb doubles
  float* delta_log;      // delta_cap floats or null
  int delta_cap;
};

constexpr int SCAL_STRIDE = 128;

int ew_blocks(size_t n);
// `base` is the first global block index of the launch: 0 for a launch over
// the whole batch, the slice's own begin index for a single-slice launch.
void launch_tv_finalize(const TVTask* t, int first, int count, int slot, hipStream_t s);
// lin: ifkt and iukt from the linear schedule's FP64 carried terms
void launch_tv_residual(const TVTask* t, int n, int base, int blocks, bool lin, hipStream_t s);
void launch_tv_apdot(const TVTask* t, int n, int base, int blocks, int j, hipStream_t s);
// lin: the linear schedule's carried iukt64 += alpha irpt as well
void launch_tv_axpy(const TVTask* t, int n, int base, int blocks, int j, bool lin, hipStream_t s);
// linear schedule: ruk64 += alpha rp over the sinogram blocks
void launch_tv_sino_axpy(const TVTask* t, int n, int base, int sblocks, int j, hipStream_t s);
void launch_tv_pupdate(const TVTask* t, int n, int base, int blocks, int j, hipStream_t s);
// upd_ifkt: the linear schedule's ifkt64 = (ifkt64 + fbp) - iukt64 as well
void launch_tv_shrink(const TVTask* t, int n, int base, int blocks, bool upd_ifkt, hipStream_t s);
// lin: ruk from the linear schedule's FP64 carried term
void launch_tv_addback(const TVTask* t, int n, int base, int sblocks, bool lin, hipStream_t s);
// linear schedule: load the FP64 carried terms from exact FP32 operator
// results, ifkt64 = ifkt, iukt64 = iukt, ruk64 = ruk
void launch_tv_linload(const TVTask* t, int n, int base, int blocks, int sbase, int sblocks, hipStream_t s);
void launch_tv_init(const TVTask* t, int n, int base, int blocks, int sbase, int sblocks,
                    hipStream_t s);

}  // namespace acu
