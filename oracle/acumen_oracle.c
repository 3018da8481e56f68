/*
 * acumen_oracle.c -- CPU oracle for the neutron-CT reconstruction path
 * (radon, iradon = ramp filter + backprojection, FBP, split-Bregman TV with
 * inner CG) of Acumen's Fortran module `Tomography`
 * (tomography/tv_hpc.F90:1-390, duplicated in fbp_hpc.F90:1-154 and
 * tv_hpc_slice.F90:1-366).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * test-suite, by the smoke check and by the benchmark's CPU-baseline leg.
 * The product library never links, loads or calls it.
 *
 * Two builds of the same body (oracle_body.h):
 *   orc_*_f32  the reference's FP32 arithmetic, one rounding per Fortran
 *              operation, no FMA contraction (compile with
 *              -ffp-contract=off), sequential FP32 reductions;
 *   orc_*_f64  the same algorithm in FP64: the accuracy yardstick.
 *
 * Differentiation (no counterpart in the reference, built on its projector):
 * orc_radon_adjoint_* is the exact transpose of orc_radon_*; orc_shift_* and
 * orc_shift_adjoint_* apply a per-angle detector shift by linear
 * interpolation and its transpose; orc_loss_grad_* evaluates
 * 1/2 |shift(radon(x)) - b|^2 + lambda TV_eps(x) and its gradients with
 * respect to the image and to the per-angle shifts.
 *
 * Threading: orc_set_threads(n) parallelises radon / ramp / backprojection
 * over output elements.  Every output element is still one sequential sum,
 * so results do not depend on the thread count.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_threads = 1;
/* 1: every precision uses the reference's FP32 per-angle sin/cos tables, so
 * the FP64 build differs from the FP32 one in arithmetic only, not in the
 * (FP32-rounded) projection angles. */
static int g_fp32_angles = 0;

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
void orc_set_fp32_angles(int on) { g_fp32_angles = on != 0; }

/* ---- FP32 build: the reference's own arithmetic ---- */
#define REAL float
#define DREAL float
#define FN(name) orc_##name##_f32
#define TVFN(name) orc_##name##_f32
#define SIN sinf
#define COS cosf
#define EXP expf
#define SQRT sqrtf
#define ATAN atanf
#define FLOOR floorf
#define FABS fabsf
#include "oracle_body.h"
#undef REAL
#undef DREAL
#undef FN
#undef TVFN
#undef SIN
#undef COS
#undef EXP
#undef SQRT
#undef ATAN
#undef FLOOR
#undef FABS

/* ---- FP64 build: the yardstick ---- */
#define REAL double
#define DREAL double
#define FN(name) orc_##name##_f64
#define TVFN(name) orc_##name##_f64
#define SIN sin
#define COS cos
#define EXP exp
#define SQRT sqrt
#define ATAN atan
#define FLOOR floor
#define FABS fabs
#include "oracle_body.h"
#undef REAL
#undef DREAL
#undef FN
#undef TVFN
#undef SIN
#undef COS
#undef EXP
#undef SQRT
#undef ATAN
#undef FLOOR
#undef FABS

/* ---- Mixed build: FP32 operators and arrays, FP64 scalar reductions and
 * step sizes (rsold, p.Ap, rsnew, alpha, beta, Err_diff).  This is the
 * precision policy of the GPU path; only tvopt differs from the FP32 build. */
#define ORC_TV_ONLY 1
#define REAL float
#define DREAL double
#define FN(name) orc_##name##_f32
#define TVFN(name) orc_##name##_f32d
#define SIN sinf
#define COS cosf
#define EXP expf
#define SQRT sqrtf
#define ATAN atanf
#define FLOOR floorf
#define FABS fabsf
#include "oracle_body.h"
#undef ORC_TV_ONLY
#undef REAL
#undef DREAL
#undef FN
#undef TVFN
#undef SIN
#undef COS
#undef EXP
#undef SQRT
#undef ATAN
#undef FLOOR
#undef FABS

/* Per-angle sin/cos tables exactly as the FP32 reference evaluates them
 * (exported so tests can check the product's tables against them). */
void orc_angle_tables_f32(int T, const float *theta, int bp, float *sn, float *cs)
{
    orc_angles_f32(T, theta, bp, sn, cs);
}

/* CPU baseline workload: `nslices` independent slices, one per thread, each
 * running one radon (tv_hpc.F90:129-159) and one iradon (:84-126) in the
 * reference's FP32 arithmetic -- the same per-slice decomposition as the
 * reference's MPI master/worker drivers (tv_hpc.F90:583-631). */
void orc_pass_batch_f32(int nslices, int N, int T, const float *theta, int wpu,
                        const float *imgs, float *sinos, float *imgs_out)
{
    const size_t NN = (size_t)N * N, NT = (size_t)N * T;
    float *h = (float *)malloc(sizeof(float) * (size_t)(2 * N + 1));
    float *cm = (float *)malloc(sizeof(float) * NN);
    orc_filter_f32(N, h);
    orc_cmask_f32(N, T, wpu, cm);
    const int saved = g_threads;
    const int nth = saved;
    g_threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth) if (nth > 1)
    for (int s = 0; s < nslices; ++s) {
        orc_radon_f32(N, T, theta, imgs + (size_t)s * NN, sinos + (size_t)s * NT);
        orc_iradon_f32(N, T, theta, h, cm, sinos + (size_t)s * NT, imgs_out + (size_t)s * NN);
    }
    g_threads = saved;
    free(h);
    free(cm);
}
