/*
 * acumen.h -- C ABI of the MI355X (gfx950) neutron-CT reconstruction library
 * libacumen_hip.so.
 *
 * It replaces the Fortran module `Tomography` of Acumen
 * (tomography/tv_hpc.F90:1-390, with copies in fbp_hpc.F90:1-154 and
 * tv_hpc_slice.F90:1-366).  The reference keeps all state in module globals
 * and OpenACC device arrays; here that state is an opaque handle:
 *
 *   reference                               this library
 *   ------------------------------------    --------------------------------
 *   int_radon()        tv_hpc.F90:11-78     acu_plan_create / acu_batch_create
 *   theta(:) fill      tv_hpc.F90:560-562   acu_batch_set_theta
 *   f(:,:) fill        tv_hpc.F90:614-618   acu_batch_set_sino
 *   radon(sino, img)   tv_hpc.F90:129-159   acu_radon / acu_radon_dev
 *   iradon(img, sino)  tv_hpc.F90:84-126    acu_iradon / acu_iradon_dev
 *   get_slice(img)     fbp_hpc.F90:129-150  acu_get_slice / acu_batch_fbp
 *   tvopt(tv, fbp, j)  tv_hpc.F90:162-363   acu_tvopt / acu_batch_tvopt
 *   fin_radon()        tv_hpc.F90:80-82     acu_batch_destroy
 *
 * Conventions
 *  - Arrays are FP32 in the reference's Fortran layout, first index fastest:
 *    an image img(x, y) is N*N floats with x fastest, a sinogram
 *    sino(detector, angle) is N*T floats with the detector fastest.
 *  - N ("nimage") is the odd internal image side = detector bins, T
 *    ("ntheta") the number of angles, theta in degrees, wpu >= 1 the width of
 *    the mask's smooth edge in pixels ("Radon_wpu").
 *  - Every function returns 0 (ACU_OK) or an ACU_ERR_* code; the message of
 *    the last failure is acu_last_error().  The reference has no error
 *    reporting at all (tomography/tv_hpc.F90: ierr is never checked).
 *  - A context owns one device and one HIP stream; every call on a context
 *    is ordered on that stream.  Functions that take or return host buffers
 *    synchronise before returning; *_dev functions and acu_batch_* compute
 *    calls are asynchronous until acu_ctx_synchronize().
 *  - One context per host thread.
 */
#ifndef ACUMEN_H
#define ACUMEN_H

#ifdef __cplusplus
extern "C" {
#endif

#define ACU_OK 0
#define ACU_ERR_ARG 1     /* invalid argument */
#define ACU_ERR_HIP 2     /* HIP runtime error */
#define ACU_ERR_STATE 3   /* call out of order (e.g. image requested before it exists) */

/* image selectors of acu_batch_get_image / acu_batch_device_ptr */
#define ACU_IMG_TV 0      /* img_rec_tv: result of tvopt (module array uk after tvopt) */
#define ACU_IMG_FBP 1     /* img_rec_fbp: FBP warm start / get_slice result */
#define ACU_SINO_F 2      /* module array f: the windowed sinogram input (N*T) */

/* tvopt execution modes */
#define ACU_TV_PLAIN 0    /* one slice after another, each reduction returned to the host
                             before the next kernel (the reference's structure, tv_hpc.F90:243-302) */
#define ACU_TV_BATCHED 1  /* all slices in every launch, scalars stay on the device */
#define ACU_TV_GRAPH 2    /* ACU_TV_BATCHED with each outer iteration replayed from a HIP graph */

/* Operator schedule of tvopt's outer iteration (acu_batch_set_tv_schedule);
 * with ninner = 1, per outer iteration: */
#define ACU_TV_SCHED_REFERENCE 0 /* 7 radon + 7 iradon, as tv_hpc.F90:205-346 */
#define ACU_TV_SCHED_REUSE 1     /* 6 + 7: the radon of :228 is the add-back's radon(uk) of :339 (same bits) */
#define ACU_TV_SCHED_LINEAR 2    /* 5 + 5 (default): radon(uk), iradon(radon(uk)) and iradon(fk) carried by linearity */

typedef struct acu_ctx acu_ctx;
typedef struct acu_batch acu_batch;

/* Last error message of the calling thread (never NULL). */
const char *acu_last_error(void);

/* Context: device selection and the stream all work is ordered on.  Calls
 * that take pageable host arrays of projections (acu_prep_sinograms,
 * acu_ingest_*) stage them through two pinned host buffers that the context
 * allocates on first use (up to 32 MiB each) and frees in acu_ctx_destroy. */
int acu_ctx_create(int device, acu_ctx **out);
int acu_ctx_destroy(acu_ctx *ctx);
int acu_ctx_synchronize(acu_ctx *ctx);

/* int_radon (tomography/tv_hpc.F90:11-78) for a single slice: filter taps,
 * mask and device arrays.  theta_deg (ntheta floats) may be NULL and be set
 * later with acu_batch_set_theta, as the reference fills theta after
 * int_radon (tv_hpc.F90:559-562). */
int acu_plan_create(acu_ctx *ctx, int nimage, int ntheta, int wpu, const float *theta_deg,
                    acu_batch **out);

/* int_radon for `nslices` independent slices, each with its own N, T, wpu
 * and angles -- the unit the batched kernels launch over.  Replaces the
 * reference's MPI master/worker slice distribution (tv_hpc.F90:583-631). */
int acu_batch_create(acu_ctx *ctx, int nslices, const int *nimage, const int *ntheta,
                     const int *wpu, const float *const *theta_deg, acu_batch **out);

/* fin_radon (tv_hpc.F90:80-82): frees every device array of the batch. */
int acu_batch_destroy(acu_batch *b);

int acu_batch_nslices(const acu_batch *b);

/* theta(1:ntheta) in degrees for one slice (tv_hpc.F90:560-562).  The
 * slice's FBP and TV images of the old angles are invalidated
 * (acu_batch_get_image then returns ACU_ERR_STATE). */
int acu_batch_set_theta(acu_batch *b, int slice, const float *theta_deg);

/* Module array f(N,T) of one slice from host memory (tv_hpc.F90:614-618 +
 * `update device(f)` at :172).  The slice's FBP and TV images of the old
 * sinogram are invalidated. */
int acu_batch_set_sino(acu_batch *b, int slice, const float *f_host);

/* Copy an image (ACU_IMG_TV or ACU_IMG_FBP, N*N floats) of one slice to host
 * memory (`update host(uk)`, tv_hpc.F90:180,357). */
int acu_batch_get_image(acu_batch *b, int slice, int which, float *img_host);

/* Device address of a slice's f, FBP or TV array, for callers that keep
 * their data resident on the GPU.  Returns NULL on error.  The f and FBP
 * addresses hold until acu_batch_destroy.  The TV address holds until a
 * tvopt with more outer iterations than any before on this batch: that
 * call moves the TV arrays, so ask for the address again afterwards. */
float *acu_batch_device_ptr(acu_batch *b, int slice, int which);

/* get_slice (fbp_hpc.F90:129-150) for every slice: img_fbp = iradon(f). */
int acu_batch_fbp(acu_batch *b);

/* tvopt (tv_hpc.F90:162-363) for every slice: FBP warm start into
 * ACU_IMG_FBP, split-Bregman TV into ACU_IMG_TV, with the module parameters
 * mu, lambda, nouter, ninnner (sic) and ncg of the reference.  mode is one of
 * ACU_TV_*.  ncg <= 60. */
int acu_batch_tvopt(acu_batch *b, float mu, float lambda, int nouter, int ninner, int ncg,
                    int mode);

/* Operator schedule of the batch's tvopt (ACU_TV_SCHED_*), any mode.  The
 * reference (tv_hpc.F90:205-346) recomputes, every outer iteration,
 * ifkt = iradon(fk) (:215), ruk = radon(uk) and iukt = iradon(ruk)
 * (:228-229) and radon(uk) again for the add-back (:339).  radon and iradon
 * are linear and the CG already forms rp = radon(p) and irpt = iradon(rp),
 * so ACU_TV_SCHED_LINEAR carries ruk += alpha rp, iukt += alpha irpt and,
 * after the add-back fk = fk + f - ruk, ifkt = ifkt + fbp - iukt (fbp =
 * iradon(f)), from exact values formed in the prologue: 5 radon + 5 iradon
 * per outer iteration instead of 7 + 7, equal in exact arithmetic and within
 * FP32 rounding in practice.  refresh > 0 recomputes the three exactly every
 * `refresh` outer iterations (0: never).  ACU_TV_SCHED_REUSE skips only the
 * radon of :228 (bit-identical to the reference schedule).  Defaults come
 * from ACUMEN_TV_SCHEDULE (linear | reuse | reference) and ACUMEN_TV_REFRESH,
 * else LINEAR with no refresh.  No counterpart in the reference. */
int acu_batch_set_tv_schedule(acu_batch *b, int sched, int refresh);
/* The schedule (ACU_TV_SCHED_*, or -ACU_ERR_ARG); *refresh (may be NULL)
 * receives the refresh period. */
int acu_batch_tv_schedule(const acu_batch *b, int *refresh);

/* Err_diff (the reference's "delta", tv_hpc.F90:327-333,350-352) after each
 * outer iteration of the last tvopt, n <= that tvopt's nouter values.
 * ACU_ERR_STATE when no tvopt has completed on the slice since its sinogram
 * or angles were last set. */
int acu_batch_get_delta(acu_batch *b, int slice, float *delta_host, int n);

/* radon(sinogram, img) (tv_hpc.F90:129-159) on device memory of slice
 * `slice`: img_dev N*N in, sino_dev N*T out.  Asynchronous. */
int acu_radon_dev(acu_batch *b, int slice, float *sino_dev, const float *img_dev);

/* iradon(img_rec, sinogram) (tv_hpc.F90:84-126) on device memory: ramp
 * filter then masked backprojection.  sino_dev N*T in, img_dev N*N out. */
int acu_iradon_dev(acu_batch *b, int slice, float *img_dev, const float *sino_dev);

/* The two halves of iradon: ramp filter (tv_hpc.F90:93-105) and masked
 * backprojection of an already filtered sinogram (:107-123). */
int acu_ramp_dev(acu_batch *b, int slice, float *out_dev, const float *sino_dev);
int acu_backproject_dev(acu_batch *b, int slice, float *img_dev, const float *filtered_dev);

/* Host-memory forms of the reference subroutines on a single-slice plan
 * (slice 0).  Synchronous. */
int acu_radon(acu_batch *b, float *sino_host, const float *img_host);
int acu_iradon(acu_batch *b, float *img_host, const float *sino_host);
int acu_get_slice(acu_batch *b, const float *f_host, float *img_fbp_host);
int acu_tvopt(acu_batch *b, const float *f_host, float mu, float lambda, int nouter, int ninner,
              int ncg, float *img_tv_host, float *img_fbp_host, float *delta_host);

/* ---- Tuning.  No counterpart in the reference (its kernels are generated
 * by the OpenACC compiler). ----
 *
 * Forward-projection variant of a slice: 0 = gathers of bilinear cells
 * from global memory; 1, 2, 4 = LDS-staged pixel boxes with 4, 8 or 16
 * angles per workgroup (1, 2, 4 per lane); 8 = LDS-staged bilinear-cell
 * boxes, 16 angles per 1024-thread workgroup (one per wave).  All compute
 * the same sums; where a chunk's box fits both LDS layouts, 4 and 8 give
 * identical bits.  A batch starts from
 * ACUMEN_FP_PATH / ACUMEN_FPL_AL or a per-shape rule; acu_batch_autotune
 * times every variant (reps launches each, after one warm-up) for every
 * distinct slice shape (N, T, angles) and keeps the fastest for all later
 * calls.  It also times the ramp-filter arithmetics that meet the FBP and
 * TV accuracy bar (ACU_RAMP_FP64, _FP32, _BF16X6, _FFT; not _BF16X3) on the
 * whole batch and keeps the fastest, and the backprojection tile height
 * (below) on the whole batch.  report (may be NULL) receives a JSON summary
 * of the timings and choices, truncated to cap bytes including the
 * terminating NUL.  The choices follow measured times, so two runs may
 * choose differently; projector variants 4 and 8 agree bit for bit where
 * their boxes fit, while the ramp arithmetics and the tile heights differ
 * at FP32 rounding level.  A caller that needs identical bits from run to
 * run stores the choices and reapplies them with the setters.  The
 * getter returns the variant, or -ACU_ERR_ARG for a bad slice; the setter
 * lets a caller reapply a stored choice. */
int acu_batch_autotune(acu_batch *b, int reps, char *report, int cap);

/* Arithmetic of the ramp filter (tv_hpc.F90:93-105) for every slice of the
 * batch: ACU_RAMP_FP64 (default: FP32 inputs, exact products, FP64 sums on
 * the FP64 matrix cores), ACU_RAMP_FP32 (FP32 matrix cores), ACU_RAMP_BF16X3
 * (operands split into bf16 hi + lo, three products, FP32 sums),
 * ACU_RAMP_BF16X6 (hi + mid + lo, six products, FP32 sums), ACU_RAMP_FFT
 * (FFT convolution in FP32 with the kernel's spectrum from FP64; N <= 4097).
 * The initial value comes from ACUMEN_RAMP = fp64 | fp32 | bf16x3 | bf16x6 |
 * fft.  The getter returns the variant or -ACU_ERR_ARG. */
#define ACU_RAMP_FP64 0
#define ACU_RAMP_FP32 1
#define ACU_RAMP_BF16X3 2
#define ACU_RAMP_BF16X6 3
#define ACU_RAMP_FFT 4
int acu_batch_set_ramp_variant(acu_batch *b, int variant);
int acu_batch_ramp_variant(const acu_batch *b);
int acu_batch_fp_variant(const acu_batch *b, int slice);
int acu_batch_set_fp_variant(acu_batch *b, int slice, int variant);

/* Backprojection tile of the batch (tv_hpc.F90:107-123 is its arithmetic):
 * 64-pixel-wide tiles of 16, 32 (default) or 64 rows, one workgroup each.
 * The angle blocks (32 angles, each block summed separately) are the same
 * at every height; with the reference's positions the heights give the
 * same FP32 result, with tile-relative positions they differ by FP32
 * rounding only.  The initial value comes from ACUMEN_BP_TILE.  The getter
 * returns the rows or -ACU_ERR_ARG. */
int acu_batch_set_bp_tile(acu_batch *b, int rows);
int acu_batch_bp_tile(const acu_batch *b);

/* Backprojection positions (tv_hpc.F90:114).  ACU_BP_POS_REFERENCE forms
 * yr = s*(x-m) + c*(y-m) + m as the reference does, each product and sum
 * rounded once in FP32, so every floor, weight and mask decision is the
 * reference's; ACU_BP_POS_TILE forms it relative to the tile origin with
 * one FMA per update (rounded at magnitude < 140 instead of N: closer to
 * FP64, about 25% faster).  fbp: the positions of FBP images, acu_iradon*
 * and acu_backproject_dev (from ACUMEN_BP_POSITIONS = reference | tile);
 * tv: those of the backprojections inside tvopt's iteration, whose warm
 * start is the FBP image (from ACUMEN_TV_BP_POSITIONS).  Both default to
 * REFERENCE: past ~25 outer iterations tvopt amplifies rounding
 * differences, and the reference's positions keep its result on the
 * reference's own FP32 trajectory (3e-5 from it after 50 outer at
 * 513^2 x 360, against 2e-4 with tile positions inside tvopt, both ~1e-4
 * from FP64); TILE inside tvopt trades that for ~6% of its throughput.
 * Setting them invalidates the batch's FBP and TV images.  No counterpart
 * in the reference. */
#define ACU_BP_POS_REFERENCE 0
#define ACU_BP_POS_TILE 1
int acu_batch_set_bp_positions(acu_batch *b, int fbp, int tv);
/* *fbp and *tv (either may be NULL) receive the positions; ACU_OK. */
int acu_batch_bp_positions(const acu_batch *b, int *fbp, int *tv);

/* ---- Differentiation.  No counterpart in the reference: these build on
 * its projector (tv_hpc.F90:129-159) and the periodic differences of its
 * tvopt (tv_hpc.F90:219-221) to make reconstruction and centre-of-rotation
 * registration gradient-based. ---- */

/* Exact transpose of radon (the same FP32 sample positions, bounds test and
 * bilinear weights), so <radon(x), y> = <x, radon_adjoint(y)> up to FP32
 * rounding: sino N*T in, img N*N out.  The _dev form is asynchronous on
 * device memory of slice `slice`; the host form works on slice 0. */
int acu_radon_adjoint_dev(acu_batch *b, int slice, float *img_dev, const float *sino_dev);
int acu_radon_adjoint(acu_batch *b, float *img_host, const float *sino_host);

/* For every slice i, with b_i its sinogram (acu_batch_set_sino) and s_i its
 * per-angle detector shifts (T values in bins; the projection row is read
 * at detector position d - s by linear interpolation, zero off the
 * detector):
 *   L_i = 1/2 |shift_{s_i}(radon(x_i)) - b_i|^2 + lambda TV_eps(x_i),
 *   TV_eps(x) = sum_p sqrt((Dx x)_p^2 + (Dy x)_p^2 + eps^2),
 *   (Dx u)(x,y) = u(x-1,y) - u(x,y), periodic.
 * loss receives 3 doubles per slice: L_i, the data term and TV_eps(x_i);
 * grad_x[i] (N*N) = dL_i/dx_i, grad_s[i] (T) = dL_i/ds_i, curv_s[i] (T) =
 * sum_d (y(d+o) - y(d+o+1))^2 with y = radon(x_i) and o = floor(-s): the
 * Gauss-Newton curvature of the data term in each shift (its exact second
 * derivative between the kinks of the interpolation).  shift (or any
 * entry) may be NULL for zero shifts; grad_x NULL skips the image gradient;
 * grad_s / curv_s (or any entry) NULL skips that output.  eps > 0.
 * Data terms and sums are accumulated in FP64.  Synchronous.  The _dev form
 * takes device pointers. */
int acu_batch_loss_grad(acu_batch *b, const float *const *x_host, const float *const *shift_host,
                        float lambda, float eps, float *const *grad_x_host,
                        float *const *grad_s_host, float *const *curv_s_host, double *loss);
int acu_batch_loss_grad_dev(acu_batch *b, const float *const *x_dev, const float *const *shift_dev,
                            float lambda, float eps, float *const *grad_x_dev,
                            float *const *grad_s_dev, float *const *curv_s_dev, double *loss);

/* ---- Volume streaming.  Replaces the reference drivers' slice loop over
 * MPI ranks (tv_hpc.F90:583-631, fbp_hpc.F90:340-403) for a volume whose
 * slices share one shape. ----
 *
 * nslices sinograms (sino_host: nslices * T * N floats, slice after slice,
 * each (T, N) detector fastest as in sinograms{js}.dat) go through the GPU
 * in batches of `batch` slices on two lanes, each with its own HIP stream
 * and pinned staging buffers, so that one lane's host staging and PCIe
 * transfers overlap the other lane's kernels.  fbp_host (nslices * N * N,
 * may be NULL) receives get_slice (fbp_hpc.F90:129-150); tv_host (may be
 * NULL) receives tvopt's TV image (tv_hpc.F90:162-363) with mu, lambda,
 * nouter >= 1, ninner, ncg.  tune = 1 runs acu_batch_autotune on the first
 * batch and uses its choices throughout.  stats (may be NULL, else 4
 * doubles) receives [0] wall milliseconds of the call, [1] slices per
 * second over the call, [2] milliseconds of set-up (lanes, pinned buffers,
 * tuning), [3] slices per second of the streaming after set-up; PCIe and
 * host staging are included in all.  Synchronous; uses device memory for 2
 * batches.  ACUMEN_COPY_THREADS (default: the host's hardware threads, at
 * most 16) sets the host threads of the staging copies. */
int acu_volume_reconstruct(acu_ctx *ctx, int nslices, int nimage, int ntheta, int wpu,
                           const float *theta_deg, const float *sino_host, float *fbp_host,
                           float *tv_host, float mu, float lambda, int nouter, int ninner, int ncg,
                           int batch, int tune, double *stats);

/* ---- Projection -> sinogram preprocessing, the GPU restatement of
 * tomography/projections_to_sinogram_hpc.F90:231-487.  A projection is
 * img(x, y), n x n floats (x = detector column fastest, y = detector row =
 * slice); Nw = n + add_nx with add_nx = 1 for even n (the reference's
 * add_Nx).  Sinogram volumes are n slices x nfiles angles x Nw, slice-major,
 * the layout of sinograms{js}.dat. ----
 *
 * 3x3x3 median of projection c with its neighbours l and r, neighbours
 * clamped into the image (:420-447).  Device pointers, asynchronous. */
int acu_prep_median27_dev(acu_ctx *ctx, int n, const float *l, const float *c, const float *r,
                          float *out);
/* Rotation / centre-of-rotation correction by clamped bilinear resampling
 * (:449-464) of `count` projections (device, n*n each) into rows job0 ..
 * job0+count-1 of every slice's sinogram in the device volume `sino`
 * (n x nfiles x Nw).  Asynchronous. */
int acu_prep_correct_dev(acu_ctx *ctx, int n, int add_nx, float theta, float shift,
                         const float *imgs, int count, int job0, int nfiles, float *sino);
/* Mirror error sum |V0(x,y) - V180(Nw-x+1,y)| of the resampled 0 and 180
 * degree projections at (theta, shift) (:260-283), summed in FP64.
 * Synchronous. */
int acu_prep_mirror_error_dev(acu_ctx *ctx, int n, int add_nx, float theta, float shift,
                              const float *p0_dev, const float *p180_dev, double *err);
/* No counterpart in the reference: the library's one entry point that
 * allocates nothing, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * mirror error of acu_prep_mirror_error_dev with its FP64 block partials in
 * a static device buffer of ACU_NOALLOC_SCRATCH_DOUBLES doubles (512 KiB per
 * GPU, held until the process ends) instead of memory allocated per call,
 * and the result written to device memory err_dev (one double), both on the
 * context's stream with no wait: read err_dev after acu_ctx_synchronize or
 * a later stream-ordered copy.  Same bits as acu_prep_mirror_error_dev.
 * ACU_ERR_ARG when ceil((n + add_nx) * n / 256) > ACU_NOALLOC_SCRATCH_DOUBLES
 * (n at most 4095).  Every call on one GPU shares that GPU's buffer: calls
 * that overlap on one GPU (other streams or threads) race. */
#define ACU_NOALLOC_SCRATCH_DOUBLES 65536
int acu_prep_mirror_error_noalloc(acu_ctx *ctx, int n, int add_nx, float theta, float shift,
                                  const float *p0_dev, const float *p180_dev, double *err_dev);
/* Centre-of-rotation and tilt search (:231-364): centroid estimate, then the
 * reference's pattern search on (theta, shift) with every trial's mirror
 * error on the GPU.  Host projections in; theta, shift and (may be NULL)
 * the number of error evaluations out.  Synchronous. */
int acu_prep_find_center(acu_ctx *ctx, int n, const float *p0_host, const float *p180_host,
                         float *theta, float *shift, int *evaluations);
/* Median filter and correction of all nfiles projections (host, nfiles*n*n)
 * into the sinogram volume (host, n*nfiles*Nw) (:384-487).  Synchronous. */
int acu_prep_sinograms(acu_ctx *ctx, int n, int nfiles, const float *projections, float theta,
                       float shift, float *sinograms);

/* ---- FITS ingest normalisation, the GPU part of tomography/fits_to_hpc.F90
 * (the FITS decoding itself is host I/O). ----
 *
 * Mean of `count` images of npix pixels (device, image after image),
 * accumulated in file order in FP32 and divided by count, as the open-beam
 * average of :203-213.  Asynchronous. */
int acu_ingest_mean_dev(acu_ctx *ctx, int count, long long npix, const float *imgs, float *out);
/* out = -log(image / max(open_beam, 1)) for `count` images (:243).
 * Asynchronous. */
int acu_ingest_normalize_dev(acu_ctx *ctx, int count, long long npix, const float *imgs,
                             const float *open_beam, float *out);

/* Host-buffer forms of the two above (staged through device memory;
 * synchronous).  These are what the ingest driver calls per batch of
 * decoded FITS images. */
int acu_ingest_mean(acu_ctx *ctx, int count, long long npix, const float *imgs, float *out);
int acu_ingest_normalize(acu_ctx *ctx, int count, long long npix, const float *imgs,
                         const float *open_beam, float *out);

/* Kernel timing: when enabled, every ramp-filter, backprojection, cells,
 * forward-projection, preprocessing and ingest launch is bracketed by HIP events on the context's
 * stream (graph mode is then run as ACU_TV_BATCHED).  acu_ctx_kernel_time
 * reports the accumulated milliseconds and launch count of one kind:
 * 0 = ramp filter, 1 = backprojection, 2 = cells pass, 3 = forward projection,
 * 4 = adjoint projection, 5 = preprocessing (median filter + correction of
 * one batch of projections), 6 = ingest (open-beam mean, normalisation). */
int acu_ctx_set_timing(acu_ctx *ctx, int enable);
int acu_ctx_kernel_time(acu_ctx *ctx, int kind, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif

#endif /* ACUMEN_H */
