/*
 * oracle_body.h -- body of the CPU restatement of Acumen's `Tomography`
 * module, included twice by acumen_oracle.c: once with REAL=float (the
 * reference's own FP32 arithmetic, operation by operation, no FMA
 * contraction) and once with REAL=double (the FP64 truth).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this.
 *
 * Array conventions follow the reference: Fortran column-major, first index
 * fastest, 1-based indices in the formulas below and 0-based in memory:
 *   image    (x, y)      -> img [(x-1) + (y-1)*N]
 *   sinogram (det, angle)-> sino[(d-1) + (a-1)*N]
 *
 * Every function cites the reference lines it restates
 * (tomography/tv_hpc.F90 unless stated otherwise).
 */

#ifndef REAL
#error "include through acumen_oracle.c"
#endif

#define RC(x) ((REAL)(x))

#ifndef ORC_TV_ONLY

/* pi = 4.0*atan(1.0)  (tv_hpc.F90:16) */
static REAL FN(pi)(void) { return RC(4.0) * ATAN(RC(1.0)); }

/* midpoint = (nimage+1)/2 with integer division (tv_hpc.F90:15) */
static REAL FN(midpoint)(int N) { return (REAL)((N + 1) / 2); }

/* Ramp filter taps, tv_hpc.F90:53-56:
 *   filter(jx) = -2.0*N / (pi**2 * (4*(jx-1-N)**2 - 1)),  jx = 1..2N+1
 * The integer expression 4k^2-1 is evaluated in integer arithmetic. */
void FN(filter)(int N, REAL *h)
{
    const REAL pi = FN(pi)();
    const int nfilt = 2 * N + 1;
    for (int jx = 1; jx <= nfilt; ++jx) {
        const int k = jx - 1 - N;
        const int q = 4 * k * k - 1;
        h[jx - 1] = (-(RC(2.0) * (REAL)N)) / ((pi * pi) * (REAL)q);
    }
}

/* Circular smooth mask, tv_hpc.F90:58-74.  wpu must be >= 1. */
void FN(cmask)(int N, int T, int wpu, REAL *cmask)
{
    const REAL pi = FN(pi)();
    const REAL rad_n = (REAL)(N - 1) / RC(2.0) - (REAL)wpu;
    const REAL half = ((REAL)N - RC(1.0)) / RC(2.0);
    for (int jx = 1; jx <= N; ++jx) {
        for (int jy = 1; jy <= N; ++jy) {
            const REAL dx = ((REAL)jx - half) - RC(1.0);
            const REAL dy = ((REAL)jy - half) - RC(1.0);
            const REAL rr = SQRT(dx * dx + dy * dy);
            const REAL rtmp = RC(1.0) - (RC(2.0) * (rr - rad_n)) / (REAL)wpu;
            REAL v;
            if (rtmp <= RC(0.0)) {
                v = RC(0.0);
            } else if (rtmp >= RC(1.0)) {
                v = pi / (REAL)T;
            } else {
                const REAL e1 = EXP(RC(1.0) - RC(1.0) / (RC(1.0) - rtmp));
                v = (pi * EXP(RC(1.0) - RC(1.0) / (RC(1.0) - e1))) / (REAL)T;
            }
            cmask[(jx - 1) + (size_t)(jy - 1) * N] = v;
        }
    }
}

/* Ramp filtering (first half of iradon), tv_hpc.F90:93-105:
 *   sinoCS(jy,t) = sum_{jx=1..N} sino(jx,t) * filter(jy-jx+1+N)
 * summed sequentially in jx.  The reference's guard 1 < jk < 2N+2 on
 * jk = jy-jx+1+N always holds for jx, jy in 1..N (jk runs over 2..2N), so it
 * is not tested.  Loop order: jx outer, jy inner, with one running sum per
 * output jy -- every output is still the sum over jx = 1..N in that order,
 * one rounding per operation, and the inner loop runs over contiguous taps. */
void FN(ramp)(int N, int T, const REAL *h, const REAL *sino, REAL *sinoCS)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int jt = 1; jt <= T; ++jt) {
        const REAL *s = sino + (size_t)(jt - 1) * N;
        REAL *out = sinoCS + (size_t)(jt - 1) * N;
        for (int jy = 0; jy < N; ++jy) out[jy] = RC(0.0);
        for (int jx = 1; jx <= N; ++jx) {
            const REAL sv = s[jx - 1];
            const REAL *hk = h + (N - jx);  /* hk[jy-1] = filter(jy-jx+1+N) */
            for (int jy = 1; jy <= N; ++jy)
                out[jy - 1] = out[jy - 1] + sv * hk[jy];
        }
    }
}

/* Per-angle sin/cos of ctheta = pi*(90-theta)/180 (backprojection,
 * tv_hpc.F90:113) or pi*(theta-90)/180 (projection, tv_hpc.F90:142).
 * sin and cos are pure functions of ctheta, so hoisting them out of the
 * pixel loops leaves every value unchanged. */
static void FN(angles)(int T, const float *theta, int bp, REAL *sn, REAL *cs)
{
    const REAL pi = FN(pi)();
    for (int j = 0; j < T; ++j) {
        const REAL th = (REAL)theta[j];
        const REAL ct = bp ? (pi * (RC(90.0) - th)) / RC(180.0)
                           : (pi * (th - RC(90.0))) / RC(180.0);
        sn[j] = SIN(ct);
        cs[j] = COS(ct);
        if (g_fp32_angles) {
            /* the reference's FP32 sin/cos of its FP32 ctheta, promoted */
            const float pf = 4.0f * atanf(1.0f), tf = theta[j];
            const float cf = bp ? (pf * (90.0f - tf)) / 180.0f : (pf * (tf - 90.0f)) / 180.0f;
            sn[j] = (REAL)sinf(cf);
            cs[j] = (REAL)cosf(cf);
        }
    }
}

/* Pixel-driven linear-interpolation backprojection (second half of
 * iradon), tv_hpc.F90:107-123:
 *   yr = sin(c)*(x-m) + cos(c)*(y-m) + m,  y1 = floor(yr), y2 = y1+1,
 *   taps used iff y1 > 1 and y2 < N+1,
 *   img(x,y) = cmask(x,y) * sum_t [S(y1)*(y2-yr) + S(y2)*(yr-y1)].
 * Loop order: blocks of BP_RB image rows x blocks of BP_AB angles, with one
 * running sum per pixel; angle blocks go in order, so every pixel is still
 * the sum over t = 1..T in that order (two additions per angle, as in the
 * reference), while a block of sinogram rows stays in cache. */
#ifndef ORC_BP_BLOCKS
#define ORC_BP_BLOCKS
enum { BP_RB = 8, BP_AB = 32 };
#endif
void FN(backproject)(int N, int T, const float *theta, const REAL *cmask,
                     const REAL *sinoCS, REAL *img)
{
    REAL *sn = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    REAL *cs = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    FN(angles)(T, theta, 1, sn, cs);
    const REAL m = FN(midpoint)(N);
    const int nrb = (N + BP_RB - 1) / BP_RB;
#pragma omp parallel num_threads(g_threads) if (g_threads > 1)
    {
        REAL *sum = (REAL *)malloc(sizeof(REAL) * (size_t)BP_RB * N);
        REAL *um = (REAL *)malloc(sizeof(REAL) * (size_t)N);
        for (int jx = 1; jx <= N; ++jx) um[jx - 1] = (REAL)jx - m;
#pragma omp for schedule(dynamic, 1)
        for (int rb = 0; rb < nrb; ++rb) {
            const int y0 = 1 + rb * BP_RB, y1e = y0 + BP_RB - 1 < N ? y0 + BP_RB - 1 : N;
            for (size_t i = 0; i < (size_t)BP_RB * N; ++i) sum[i] = RC(0.0);
            for (int t0 = 1; t0 <= T; t0 += BP_AB) {
                const int t1 = t0 + BP_AB - 1 < T ? t0 + BP_AB - 1 : T;
                for (int jy = y0; jy <= y1e; ++jy) {
                    REAL *row = sum + (size_t)(jy - y0) * N;
                    const REAL vy = (REAL)jy - m;
                    for (int jt = t0; jt <= t1; ++jt) {
                        const REAL *S = sinoCS + (size_t)(jt - 1) * N;
                        const REAL st = sn[jt - 1], cy = cs[jt - 1] * vy;
                        for (int jx = 1; jx <= N; ++jx) {
                            const REAL yr = (st * um[jx - 1] + cy) + m;
                            const int ya = (int)FLOOR(yr);
                            const int yb = ya + 1;
                            if (ya > 1 && yb < N + 1) {
                                REAL v = row[jx - 1] + S[ya - 1] * ((REAL)yb - yr);
                                row[jx - 1] = v + S[yb - 1] * (yr - (REAL)ya);
                            }
                        }
                    }
                }
            }
            for (int jy = y0; jy <= y1e; ++jy)
                for (int jx = 1; jx <= N; ++jx) {
                    const size_t o = (size_t)(jx - 1) + (size_t)(jy - 1) * N;
                    img[o] = cmask[o] * sum[(size_t)(jy - y0) * N + (jx - 1)];
                }
        }
        free(um);
        free(sum);
    }
    free(sn);
    free(cs);
}

/* iradon = ramp filter + backprojection (tv_hpc.F90:84-126). */
void FN(iradon)(int N, int T, const float *theta, const REAL *h, const REAL *cmask,
                const REAL *sino, REAL *img)
{
    REAL *cs = (REAL *)malloc(sizeof(REAL) * (size_t)N * T);
    FN(ramp)(N, T, h, sino, cs);
    FN(backproject)(N, T, theta, cmask, cs, img);
    free(cs);
}

/* Ray-driven bilinear forward projection, tv_hpc.F90:129-159:
 *   xr = cos(c)*(x-m) - sin(c)*(y-m) + m,  yr = sin(c)*(x-m) + cos(c)*(y-m) + m
 *   taps used iff min(x1,y1) > 1 and max(x2,y2) < N+1,
 *   sino(y,t) = (1/N) * sum_{x=1..N} bilinear(img; xr, yr), x in order.
 * Loop order: per angle, blocks of FP_RB neighbouring rays step together
 * (x outer, rays inner), one running sum per ray: each ray is still summed
 * over x = 1..N in that order, while the block's samples at one step are
 * neighbouring pixels that stay in cache from one step to the next.  The
 * products c*(x-m), s*(x-m), s*(y-m) and c*(y-m) are the reference's
 * (each rounded once), computed once and reused. */
#ifndef ORC_FP_BLOCKS
#define ORC_FP_BLOCKS
enum { FP_RB = 16 };
#endif
void FN(radon)(int N, int T, const float *theta, const REAL *img, REAL *sino)
{
    REAL *sn = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    REAL *cs = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    FN(angles)(T, theta, 0, sn, cs);
    const REAL m = FN(midpoint)(N);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) if (g_threads > 1)
    for (int jt = 1; jt <= T; ++jt) {
        const REAL s = sn[jt - 1], c = cs[jt - 1];
        for (int y0 = 1; y0 <= N; y0 += FP_RB) {
            const int nr = y0 + FP_RB - 1 <= N ? FP_RB : N - y0 + 1;
            REAL sum[FP_RB], sy[FP_RB], cy[FP_RB];
            for (int k = 0; k < nr; ++k) {
                const REAL v = (REAL)(y0 + k) - m;
                sum[k] = RC(0.0);
                sy[k] = s * v;
                cy[k] = c * v;
            }
            for (int jx = 1; jx <= N; ++jx) {
                const REAL u = (REAL)jx - m;
                const REAL cu = c * u, su = s * u;
                for (int k = 0; k < nr; ++k) {
                    const REAL xr = (cu - sy[k]) + m;
                    const REAL yr = (su + cy[k]) + m;
                    const int y1 = (int)FLOOR(yr), y2 = y1 + 1;
                    const int x1 = (int)FLOOR(xr), x2 = x1 + 1;
                    const int mn = x1 < y1 ? x1 : y1;
                    const int mx = x2 > y2 ? x2 : y2;
                    if (mn > 1 && mx < N + 1) {
                        const REAL *r1 = img + (size_t)(y1 - 1) * N;
                        const REAL *r2 = r1 + N;
                        const REAL ax = (REAL)x2 - xr, bx = xr - (REAL)x1;
                        const REAL ay = (REAL)y2 - yr, by = yr - (REAL)y1;
                        REAL v = sum[k] + (r1[x1 - 1] * ax) * ay;
                        v = v + (r1[x2 - 1] * bx) * ay;
                        v = v + (r2[x1 - 1] * ax) * by;
                        sum[k] = v + (r2[x2 - 1] * bx) * by;
                    }
                }
            }
            for (int k = 0; k < nr; ++k)
                sino[(size_t)(y0 + k - 1) + (size_t)(jt - 1) * N] = sum[k] / (REAL)N;
        }
    }
    free(sn);
    free(cs);
}

/* ---- Differentiation of the projector (no counterpart in the reference;
 * built on its radon, tv_hpc.F90:129-159) ---- */

/* Exact transpose of FN(radon): the same sample positions, bounds test and
 * bilinear weights (tv_hpc.F90:143-153), scattered into the image instead of
 * gathered, so that <radon(x), y> = <x, radon_adjoint(y)>.  Angles are split
 * over threads, each with a private image, summed at the end. */
void FN(radon_adjoint)(int N, int T, const float *theta, const REAL *sino, REAL *img)
{
    REAL *sn = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    REAL *cs = (REAL *)malloc(sizeof(REAL) * (size_t)T);
    FN(angles)(T, theta, 0, sn, cs);
    const REAL m = FN(midpoint)(N);
    const size_t NN = (size_t)N * N;
    const int nt = g_threads > T ? (T > 0 ? T : 1) : g_threads;
    REAL *part = (REAL *)calloc((size_t)nt * NN, sizeof(REAL));
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int k = 0; k < nt; ++k) {
        REAL *acc = part + (size_t)k * NN;
        for (int jt = 1 + k; jt <= T; jt += nt) {
            const REAL s = sn[jt - 1], c = cs[jt - 1];
            for (int jy = 1; jy <= N; ++jy) {
                const REAL sv = sino[(jy - 1) + (size_t)(jt - 1) * N] / (REAL)N;
                for (int jx = 1; jx <= N; ++jx) {
                    const REAL xr = (c * ((REAL)jx - m) - s * ((REAL)jy - m)) + m;
                    const REAL yr = (s * ((REAL)jx - m) + c * ((REAL)jy - m)) + m;
                    const int y1 = (int)FLOOR(yr), y2 = y1 + 1;
                    const int x1 = (int)FLOOR(xr), x2 = x1 + 1;
                    const int mn = x1 < y1 ? x1 : y1;
                    const int mx = x2 > y2 ? x2 : y2;
                    if (mn > 1 && mx < N + 1) {
                        const REAL ax = (REAL)x2 - xr, bx = xr - (REAL)x1;
                        const REAL ay = (REAL)y2 - yr, by = yr - (REAL)y1;
                        acc[(x1 - 1) + (size_t)(y1 - 1) * N] += (sv * ax) * ay;
                        acc[(x2 - 1) + (size_t)(y1 - 1) * N] += (sv * bx) * ay;
                        acc[(x1 - 1) + (size_t)(y2 - 1) * N] += (sv * ax) * by;
                        acc[(x2 - 1) + (size_t)(y2 - 1) * N] += (sv * bx) * by;
                    }
                }
            }
        }
    }
    for (size_t i = 0; i < NN; ++i) {
        REAL v = RC(0.0);
        for (int k = 0; k < nt; ++k) v += part[(size_t)k * NN + i];
        img[i] = v;
    }
    free(part);
    free(sn);
    free(cs);
}

/* Per-angle centre-of-rotation shift of one projection row: out(d) is the
 * row sampled at detector position d - s by linear interpolation, zero
 * outside 1..N.  With o = floor(-s) and w = -s - o (exact in FP32):
 *   out(d) = (1 - w) y(d + o) + w y(d + o + 1). */
static void FN(shift_weights)(float s, int *o, REAL *w)
{
    const REAL sg = -(REAL)s;
    const REAL fl = FLOOR(sg);
    *o = (int)fl;
    *w = sg - fl;
}

static inline REAL FN(at)(const REAL *row, int N, int i) { return (i >= 1 && i <= N) ? row[i - 1] : RC(0.0); }

void FN(shift)(int N, int T, const float *shift, const REAL *sino, REAL *out)
{
    for (int jt = 0; jt < T; ++jt) {
        int o; REAL w;
        FN(shift_weights)(shift[jt], &o, &w);
        const REAL *row = sino + (size_t)jt * N;
        for (int d = 1; d <= N; ++d)
            out[(size_t)jt * N + d - 1] = (RC(1.0) - w) * FN(at)(row, N, d + o) + w * FN(at)(row, N, d + o + 1);
    }
}

/* Transpose of FN(shift): out(k) = (1 - w) r(k - o) + w r(k - o - 1). */
void FN(shift_adjoint)(int N, int T, const float *shift, const REAL *sino, REAL *out)
{
    for (int jt = 0; jt < T; ++jt) {
        int o; REAL w;
        FN(shift_weights)(shift[jt], &o, &w);
        const REAL *row = sino + (size_t)jt * N;
        for (int k = 1; k <= N; ++k)
            out[(size_t)jt * N + k - 1] = (RC(1.0) - w) * FN(at)(row, N, k - o) + w * FN(at)(row, N, k - o - 1);
    }
}

#endif /* ORC_TV_ONLY */

/* Periodic neighbour indices of tvopt (tv_hpc.F90:220,234-237), derived from
 * FP32 floor formulas that reduce exactly to wrap-around for N < 2^24. */
static inline int TVFN(nxt)(int j, int N) { return j == N ? 1 : j + 1; }
static inline int TVFN(prv)(int j, int N) { return j == 1 ? N : j - 1; }

#ifndef ORC_TV_ONLY
/* Objective of gradient-based reconstruction and registration:
 *   L(x, s) = 1/2 |shift_s(radon(x)) - b|^2 + lambda TV_eps(x),
 *   TV_eps(x) = sum_p sqrt((Dx x)_p^2 + (Dy x)_p^2 + eps^2),
 * with the module's periodic differences (Dx u)(x,y) = u(x-1,y) - u(x,y)
 * (tv_hpc.F90:219-221).  Gradients:
 *   dL/dx   = radon_adjoint(shift_adjoint(r)) + lambda (Dx^T qx + Dy^T qy),
 *             r = shift_s(radon(x)) - b, q = Dx / sqrt(.), Dx^T v = v(x+1) - v(x);
 *   dL/ds_t = sum_d r(d,t) (y(d+o,t) - y(d+o+1,t)),  y = radon(x),
 *   hs_t    = sum_d (y(d+o,t) - y(d+o+1,t))^2, the Gauss-Newton curvature
 *             of the data term in s_t (its exact second derivative between
 *             the kinks of the interpolation, where o is constant).
 * loss[0..2] = total, data term, TV_eps; shift, gx, gs, hs may be NULL. */
void FN(loss_grad)(int N, int T, const float *theta, const REAL *x, const float *shift,
                   const REAL *b, double lambda, double eps, double *loss, REAL *gx, REAL *gs,
                   REAL *hs)
{
    const size_t NN = (size_t)N * N, NT = (size_t)N * T;
    REAL *y = (REAL *)malloc(sizeof(REAL) * NT);
    REAL *r = (REAL *)malloc(sizeof(REAL) * NT);
    FN(radon)(N, T, theta, x, y);
    double data = 0.0;
    for (int jt = 0; jt < T; ++jt) {
        int o = 0; REAL w = RC(0.0);
        if (shift) FN(shift_weights)(shift[jt], &o, &w);
        const REAL *row = y + (size_t)jt * N;
        double g = 0.0, h = 0.0;
        for (int d = 1; d <= N; ++d) {
            const REAL ya = FN(at)(row, N, d + o), yb = FN(at)(row, N, d + o + 1);
            const REAL rv = ((RC(1.0) - w) * ya + w * yb) - b[(size_t)jt * N + d - 1];
            r[(size_t)jt * N + d - 1] = rv;
            data += 0.5 * (double)rv * (double)rv;
            g += (double)rv * ((double)ya - (double)yb);
            h += ((double)ya - (double)yb) * ((double)ya - (double)yb);
        }
        if (gs) gs[jt] = (REAL)g;
        if (hs) hs[jt] = (REAL)h;
    }
    double tv = 0.0;
    REAL *qx = (REAL *)malloc(sizeof(REAL) * NN);
    REAL *qy = (REAL *)malloc(sizeof(REAL) * NN);
    for (int jy = 1; jy <= N; ++jy)
        for (int jx = 1; jx <= N; ++jx) {
            const size_t i = (size_t)(jx - 1) + (size_t)(jy - 1) * N;
            const REAL dx = x[(size_t)(TVFN(prv)(jx, N) - 1) + (size_t)(jy - 1) * N] - x[i];
            const REAL dy = x[(size_t)(jx - 1) + (size_t)(TVFN(prv)(jy, N) - 1) * N] - x[i];
            const REAL n = SQRT(dx * dx + dy * dy + (REAL)(eps * eps));
            tv += (double)n;
            qx[i] = dx / n;
            qy[i] = dy / n;
        }
    if (gx) {
        REAL *rs = y;  /* y is no longer needed */
        if (shift) FN(shift_adjoint)(N, T, shift, r, rs);
        else memcpy(rs, r, sizeof(REAL) * NT);
        FN(radon_adjoint)(N, T, theta, rs, gx);
        for (int jy = 1; jy <= N; ++jy)
            for (int jx = 1; jx <= N; ++jx) {
                const size_t i = (size_t)(jx - 1) + (size_t)(jy - 1) * N;
                const REAL t = (qx[(size_t)(TVFN(nxt)(jx, N) - 1) + (size_t)(jy - 1) * N] - qx[i])
                             + (qy[(size_t)(jx - 1) + (size_t)(TVFN(nxt)(jy, N) - 1) * N] - qy[i]);
                gx[i] = gx[i] + (REAL)lambda * t;
            }
    }
    loss[0] = data + lambda * tv;
    loss[1] = data;
    loss[2] = tv;
    free(qx); free(qy); free(y); free(r);
}
#endif /* ORC_TV_ONLY */

/* Split-Bregman TV with inner CG, tv_hpc.F90:162-363.
 *   f         in  (N,T) windowed sinogram (module array f)
 *   img_tv    out (N,N) TV result; img_fbp out (N,N) FBP warm start
 *   delta_log out (nouter) optional Err_diff after each outer iteration
 * The scalar reductions (rsold, num, rsnew, Err_diff) are sequential sums in
 * the array's memory order, in REAL, as in the reference. */
void TVFN(tvopt)(int N, int T, const float *theta, int wpu, const float *f_in,
               REAL mu, REAL lambda, int nouter, int ninner, int ncg,
               REAL *img_tv, REAL *img_fbp, REAL *delta_log)
{
    const size_t NN = (size_t)N * N, NT = (size_t)N * T;
    REAL *h = (REAL *)malloc(sizeof(REAL) * (size_t)(2 * N + 1));
    REAL *cmask = (REAL *)malloc(sizeof(REAL) * NN);
    FN(filter)(N, h);
    FN(cmask)(N, T, wpu, cmask);
    REAL *uk = (REAL *)calloc(NN, sizeof(REAL)), *ukp = (REAL *)calloc(NN, sizeof(REAL));
    REAL *dkx = (REAL *)calloc(NN, sizeof(REAL)), *dky = (REAL *)calloc(NN, sizeof(REAL));
    REAL *bkx = (REAL *)calloc(NN, sizeof(REAL)), *bky = (REAL *)calloc(NN, sizeof(REAL));
    REAL *ifkt = (REAL *)calloc(NN, sizeof(REAL)), *rhs = (REAL *)calloc(NN, sizeof(REAL));
    REAL *iukt = (REAL *)calloc(NN, sizeof(REAL)), *r = (REAL *)calloc(NN, sizeof(REAL));
    REAL *p = (REAL *)calloc(NN, sizeof(REAL)), *irpt = (REAL *)calloc(NN, sizeof(REAL));
    REAL *Ap = (REAL *)calloc(NN, sizeof(REAL));
    REAL *f = (REAL *)malloc(sizeof(REAL) * NT), *fk = (REAL *)malloc(sizeof(REAL) * NT);
    REAL *ruk = (REAL *)malloc(sizeof(REAL) * NT), *rp = (REAL *)malloc(sizeof(REAL) * NT);
    for (size_t i = 0; i < NT; ++i) f[i] = (REAL)f_in[i];

#define IX(x, y) ((size_t)((x) - 1) + (size_t)((y) - 1) * N)
    /* FBP warm start, :176-184 */
    FN(iradon)(N, T, theta, h, cmask, f, uk);
    for (size_t i = 0; i < NN; ++i) img_fbp[i] = uk[i];
    /* :187-202 */
    for (size_t i = 0; i < NN; ++i) dkx[i] = dky[i] = bkx[i] = bky[i] = RC(0.0);
    for (size_t i = 0; i < NT; ++i) fk[i] = f[i];

    const REAL four = RC(4.0);
    const REAL inv_lambda = RC(1.0) / lambda;
    DREAL err_diff = (DREAL)0;
    for (int jouter = 1; jouter <= nouter; ++jouter) {
        for (int jinner = 1; jinner <= ninner; ++jinner) {
            for (size_t i = 0; i < NN; ++i) ukp[i] = uk[i];           /* :207-212 */
            FN(iradon)(N, T, theta, h, cmask, fk, ifkt);               /* :215 */
            for (int y = 1; y <= N; ++y)                               /* :217-225 */
                for (int x = 1; x <= N; ++x) {
                    const int xl = TVFN(nxt)(x, N), yl = TVFN(nxt)(y, N);
                    REAL t = dkx[IX(xl, y)] - bkx[IX(xl, y)];
                    t = t - dkx[IX(x, y)];
                    t = t + bkx[IX(x, y)];
                    t = t + dky[IX(x, yl)];
                    t = t - bky[IX(x, yl)];
                    t = t - dky[IX(x, y)];
                    t = t + bky[IX(x, y)];
                    rhs[IX(x, y)] = mu * ifkt[IX(x, y)] + lambda * t;
                }
            FN(radon)(N, T, theta, uk, ruk);                           /* :228 */
            FN(iradon)(N, T, theta, h, cmask, ruk, iukt);              /* :229 */
            for (int y = 1; y <= N; ++y)                               /* :231-241 */
                for (int x = 1; x <= N; ++x) {
                    const int xr = TVFN(nxt)(x, N), yr = TVFN(nxt)(y, N);
                    const int xl = TVFN(prv)(x, N), yl = TVFN(prv)(y, N);
                    REAL lap = four * uk[IX(x, y)] - uk[IX(xl, y)];
                    lap = lap - uk[IX(xr, y)];
                    lap = lap - uk[IX(x, yl)];
                    lap = lap - uk[IX(x, yr)];
                    const REAL v = (rhs[IX(x, y)] - mu * iukt[IX(x, y)]) - lambda * lap;
                    r[IX(x, y)] = v;
                    p[IX(x, y)] = v;
                }
            DREAL rsold = (DREAL)0;                                    /* :243-249 */
            for (size_t i = 0; i < NN; ++i) rsold = rsold + (DREAL)r[i] * (DREAL)r[i];

            for (int jncg = 1; jncg <= ncg; ++jncg) {                  /* :252-303 */
                FN(radon)(N, T, theta, p, rp);
                FN(iradon)(N, T, theta, h, cmask, rp, irpt);
                for (int y = 1; y <= N; ++y)
                    for (int x = 1; x <= N; ++x) {
                        const int xr = TVFN(nxt)(x, N), yr = TVFN(nxt)(y, N);
                        const int xl = TVFN(prv)(x, N), yl = TVFN(prv)(y, N);
                        REAL lap = four * p[IX(x, y)] - p[IX(xl, y)];
                        lap = lap - p[IX(xr, y)];
                        lap = lap - p[IX(x, yl)];
                        lap = lap - p[IX(x, yr)];
                        Ap[IX(x, y)] = mu * irpt[IX(x, y)] + lambda * lap;
                    }
                DREAL num = (DREAL)0;
                for (size_t i = 0; i < NN; ++i) num = num + (DREAL)p[i] * (DREAL)Ap[i];
                const REAL alpha = (REAL)(rsold / num);
                for (size_t i = 0; i < NN; ++i) {
                    uk[i] = uk[i] + alpha * p[i];
                    r[i] = r[i] - alpha * Ap[i];
                }
                DREAL rsnew = (DREAL)0;
                for (size_t i = 0; i < NN; ++i) rsnew = rsnew + (DREAL)r[i] * (DREAL)r[i];
                const REAL beta = (REAL)(rsnew / rsold);
                for (size_t i = 0; i < NN; ++i) p[i] = r[i] + beta * p[i];
                rsold = rsnew;
            }
            for (int y = 1; y <= N; ++y)                               /* :305-325 */
                for (int x = 1; x <= N; ++x) {
                    const int xl = TVFN(prv)(x, N), yl = TVFN(prv)(y, N);
                    REAL t = (uk[IX(xl, y)] - uk[IX(x, y)]) + bkx[IX(x, y)];
                    REAL d;
                    if (t == RC(0.0)) d = RC(0.0);
                    else {
                        REAL a = FABS(t) - inv_lambda;
                        if (!(a > RC(0.0))) a = RC(0.0);
                        d = (a * t) / FABS(t);
                    }
                    dkx[IX(x, y)] = d;
                    bkx[IX(x, y)] = t - d;
                    t = (bky[IX(x, y)] + uk[IX(x, yl)]) - uk[IX(x, y)];
                    if (t == RC(0.0)) d = RC(0.0);
                    else {
                        REAL a = FABS(t) - inv_lambda;
                        if (!(a > RC(0.0))) a = RC(0.0);
                        d = (a * t) / FABS(t);
                    }
                    dky[IX(x, y)] = d;
                    bky[IX(x, y)] = t - d;
                }
            err_diff = (DREAL)0;                                       /* :327-333 */
            for (size_t i = 0; i < NN; ++i) err_diff = err_diff + (DREAL)FABS(uk[i] - ukp[i]);
        }
        FN(radon)(N, T, theta, uk, ruk);                               /* :339 */
        for (size_t i = 0; i < NT; ++i) fk[i] = (fk[i] + f[i]) - ruk[i]; /* :341-346 */
        if (delta_log) delta_log[jouter - 1] = (REAL)err_diff;
    }
#undef IX
    for (size_t i = 0; i < NN; ++i) img_tv[i] = uk[i];
    free(h); free(cmask); free(uk); free(ukp); free(dkx); free(dky); free(bkx); free(bky);
    free(ifkt); free(rhs); free(iukt); free(r); free(p); free(irpt); free(Ap);
    free(f); free(fk); free(ruk); free(rp);
}

#undef RC
