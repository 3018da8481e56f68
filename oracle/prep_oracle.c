/*
 * prep_oracle.c -- CPU oracle of the projection -> sinogram preprocessing of
 * Acumen (tomography/projections_to_sinogram_hpc.F90), in the reference's
 * FP32 arithmetic (one rounding per Fortran operation, no contraction).
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by the test-suite; the product library
 * never links, loads or calls it.
 *
 * Arrays follow the reference's Fortran layout, first index fastest: a
 * projection is image(x, y) of nimage x nimage (x = detector column, y =
 * detector row = slice); a corrected projection is sino(x, y) of
 * Nw x nimage with Nw = nimage + add_Nx (add_Nx = 1 for even nimage).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }

/* 3x3x3 median of a projection and its two neighbours, clamped at the image
 * edges (projections_to_sinogram_hpc.F90:420-447): the 14th of the 27 values
 * in descending order.
 *
 * This is the median the reference intends, not the one its code computes.
 * It stores neighbour (jxs, jys) of datal at register slot
 * 3*(jxs+1-1)+jys+2 = 3*jxs+jys+2 (:433), which runs from -2 to 6, of datac
 * at +9 (7..15) and of datar at +18 (16..24), in an array allocated
 * register(nimage,nimage,27) (:134).  So the jxs = -1 column of datal is
 * written out of bounds (slots -2..0, before the array), and slots 25-27 are
 * never written: the sort (:438-447) then ranks 24 neighbours with 3 values
 * left over from the previous projection at that pixel (or never set, at the
 * first).  That result depends on undefined memory, so the oracle and the
 * GPU kernel take the 27-neighbour median the loop nest describes. */
void orc_prep_median27(int n, const float *l, const float *c, const float *r, float *out)
{
    for (int jy = 1; jy <= n; ++jy)
        for (int jx = 1; jx <= n; ++jx) {
            float reg[27];
            int k = 0;
            for (int jxs = -1; jxs <= 1; ++jxs)
                for (int jys = -1; jys <= 1; ++jys) {
                    const int cx = clampi(jx + jxs, 1, n), cy = clampi(jy + jys, 1, n);
                    const size_t i = (size_t)(cx - 1) + (size_t)(cy - 1) * n;
                    reg[k] = l[i];
                    reg[k + 9] = c[i];
                    reg[k + 18] = r[i];
                    ++k;
                }
            for (int a = 0; a < 26; ++a) {   /* the reference's exchange sort, descending */
                float ra = reg[a];
                for (int b = a + 1; b < 27; ++b) {
                    const float rb = reg[b];
                    if (rb > ra) { reg[b] = ra; reg[a] = rb; ra = rb; }
                }
            }
            out[(size_t)(jx - 1) + (size_t)(jy - 1) * n] = reg[13];
        }
}

/* The rotation / shift resampling of :449-464 (and of :260-273 with the
 * search's trial angles): sample (jx, jy), jx = 1..Nw, jy = 1..n. */
static float prep_sample(int n, float ct, float st, float mx, float my, float S, float shift,
                         const float *img, int jx, int jy)
{
    const float u = ((float)jx - mx) - S * shift;
    const float xr = clampf((ct * u - st * ((float)jy - my)) + mx, 1.0f, (float)n);
    const float yr = clampf((st * u + ct * ((float)jy - my)) + my, 1.0f, (float)n);
    const int y1 = clampi((int)floorf(yr), 1, n), y2 = clampi(y1 + 1, 1, n);
    const int x1 = clampi((int)floorf(xr), 1, n), x2 = clampi(x1 + 1, 1, n);
    const float ax = (float)x2 - xr, bx = xr - (float)x1, ay = (float)y2 - yr, by = yr - (float)y1;
#define IM(x, y) img[(size_t)((x) - 1) + (size_t)((y) - 1) * n]
    float v = (IM(x1, y1) * ax) * ay;
    v = v + (IM(x2, y1) * bx) * ay;
    v = v + (IM(x1, y2) * ax) * by;
    v = v + (IM(x2, y2) * bx) * by;
#undef IM
    return v;
}

void orc_prep_correct(int n, int add_nx, float theta, float shift, const float *img, float *out)
{
    const int nw = n + add_nx;
    const float mx = (float)(nw - 1) / 2.0f, my = (float)(n - 1) / 2.0f, S = (float)n;
    const float ct = cosf(theta), st = sinf(theta);
    for (int jy = 1; jy <= n; ++jy)
        for (int jx = 1; jx <= nw; ++jx)
            out[(size_t)(jx - 1) + (size_t)(jy - 1) * nw] = prep_sample(n, ct, st, mx, my, S, shift, img, jx, jy);
}

/* Mirror error of :278-283 for one (theta, shift): sum |V0(jx,jy) - V180(Nw-jx+1,jy)|
 * with V = the resampled 0 and 180 degree projections; summed in FP64. */
double orc_prep_mirror_error(int n, int add_nx, float theta, float shift, const float *p0, const float *p180)
{
    const int nw = n + add_nx;
    const float mx = (float)(nw - 1) / 2.0f, my = (float)(n - 1) / 2.0f, S = (float)n;
    const float ct = cosf(theta), st = sinf(theta);
    double e = 0.0;
    for (int jy = 1; jy <= n; ++jy)
        for (int jx = 1; jx <= nw; ++jx) {
            const float v0 = prep_sample(n, ct, st, mx, my, S, shift, p0, jx, jy);
            const float v1 = prep_sample(n, ct, st, mx, my, S, shift, p180, nw - jx + 1, jy);
            e += (double)fabsf(v0 - v1);
        }
    return e;
}

/* First centre-of-rotation estimate from the centroids of the 0 and 180
 * degree projections (:231-248). */
float orc_prep_centroid_shift(int n, int add_nx, const float *p0, const float *p180)
{
    const int nw = n + add_nx;
    double tm = 0.0, c0 = 0.0, c180 = 0.0;
    for (int jx = 1; jx <= n; ++jx)
        for (int jy = 1; jy <= n; ++jy) {
            const size_t i = (size_t)(jx - 1) + (size_t)(jy - 1) * n;
            const float w = (2.0f * (float)nw * (float)(jx - 1)) / (float)(nw - 1) - (float)n - (float)add_nx;
            tm += (double)p0[i] + (double)p180[i];
            c0 += (double)p0[i] * (double)w;
            c180 += (double)p180[i] * (double)w;
        }
    return -((float)(c0 / tm) + (float)(c180 / tm)) / (2.0f * (float)n);
}

/* The pattern search of :250-364 on (theta, shift), FP32 parameters and
 * steps as in the reference, error sums in FP64. */
void orc_prep_find_center(int n, int add_nx, const float *p0, const float *p180, float *theta_out,
                          float *shift_out)
{
    float best_shift = orc_prep_centroid_shift(n, add_nx, p0, p180), best_theta = 0.0f;
    float dshift = 1e-4f, dtheta = 1e-4f;
    const int n_sp = 6;
    const float n_mag = 1.05f, n_min = 2.0f / (float)n_sp;
    double best_error = orc_prep_mirror_error(n, add_nx, best_theta, best_shift, p0, p180);
    while ((dtheta > dshift ? dtheta : dshift) > 1e-8f) {
        double perror = best_error;
        for (int jsp = 1; jsp <= n_sp; ++jsp) {
            const float th = best_theta + dtheta * (float)(jsp - n_sp / 2);
            const double e = orc_prep_mirror_error(n, add_nx, th, best_shift, p0, p180);
            if (e < best_error) { best_theta = th; best_error = e; }
        }
        dtheta = best_error < perror ? n_mag * dtheta : n_min * dtheta;
        perror = best_error;
        for (int jsp = 1; jsp <= n_sp; ++jsp) {
            const float sh = best_shift + dshift * (float)(jsp - n_sp / 2);
            const double e = orc_prep_mirror_error(n, add_nx, best_theta, sh, p0, p180);
            if (e < best_error) { best_shift = sh; best_error = e; }
        }
        dshift = best_error < perror ? n_mag * dshift : n_min * dshift;
    }
    *theta_out = best_theta;
    *shift_out = best_shift;
}

/* ---- FITS ingest normalisation (tomography/fits_to_hpc.F90) ---- */

/* Open-beam average of :203-215: bg = 0; bg = bg + image for each file in
 * order; bg = bg / nfiles_bg, all FP32. */
void orc_ingest_mean(int count, long long npix, const float *imgs, float *out)
{
    for (long long i = 0; i < npix; ++i) {
        float s = 0.0f;
        for (int k = 0; k < count; ++k) s = s + imgs[(long long)k * npix + i];
        out[i] = s / (float)count;
    }
}

/* projection = -log(image / max(bg, 1.0)) of :243; gfortran evaluates the
 * FP32 log with the C library's logf, as here. */
void orc_ingest_normalize(int count, long long npix, const float *imgs, const float *bg, float *out)
{
    for (int k = 0; k < count; ++k)
        for (long long i = 0; i < npix; ++i) {
            const float b = bg[i] > 1.0f ? bg[i] : 1.0f;
            out[(long long)k * npix + i] = -logf(imgs[(long long)k * npix + i] / b);
        }
}
