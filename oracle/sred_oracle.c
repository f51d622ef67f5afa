/*
 * sred_oracle.c -- CPU restatement of SReD's 2D block-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See sred_oracle.h for the two
 * forms (refsim float32-faithful, oracle64 FP64) and their rules.  Built
 * with -ffp-contract=off so every float expression rounds exactly as written
 * (Java never contracts; the reference's OpenCL compiler may, which is the
 * one unpinned freedom of the kernels' float order).
 *
 * Paths in citations are relative to the SReD repository root.
 */
#include "sred_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS_F 1e-7f   /* Utils.EPSILON, Utils.java:20 */

/* ---- footprint: GlobalRepetition2D_.java:102-111, CLUtils.java:391-400 ---- */
int or_footprint(int bw, int bh, int *dx, int *dy)
{
    int brw = bw / 2, brh = bh / 2, k = 0;
    for (int y = 0; y < bh; y++) {
        for (int x = 0; x < bw; x++) {
            float fx = (float)(x - brw), fy = (float)(y - brh);
            /* float division by (float)(r*r): with r=0 this is x/0 -> inf/nan,
             * exactly as the reference (only odd sizes >= 3 are meaningful) */
            if (((fx * fx) / (float)(brw * brw)) + ((fy * fy) / (float)(brh * brh)) <= 1.0f) {
                if (dx) { dx[k] = x - brw; dy[k] = y - brh; }
                k++;
            }
        }
    }
    return k;
}

/* ---- Java ""+float -> OpenCL double literal (CLUtils.java:1705-1712) ---- */
double or_java_float_literal(float f)
{
    char buf[64];
    if (f == 0.0f) return 0.0;
    for (int p = 1; p <= 9; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, (double)f);
        if (strtof(buf, NULL) == f) return strtod(buf, NULL);
    }
    return (double)f;
}

/* ---- Utils.java:708-722 ---- */
void or_normalize_input(float *img, int64_t n)
{
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int64_t i = 0; i < n; i++) {
        mn = fminf(mn, img[i]);
        mx = fmaxf(mx, img[i]);
    }
    for (int64_t i = 0; i < n; i++) img[i] = (img[i] - mn) / (mx - mn + EPS_F);
}

/* ---- Utils.java:557-639 ---- */
int or_reference_block(const float *blk, int bw, int bh, float *px, float *mean, float *std_)
{
    if (bw % 2 == 0 || bh % 2 == 0) return -1;               /* Utils.java:574-577 */
    int brw = bw / 2, brh = bh / 2, k = 0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            float fx = (float)(i - brw), fy = (float)(j - brh);
            if (((fx * fx) / (float)(brw * brw)) + ((fy * fy) / (float)(brh * brh)) <= 1.0f) px[k++] = blk[j * bw + i];
        }
    float bmin = FLT_MAX, bmax = -FLT_MAX;
    for (int i = 0; i < k; i++) { bmin = fminf(bmin, px[i]); bmax = fmaxf(bmax, px[i]); }
    float m = 0.0f;
    for (int i = 0; i < k; i++) {
        px[i] = (px[i] - bmin) / (bmax - bmin + EPS_F);
        m += px[i];
    }
    m /= (float)k;
    for (int i = 0; i < k; i++) px[i] = px[i] - m;
    float s = 0.0f;
    for (int i = 0; i < k; i++) s += (px[i] - m) * (px[i] - m);  /* Utils.java:632-634: mean subtracted twice */
    s = (float)sqrt((double)(s / (float)(k - 1)));
    *mean = m;
    *std_ = s;
    return k;
}

/* Entry points with per-thread block buffers of OR_MAX_K elements refuse
 * larger footprints by filling their outputs with NaN (or returning -1). */
static void or_fill_nan(float *a, float *b, int64_t n)
{
    for (int64_t i = 0; i < n; i++) { a[i] = NAN; b[i] = NAN; }
}
static void or_fill_nan64(double *a, double *b, int64_t n)
{
    for (int64_t i = 0; i < n; i++) { a[i] = NAN; b[i] = NAN; }
}

/* ---- kernelGetLocalStatistics2D.cl:9-55 ---- */
void or_local_stats_f32(const float *img, int w, int h, int bw, int bh, float *means, float *stds)
{
    int brw = bw / 2, brh = bh / 2;
    int K = or_footprint(bw, bh, NULL, NULL);
    if (K > OR_MAX_K) { or_fill_nan(means, stds, (int64_t)w * h); return; }
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < h; gy++) {
        float blk[OR_MAX_K];
        for (int gx = 0; gx < w; gx++) {
            int64_t o = (int64_t)gy * w + gx;
            if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) { means[o] = 0.0f; stds[o] = 0.0f; continue; }
            for (int i = 0; i < K; i++) blk[i] = img[(int64_t)(gy + dy[i]) * w + gx + dx[i]];
            float m = 0.0f;
            for (int i = 0; i < K; i++) m += blk[i];
            m /= (float)K;
            float v = 0.0f;
            for (int i = 0; i < K; i++) v += (blk[i] - m) * (blk[i] - m);
            v /= (float)(K - 1);
            means[o] = m;
            stds[o] = sqrtf(v);
        }
    }
    free(dx); free(dy);
}

void or_local_stats_f64(const float *img, int w, int h, int bw, int bh, double *means, double *vars)
{
    int brw = bw / 2, brh = bh / 2;
    int K = or_footprint(bw, bh, NULL, NULL);
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < h; gy++)
        for (int gx = 0; gx < w; gx++) {
            int64_t o = (int64_t)gy * w + gx;
            if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) { means[o] = 0.0; vars[o] = 0.0; continue; }
            double m = 0.0;
            for (int i = 0; i < K; i++) m += img[(int64_t)(gy + dy[i]) * w + gx + dx[i]];
            m /= K;
            double v = 0.0;
            for (int i = 0; i < K; i++) {
                double d = img[(int64_t)(gy + dy[i]) * w + gx + dx[i]] - m;
                v += d * d;
            }
            means[o] = m;
            vars[o] = v / (K - 1);
        }
    free(dx); free(dy);
}

/* ---- Utils.java:854-900 ---- */
int or_mean_var_block(const float *img, int w, int h, int x0, int y0, int x1, int y1, float *mean, float *var)
{
    int bw = x1 - x0, bh = y1 - y0, n = bw * bh;
    if (bw <= 0 || bh <= 0 || n <= 0) return -1;
    if (x0 < 0 || y0 < 0 || x1 > w || y1 > h) return -1;
    float m = 0.0f, sq = 0.0f;
    for (int y = y0; y < y1; y++)
        for (int x = x0; x < x1; x++) {
            float v = img[(int64_t)y * w + x];
            if (isnan(v) || isinf(v)) return -1;
            m += v;
            sq += v * v;
        }
    m /= (float)n;
    float va = sq / (float)n - m * m;
    if (isnan(va)) return -1;
    *mean = m;
    *var = va;
    return 0;
}

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* ---- Utils.java:790-837 ---- */
float or_noise_var(const float *img, int w, int h)
{
    int bw, bh;
    if (w * h <= 352 * 288) { bw = 8; bh = 8; } else { bw = 16; bh = 16; }
    int nbx = w / bw, nby = h / bh, nb = nbx * nby;
    float *lv = calloc(nb > 0 ? nb : 1, sizeof(float));
    int idx = 0;
    for (int y = 0; y < nby; y++)
        for (int x = 0; x < nbx; x++) {
            float m, v;
            or_mean_var_block(img, w, h, x * bw, y * bh, (x + 1) * bw, (y + 1) * bh, &m, &v);
            lv[idx++] = v;
        }
    qsort(lv, nb, sizeof(float), cmp_float);
    int nvars = (int)(0.03f * (float)nb + 1.0f);
    float nv = 0.0f;
    for (int i = 0; i < nvars && i < nb; i++) nv += lv[i];
    nv = fabsf(nv / (float)nvars);
    nv = (1.0f + 0.001f * (nv - 40.0f)) * nv;
    free(lv);
    return nv;
}

/* ---- Utils.java:944-968 and GlobalRepetition2D_.java:133-140 ---- */
void or_relevance_mask(const float *stds, int w, int h, int brw, int brh, float noise_var, float c,
                       float *mask, float *threshold, float *n_pixels)
{
    float thr = noise_var * c;
    memset(mask, 0, sizeof(float) * (size_t)w * h);
    float np = 0.0f;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            float s = stds[(int64_t)y * w + x];
            mask[(int64_t)y * w + x] = (s * s <= thr) ? 0.0f : 1.0f;
            if (mask[(int64_t)y * w + x] > 0.0f) np += 1.0f;
        }
    *threshold = thr;
    *n_pixels = np;
}

void or_apply_mask(float *map, int64_t n, const float *mask)
{
    for (int64_t i = 0; i < n; i++) map[i] *= mask[i];
}

/* ---- Utils.java:985-1031 ---- */
int or_normalize_map(float *map, int w, int h, int brw, int brh, const float *mask)
{
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask) {
                if (mask[i] != 0.0f && mask[i] != 1.0f) return -1;
                if (!(mask[i] > 0.0f)) continue;
            }
            mn = fminf(map[i], mn);
            mx = fmaxf(map[i], mx);
        }
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask && !(mask[i] > 0.0f)) continue;
            map[i] = (map[i] - mn) / (mx - mn + EPS_F);
        }
    return 0;
}

int or_normalize_map_f64(double *map, int w, int h, int brw, int brh, const float *mask)
{
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask) {
                if (mask[i] != 0.0f && mask[i] != 1.0f) return -1;
                if (!(mask[i] > 0.0f)) continue;
            }
            mn = fmin(map[i], mn);
            mx = fmax(map[i], mx);
        }
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask && !(mask[i] > 0.0f)) continue;
            map[i] = (map[i] - mn) / (mx - mn + 1e-7);
        }
    return 0;
}

/* ---- global accumulators, float32 refsim ----
 * kernelGetGlobalPearson2D.cl:67-166, kernelGetGlobalSsim2D.cl:11-118,
 * kernelGetGlobalNrmse2D.cl:11-125, kernelGetGlobalAbsDiffStds2D.cl:10-53.
 * The kernels re-evaluate the ellipse predicate per pair; its outcome is
 * data-independent, so the offsets are evaluated once (identical membership
 * and order).  Accumulation over test pixels is sequential in row-major
 * order, as in each work-item of the reference. */
void or_global_accum_f32(int metric, const float *img, const float *means, const float *stds, int w, int h,
                         int bw, int bh, float filter_param, float threshold, float *S, float *Wsum,
                         const int32_t *rows, int64_t nrows, int nthreads)
{
    const int brw = bw / 2, brh = bh / 2;
    const int K = or_footprint(bw, bh, NULL, NULL);
    if (K > OR_MAX_K) { or_fill_nan(S, Wsum, (int64_t)w * h); return; }
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
    const double fp_d = or_java_float_literal(filter_param);
    const double thr_d = or_java_float_literal(threshold);
    const double eps_d = or_java_float_literal(EPS_F);
    int64_t *off = malloc(sizeof(int64_t) * K);
    for (int i = 0; i < K; i++) off[i] = (int64_t)dy[i] * w + dx[i];
    int64_t total = rows ? nrows : (int64_t)w * h;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int64_t r = 0; r < total; r++) {
        int64_t o = rows ? rows[r] : r;
        int gx = (int)(o % w), gy = (int)(o / w);
        float acc = 0.0f, wacc = 0.0f;
        float rb[OR_MAX_K], tb[OR_MAX_K];
        if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
            /* border: map written 0 (NRMSE leaves it unwritten), weights untouched */
            S[o] = 0.0f; Wsum[o] = 0.0f;
            continue;
        }
        const float rs = stds[o];
        const float rv = rs * rs;
        if ((double)rv < thr_d) { S[o] = 0.0f; Wsum[o] = 0.0f; continue; }
        const float rm = (metric == OR_ABSDIFF) ? 0.0f : means[o];
        if (metric != OR_ABSDIFF) {
            for (int i = 0; i < K; i++) rb[i] = img[o + off[i]];
            for (int i = 0; i < K; i++) rb[i] = rb[i] - rm;
        }
        for (int y = brh; y < h - brh; y++) {
            for (int x = brw; x < w - brw; x++) {
                int64_t t = (int64_t)y * w + x;
                const float ts = stds[t];
                const float tv = ts * ts;
                if ((double)tv < thr_d) continue;
                if (metric == OR_ABSDIFF) {
                    /* kernelGetGlobalAbsDiffStds2D.cl:46-50: exp evaluated in double */
                    float d = rs - ts;
                    float wt = (float)exp(-1.0 * ((double)(d * d) / (fp_d + eps_d)));
                    wacc += wt;
                    acc += fabsf(rs - ts) * wt;
                    continue;
                }
                const float tm = means[t];
                for (int i = 0; i < K; i++) tb[i] = img[t + off[i]];
                for (int i = 0; i < K; i++) tb[i] = tb[i] - tm;
                float d = rs - ts;
                float q = (float)((double)(d * d) / (fp_d + eps_d));
                float wt = (float)expf((float)(-1.0f) * q);
                wacc += wt;
                if (metric == OR_NRMSE) {
                    float mse = 0.0f;
                    for (int i = 0; i < K; i++) mse += (rb[i] - tb[i]) * (rb[i] - tb[i]);
                    mse /= (float)K;
                    acc += sqrtf(mse);
                    continue;
                }
                float cov = 0.0f;
                for (int i = 0; i < K; i++) cov += rb[i] * tb[i];
                cov /= (float)(K - 1);
                if (metric == OR_PEARSON) {
                    if (rs == 0.0f && ts == 0.0f) acc += 1.0f * wt;
                    else if (rs == 0.0f || ts == 0.0f) acc += 0.0f;
                    else {
                        float r = (float)((double)cov / ((double)(rs * ts) + eps_d));
                        acc += fmaxf(0.0f, r) * wt;
                    }
                } else { /* SSIM, kernelGetGlobalSsim2D.cl:100-115 */
                    const float c1 = 0.0001f, c2 = 0.0009f;
                    if (rs == 0.0f && ts == 0.0f) acc += 1.0f * wt;
                    else if (rs == 0.0f || ts == 0.0f) acc += 0.0f;
                    else {
                        float num = (2.0f * rm * tm + c1) * (2.0f * cov + c2);
                        float den = ((rm * rm) * (tm * tm) + c1) * ((rs * rs) + (ts * ts) + c2);
                        float ssim = num / den * wt;
                        acc += fmaxf(0.0f, ssim);
                    }
                }
            }
        }
        S[o] = acc;
        Wsum[o] = wacc;
    }
    free(dx); free(dy); free(off);
}

/* ---- global accumulators, oracle64 ---- */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum *k, double v)
{   /* Neumaier compensated summation */
    double t = k->s + v;
    if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v; else k->c += (v - t) + k->s;
    k->s = t;
}

void or_global_accum_f64(int metric, const float *img, const float *stds32, int w, int h, int bw, int bh,
                         float filter_param, float threshold, double *S, double *Wsum,
                         const int32_t *rows, int64_t nrows, int nthreads)
{
    const int brw = bw / 2, brh = bh / 2;
    const int K = or_footprint(bw, bh, NULL, NULL);
    if (K > OR_MAX_K) { or_fill_nan64(S, Wsum, (int64_t)w * h); return; }
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
    const double fp_d = or_java_float_literal(filter_param);
    const double thr_d = or_java_float_literal(threshold);
    const double eps = 1e-7;
    int64_t n = (int64_t)w * h;
    double *m64 = malloc(sizeof(double) * n), *v64 = malloc(sizeof(double) * n);
    or_local_stats_f64(img, w, h, bw, bh, m64, v64);
    /* interior valid columns, compacted, with centred FP64 block vectors */
    int64_t nval = 0;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            float s = stds32[(int64_t)y * w + x];
            if (!((double)(s * s) < thr_d)) nval++;
        }
    int64_t *vidx = malloc(sizeof(int64_t) * (nval ? nval : 1));
    double *vec = malloc(sizeof(double) * (nval ? nval : 1) * K);
    nval = 0;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t t = (int64_t)y * w + x;
            float s = stds32[t];
            if ((double)(s * s) < thr_d) continue;
            for (int i = 0; i < K; i++) vec[nval * K + i] = (double)img[t + (int64_t)dy[i] * w + dx[i]] - m64[t];
            vidx[nval++] = t;
        }
    int64_t total = rows ? nrows : n;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int64_t r = 0; r < total; r++) {
        int64_t o = rows ? rows[r] : r;
        int gx = (int)(o % w), gy = (int)(o / w);
        if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) { S[o] = 0.0; Wsum[o] = 0.0; continue; }
        float rs32 = stds32[o];
        if ((double)(rs32 * rs32) < thr_d) { S[o] = 0.0; Wsum[o] = 0.0; continue; }
        double rs = sqrt(v64[o]), rm = m64[o];
        double rb[OR_MAX_K];
        for (int i = 0; i < K; i++) rb[i] = (double)img[o + (int64_t)dy[i] * w + dx[i]] - rm;
        ksum acc = {0, 0}, wacc = {0, 0};
        for (int64_t j = 0; j < nval; j++) {
            int64_t t = vidx[j];
            double ts = sqrt(v64[t]), tm = m64[t];
            double d = rs - ts;
            double wt = exp(-(d * d) / (fp_d + eps));
            kadd(&wacc, wt);
            if (metric == OR_ABSDIFF) { kadd(&acc, fabs(d) * wt); continue; }
            const double *tb = vec + j * K;
            if (metric == OR_NRMSE) {
                double mse = 0.0;
                for (int i = 0; i < K; i++) { double e = rb[i] - tb[i]; mse += e * e; }
                kadd(&acc, sqrt(mse / K));
                continue;
            }
            double cov = 0.0;
            for (int i = 0; i < K; i++) cov += rb[i] * tb[i];
            cov /= (K - 1);
            int rz = (rs32 == 0.0f), tz = (stds32[t] == 0.0f);
            if (rz && tz) { kadd(&acc, wt); continue; }
            if (rz || tz) continue;
            if (metric == OR_PEARSON) {
                double rr = cov / (rs * ts + eps);
                kadd(&acc, (rr > 0.0 ? rr : 0.0) * wt);
            } else {
                const double c1 = 0.0001f, c2 = 0.0009f;  /* float constants, as in the kernel */
                double ss = (2.0 * rm * tm + c1) * (2.0 * cov + c2) / ((rm * rm) * (tm * tm) + c1) /
                            (rs * rs + ts * ts + c2) * wt;
                kadd(&acc, ss > 0.0 ? ss : 0.0);
            }
        }
        S[o] = acc.s + acc.c;
        Wsum[o] = wacc.s + wacc.c;
    }
    free(dx); free(dy); free(m64); free(v64); free(vidx); free(vec);
}

/* ---- host finalize: CLUtils.java:1747-1770 (Pearson), 1865-1887 (SSIM),
 *      1983-2020 (NRMSE), 2112-2149 (AbsDiffStds) ---- */
int or_global_finalize(int metric, const float *S, const float *Wsum, const float *mask, float n_pixels,
                       float c, int normalize, int w, int h, int bw, int bh, float *out)
{
    int brw = bw / 2, brh = bh / 2;
    memset(out, 0, sizeof(float) * (size_t)w * h);
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            out[i] = S[i] / (Wsum[i] * n_pixels + EPS_F);
        }
    if (metric == OR_NRMSE || metric == OR_ABSDIFF) {
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) {
                int64_t i = (int64_t)y * w + x;
                if (mask[i] > 0.0f) out[i] = (out[i] == 0.0f) ? 1.0f : 1.0f / out[i];
            }
    }
    if (c > 0.0f) or_apply_mask(out, (int64_t)w * h, mask);
    if (normalize) return or_normalize_map(out, w, h, brw, brh, mask);
    return 0;
}

int or_global_finalize_f64(int metric, const double *S, const double *Wsum, const float *mask, double n_pixels,
                           float c, int normalize, int w, int h, int bw, int bh, double *out)
{
    int brw = bw / 2, brh = bh / 2;
    memset(out, 0, sizeof(double) * (size_t)w * h);
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            out[i] = S[i] / (Wsum[i] * n_pixels + 1e-7);
        }
    if (metric == OR_NRMSE || metric == OR_ABSDIFF) {
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) {
                int64_t i = (int64_t)y * w + x;
                if (mask[i] > 0.0f) out[i] = (out[i] == 0.0) ? 1.0 : 1.0 / out[i];
            }
    }
    if (c > 0.0f)
        for (int64_t i = 0; i < (int64_t)w * h; i++) out[i] *= mask[i];
    if (normalize) return or_normalize_map_f64(out, w, h, brw, brh, mask);
    return 0;
}

/* ---- GlobalRepetition2D_.java:124-145 (VST off, input already normalised) ---- */
int or_global_map(int metric, int fp64, const float *img, int w, int h, int bw, int bh, float c, int normalize,
                  float *out32, double *out64, int nthreads)
{
    int64_t n = (int64_t)w * h;
    float *means = malloc(sizeof(float) * n), *stds = malloc(sizeof(float) * n), *mask = malloc(sizeof(float) * n);
    or_local_stats_f32(img, w, h, bw, bh, means, stds);
    float nv = or_noise_var(img, w, h), thr, np;
    or_relevance_mask(stds, w, h, bw / 2, bh / 2, nv, c, mask, &thr, &np);
    int rc;
    if (!fp64) {
        float *S = malloc(sizeof(float) * n), *W = malloc(sizeof(float) * n);
        or_global_accum_f32(metric, img, means, stds, w, h, bw, bh, nv, thr, S, W, NULL, 0, nthreads);
        rc = or_global_finalize(metric, S, W, mask, np, c, normalize, w, h, bw, bh, out32);
        free(S); free(W);
    } else {
        double *S = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n);
        or_global_accum_f64(metric, img, stds, w, h, bw, bh, nv, thr, S, W, NULL, 0, nthreads);
        rc = or_global_finalize_f64(metric, S, W, mask, (double)np, c, normalize, w, h, bw, bh, out64);
        free(S); free(W);
    }
    free(means); free(stds); free(mask);
    return rc;
}

/* ---- block repetition kernels: kernelGetBlockPearson2D.cl:12-90,
 *      kernelGetBlockSsim2D.cl:103-172, kernelGetBlockNrmse2D.cl:12-75,
 *      kernelGetBlockAbsDiffStds2D.cl:84-106 ---- */
int or_block_raw(int metric, const float *img, const float *means, const float *stds, int w, int h,
                 const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std, float *out)
{
    const int brw = bw / 2, brh = bh / 2;
    const int K = or_footprint(bw, bh, NULL, NULL);
    if (K > OR_MAX_K) return -1;
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
    const double eps_d = or_java_float_literal(EPS_F);
    const double std_d = or_java_float_literal(tmpl_std);
    const double mean_d = or_java_float_literal(tmpl_mean);
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < h; gy++) {
        float cb[OR_MAX_K];
        for (int gx = 0; gx < w; gx++) {
            int64_t o = (int64_t)gy * w + gx;
            if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) { out[o] = 0.0f; continue; }
            if (metric == OR_ABSDIFF) { out[o] = fabsf(tmpl_std - stds[o]); continue; }
            const float cm = means[o];
            for (int i = 0; i < K; i++) cb[i] = img[(int64_t)(gy + dy[i]) * w + gx + dx[i]];
            for (int i = 0; i < K; i++) cb[i] = cb[i] - cm;
            if (metric == OR_NRMSE) {
                float mse = 0.0f;
                for (int i = 0; i < K; i++) mse += (tmpl[i] - cb[i]) * (tmpl[i] - cb[i]);
                mse /= (float)K;
                out[o] = sqrtf(mse);
                continue;
            }
            float cov = 0.0f;
            for (int i = 0; i < K; i++) cov += tmpl[i] * cb[i];
            cov /= (float)(K - 1);
            const float cs = stds[o];
            if (std_d == 0.0 && cs == 0.0f) { out[o] = 1.0f; continue; }
            if (std_d == 0.0 || cs == 0.0f) { out[o] = 0.0f; continue; }
            if (metric == OR_PEARSON) {
                float r = (float)((double)cov / ((std_d * (double)cs) + eps_d));
                out[o] = fmaxf(0.0f, r);
            } else {
                const float c1 = 0.0001f, c2 = 0.0009f;
                /* ref_mean/ref_std are double literals: the expression is double */
                double num = (2.0 * mean_d * (double)cm + (double)c1) * (double)(2.0f * cov + c2);
                double den = ((mean_d * mean_d) + (double)(cm * cm) + (double)c1) *
                             ((std_d * std_d) + (double)(cs * cs) + (double)c2);
                float ss = (float)(num / den);
                out[o] = fmaxf(0.0f, ss);
            }
        }
    }
    free(dx); free(dy);
    return 0;
}

/* CLUtils.java:486-981 host sequences around the block kernels */
int or_block_map(int metric, const float *img, int w, int h, const float *tmpl, int bw, int bh, float tmpl_mean,
                 float tmpl_std, float c, int normalize, float *out)
{
    const int brw = bw / 2, brh = bh / 2;
    int64_t n = (int64_t)w * h;
    float *means = malloc(sizeof(float) * n), *stds = malloc(sizeof(float) * n);
    or_local_stats_f32(img, w, h, bw, bh, means, stds);
    int rc = or_block_raw(metric, img, means, stds, w, h, tmpl, bw, bh, tmpl_mean, tmpl_std, out);
    float *mask = NULL;
    float thr, np;
    if (rc) goto done;
    if (metric == OR_PEARSON || metric == OR_SSIM) {            /* CLUtils.java:565-577 */
        if (c > 0.0f) {
            mask = malloc(sizeof(float) * n);
            or_relevance_mask(stds, w, h, brw, brh, or_noise_var(img, w, h), c, mask, &thr, &np);
            or_apply_mask(out, n, mask);
        }
        if (normalize) rc = or_normalize_map(out, w, h, brw, brh, mask);
    } else if (metric == OR_NRMSE) {                            /* CLUtils.java:788-841 */
        /* readback copies only the interior; border stays 0 (CLUtils.java:789-795) */
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                if (x < brw || x >= w - brw || y < brh || y >= h - brh) out[(int64_t)y * w + x] = 0.0f;
        mask = malloc(sizeof(float) * n);
        if (c == 0.0f) {
            memset(mask, 0, sizeof(float) * n);
            for (int y = brh; y < h - brh; y++)
                for (int x = brw; x < w - brw; x++) mask[(int64_t)y * w + x] = 1.0f;
        } else {
            or_relevance_mask(stds, w, h, brw, brh, or_noise_var(img, w, h), c, mask, &thr, &np);
        }
        or_apply_mask(out, n, mask);
        rc = or_normalize_map(out, w, h, brw, brh, mask);
        for (int y = brh; y < h - brh && !rc; y++)
            for (int x = brw; x < w - brw; x++) {
                int64_t i = (int64_t)y * w + x;
                if (mask[i] > 0.0f) out[i] = (out[i] == 0.0f) ? 1.0f : 1.0f / out[i];
            }
        if (normalize && !rc) rc = or_normalize_map(out, w, h, brw, brh, mask);
    } else {                                                    /* CLUtils.java:938-975 */
        float mn = FLT_MAX, mx = FLT_TRUE_MIN;                 /* Float.MIN_VALUE, CLUtils.java:940 */
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) {
                mn = fminf(mn, out[(int64_t)y * w + x]);
                mx = fmaxf(mx, out[(int64_t)y * w + x]);
            }
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) {
                int64_t i = (int64_t)y * w + x;
                out[i] = (out[i] - mn) / (mx - mn + EPS_F);
            }
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) out[(int64_t)y * w + x] = 1.0f - out[(int64_t)y * w + x];
        if (c > 0.0f) {
            mask = malloc(sizeof(float) * n);
            or_relevance_mask(stds, w, h, brw, brh, or_noise_var(img, w, h), c, mask, &thr, &np);
            or_apply_mask(out, n, mask);
        }
        if (normalize) rc = or_normalize_map(out, w, h, brw, brh, mask);
    }
done:
    free(means); free(stds); free(mask);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Hu template metric, kernelGetPatchHu2D.cl                                 */
/* ------------------------------------------------------------------------ */

/* float ellipse predicate of the kernel, e.g. kernelGetPatchHu2D.cl:38-40 */
static int hu_in_f(int i, int j, int brw, int brh)
{
    float dx = (float)(i - brw) / (float)brw;
    float dy = (float)(j - brh) / (float)brh;
    float a = dx * dx;
    float b = dy * dy;
    return a + b <= 1.0f;
}
/* double ellipse predicate, kernelGetPatchHu2D.cl:58-60 */
static int hu_in_d(int i, int j, int brw, int brh)
{
    double dx = (double)(i - brw) / brw;
    double dy = (double)(j - brh) / brh;
    double a = dx * dx;
    double b = dy * dy;
    return a + b <= 1.0;
}

/* hu1 of a patch held in p[bw*bh] over the cells where in[] is set,
 * kernelGetPatchHu2D.cl:52-101 and :190-239 */
static double hu1_of(const double *p, const unsigned char *in, int bw, int bh, double eps)
{
    double M00 = 0.0, M10 = 0.0, M01 = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (in[j * bw + i]) {
                double v = p[j * bw + i];
                M00 += v;
                M10 += (double)(i + 1) * v;
                M01 += (double)(j + 1) * v;
            }
    double cx = M10 / (M00 + eps);
    double cy = M01 / (M00 + eps);
    double mu20 = 0.0, mu02 = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (in[j * bw + i]) {
                double v = p[j * bw + i];
                double ddx = (double)i - cx;
                double ddy = (double)j - cy;
                mu20 += (ddx * ddx) * v;
                mu02 += (ddy * ddy) * v;
            }
    double m2 = M00 * M00;
    double nu20 = mu20 / (m2 + eps);
    double nu02 = mu02 / (m2 + eps);
    return nu20 + nu02;
}

double or_hu1_template(const float *tmpl, int bw, int bh, double eps)
{
    int brw = bw / 2, brh = bh / 2;
    double *p = (double *)calloc((size_t)bw * bh, sizeof(double));
    unsigned char *in = (unsigned char *)calloc((size_t)bw * bh, 1);
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            if (hu_in_f(i, j, brw, brh)) p[j * bw + i] = (double)tmpl[j * bw + i];  /* :36-44 */
            in[j * bw + i] = (unsigned char)hu_in_d(i, j, brw, brh);              /* :58-60 */
        }
    double r = hu1_of(p, in, bw, bh, eps);
    free(p);
    free(in);
    return r;
}

/* the window steps of kernelGetPatchHu2D.cl:127-182 on a patch whose
 * rectangle pixel (i, j) is get(i, j); fills p (FP64) and in (float ellipse) */
static void hu_window(const float *src, int pitch, int bw, int bh, int K, double eps, double *p, unsigned char *in)
{
    int brw = bw / 2, brh = bh / 2;
    double mean = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            /* (float)((i-gx)/bRW): integer division, so only the 4 corners fail */
            float dx = (float)((i - brw) / brw);
            float dy = (float)((j - brh) / brh);
            if (dx * dx + dy * dy <= 1.0f) {
                double v = (double)src[j * pitch + i];
                p[j * bw + i] = v;
                mean += v;
            }
        }
    mean /= (double)K;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) in[j * bw + i] = (unsigned char)hu_in_f(i, j, brw, brh);
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int k = 0; k < bw * bh; k++)
        if (in[k]) {
            p[k] -= mean;
            mn = p[k] < mn ? p[k] : mn;
            mx = p[k] > mx ? p[k] : mx;
        }
    for (int k = 0; k < bw * bh; k++)
        if (in[k]) p[k] = (p[k] - mn) / (mx - mn + eps);
}

void or_block_hu(const float *img, int w, int h, const float *tmpl, int bw, int bh, double eps, float *out)
{
    int brw = bw / 2, brh = bh / 2;
    int K = or_footprint(bw, bh, NULL, NULL);
    double ref = or_hu1_template(tmpl, bw, bh, eps);
#pragma omp parallel
    {
        double *p = (double *)calloc((size_t)bw * bh, sizeof(double));
        unsigned char *in = (unsigned char *)calloc((size_t)bw * bh, 1);
#pragma omp for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                if (x < brw || x >= w - brw || y < brh || y >= h - brh) {
                    out[(int64_t)y * w + x] = 0.0f;
                    continue;
                }
                hu_window(img + (int64_t)(y - brh) * w + (x - brw), w, bw, bh, K, eps, p, in);
                double c = hu1_of(p, in, bw, bh, eps);
                out[(int64_t)y * w + x] = (float)sqrt(ref * c);
            }
        free(p);
        free(in);
    }
}

void or_hu_prepare_template(const float *blk, int bw, int bh, double eps, float *out)
{
    int K = or_footprint(bw, bh, NULL, NULL);
    double *p = (double *)calloc((size_t)bw * bh, sizeof(double));
    unsigned char *in = (unsigned char *)calloc((size_t)bw * bh, 1);
    hu_window(blk, bw, bw, bh, K, eps, p, in);
    for (int k = 0; k < bw * bh; k++) out[k] = in[k] ? (float)p[k] : 0.0f;
    free(p);
    free(in);
}

/* ------------------------------------------------------------------------ */
/* Global Hu map, kernelGetHuMap.cl:15-172                                   */
/* ------------------------------------------------------------------------ */

/* getInvariant(p, q) of kernelGetHuMap.cl:146-172 for (2, 0) and (0, 2) on
 * the bw x bh rectangle around (x0, y0), patch = pixel - mean (:62-72), in
 * the kernel's float order.  Both calls sum mu_00 identically.  x_avg is the
 * float sum 1..bw over bw (:147-151); pow(d, 2) is d*d (d is an integer or
 * half-integer, so exact), pow(d, 0) is 1; pow(mu_00, 1 + 2/2) = mu_00^2
 * rounded to float; "+ EPSILON" promotes to double (the macro is an
 * unsuffixed literal), and the quotient is rounded back to float (:169-171). */
float or_hu_invariant(const float *img, int w, int bw, int bh, int x0, int y0, float mean, double eps)
{
    int brw = bw / 2, brh = bh / 2;
    float xavg = 0.0f, yavg = 0.0f;
    for (int i = 1; i <= bw; i++) xavg += (float)i;
    xavg /= (float)bw;
    for (int i = 1; i <= bh; i++) yavg += (float)i;
    yavg /= (float)bh;
    float mu20 = 0.0f, mu02 = 0.0f, mu00 = 0.0f;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            float v = img[(int64_t)(y0 - brh + j) * w + (x0 - brw + i)] - mean;
            float dx = (float)i + 1.0f - xavg;
            float dy = (float)j + 1.0f - yavg;
            mu20 += ((dx * dx) * 1.0f) * v;
            mu02 += (1.0f * (dy * dy)) * v;
            mu00 += v;
        }
    float m2 = mu00 * mu00;
    float inv20 = (float)((double)mu20 / ((double)m2 + eps));
    float inv02 = (float)((double)mu02 / ((double)m2 + eps));
    return inv20 + inv02;
}

void or_hu_invariants(const float *img, const float *means, int w, int h, int bw, int bh, double eps, float *hu)
{
    int brw = bw / 2, brh = bh / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t p = (int64_t)y * w + x;
            if (x < brw || x >= w - brw || y < brh || y >= h - brh)
                hu[p] = 0.0f;
            else
                hu[p] = or_hu_invariant(img, w, bw, bh, x, y, means[p], eps);
        }
}

/* getExpDecayWeight of kernelGetHuMap.cl:136-144 in FP64:
 * (100^(1 - |a-b| / (|a| + |b| + eps)) - 1) / 99. */
double or_hu_weight(float a, float b, double eps)
{
    double sim = 1.0 - fabs((double)a - (double)b) / (fabs((double)a) + fabs((double)b) + eps);
    return (pow(100.0, sim) - 1.0) / 99.0;
}

/* hu_map[i] = sum over interior j (raster order) of |hu_j - hu_i| w(s_i, s_j)
 * (kernelGetHuMap.cl:87-133), FP64 with Neumaier compensation, for the
 * interior reference pixels in rows (NULL: all interior pixels; the
 * kernel's uniqueStdCoords list, :27-39).  Other pixels of out are 0. */
void or_global_hu_sums(const float *hu, const float *stds, int w, int h, int bw, int bh, double eps,
                       const int32_t *rows, int64_t nrows, double *out, int nthreads)
{
    int brw = bw / 2, brh = bh / 2;
    int64_t n = (int64_t)w * h;
    for (int64_t p = 0; p < n; p++) out[p] = 0.0;
    int64_t R = rows ? nrows : n;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t r = 0; r < R; r++) {
        int64_t p = rows ? rows[r] : r;
        if (p < 0 || p >= n) continue;
        int x0 = (int)(p % w), y0 = (int)(p / w);
        if (x0 < brw || x0 >= w - brw || y0 < brh || y0 >= h - brh) continue;
        ksum s = {0.0, 0.0};
        for (int y1 = brh; y1 < h - brh; y1++)
            for (int x1 = brw; x1 < w - brw; x1++) {
                int64_t q = (int64_t)y1 * w + x1;
                kadd(&s, fabs((double)hu[q] - (double)hu[p]) * or_hu_weight(stds[p], stds[q], eps));
            }
        out[p] = s.s + s.c;
    }
}

/* ------------------------------------------------------------------------ */
/* Phase correlation map, kernelGetPhaseCorrelationMap.cl:13-179             */
/* ------------------------------------------------------------------------ */
/* The reference kernel does not compile (nUnique undefined at :27) and its
 * DFT phases use integer division (:66-67); the specification restated here
 * is textbook phase correlation on the same blocks:
 *   b(x, y)  = (pixel - mean) / (std + eps) on the bw x bh rectangle (:50-57);
 *   F        = 2D DFT of b, F(k) = sum b(x) exp(-2 pi i (kx x / bw + ky y / bh));
 *   U        = F / (|F| + eps), the block's unit spectrum (0 where F is 0);
 *   R(k)     = U_i(k) conj(U_j(k)), the normalised cross power spectrum
 *              (F_i conj F_j / |F_i conj F_j| wherever F_i, F_j >> eps),
 *              with i the lower and j the higher interior index of the pair,
 *              so each unordered pair has one displacement (|d| is the same
 *              for map[i] and map[j]);
 *   r(x, y)  = Re IDFT(R) = (1/n) Re sum_k R(k) exp(+2 pi i (...));
 *   peak     = first strict maximum of r over (y, x) in raster order, counted
 *              only if positive (:145-157, "zero ... no correlation");
 *   d        = peak coordinates wrapped to [-bw/2, bw/2] x [-bh/2, bh/2]
 *              (a circular shift s of block j against block i peaks at -s);
 *   term     = |d| * w(s_i, s_j), w the kernel's getExpDecayWeight
 *              (:167-178): 1 if s_i == s_j, else 1 - 2 |s_i - s_j| / (|s_i| + |s_j|),
 *              in float;
 *   map[i]   = sum over interior j of term (:75-164), FP64.
 * Angles are 2 pi q / n with the integer q = ((kx x mod bw) bh + (ky y mod bh)
 * bw) mod n, n = bw bh, so the twiddles are a table of n cosines and sines. */

#define OR_PI 3.14159265358979323846

static void phase_twiddles(int bw, int bh, double *c, double *s)
{
    int n = bw * bh;
    for (int q = 0; q < n; q++) {
        c[q] = cos(2.0 * OR_PI * (double)q / (double)n);
        s[q] = sin(2.0 * OR_PI * (double)q / (double)n);
    }
}

static int phase_q(int kx, int ky, int x, int y, int bw, int bh)
{
    return (((kx * x) % bw) * bh + ((ky * y) % bh) * bw) % (bw * bh);
}

/* unit spectrum U = F / (|F| + eps) of a bw x bh real block (row-major),
 * F its full 2D DFT summed in raster order */
static void phase_dft(const double *b, int bw, int bh, double eps, const double *c, const double *s, double *re,
                      double *im)
{
    for (int ky = 0; ky < bh; ky++)
        for (int kx = 0; kx < bw; kx++) {
            double a = 0.0, bi = 0.0;
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++) {
                    int q = phase_q(kx, ky, x, y, bw, bh);
                    a += b[y * bw + x] * c[q];
                    bi -= b[y * bw + x] * s[q];
                }
            double inv = 1.0 / (sqrt(a * a + bi * bi) + eps);
            re[ky * bw + kx] = a * inv;
            im[ky * bw + kx] = bi * inv;
        }
}

/* r of two unit spectra; returns the wrapped peak (dx, dy) and 1, or 0 when
 * no value of r is positive (then dx = dy = 0).  r_out (n) may be NULL. */
static int phase_peak(const double *ri, const double *ii, const double *rj, const double *ij, int bw, int bh,
                      const double *c, const double *s, double *Rr, double *Ri, double *r_out, int *dx, int *dy)
{
    int n = bw * bh;
    for (int k = 0; k < n; k++) {
        Rr[k] = ri[k] * rj[k] + ii[k] * ij[k];   /* U_i conj(U_j) */
        Ri[k] = ii[k] * rj[k] - ri[k] * ij[k];
    }
    double best = 0.0;
    int px = 0, py = 0, found = 0;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            double v = 0.0;
            for (int ky = 0; ky < bh; ky++)
                for (int kx = 0; kx < bw; kx++) {
                    int q = phase_q(kx, ky, x, y, bw, bh);
                    v += Rr[ky * bw + kx] * c[q] - Ri[ky * bw + kx] * s[q];
                }
            v /= (double)n;
            if (r_out) r_out[y * bw + x] = v;
            if (v > best) {
                best = v;
                px = x;
                py = y;
                found = 1;
            }
        }
    *dx = px <= bw / 2 ? px : px - bw;
    *dy = py <= bh / 2 ? py : py - bh;
    return found;
}

int or_phase_peak_blocks(const double *bi, const double *bj, int bw, int bh, double eps, double *r_out, int *dx,
                         int *dy)
{
    int n = bw * bh;
    double *buf = (double *)malloc(sizeof(double) * 8 * (size_t)n);
    double *c = buf, *s = buf + n, *ri = buf + 2 * n, *ii = buf + 3 * n, *rj = buf + 4 * n, *ij = buf + 5 * n;
    double *Rr = buf + 6 * n, *Ri = buf + 7 * n;
    phase_twiddles(bw, bh, c, s);
    phase_dft(bi, bw, bh, eps, c, s, ri, ii);
    phase_dft(bj, bw, bh, eps, c, s, rj, ij);
    int f = phase_peak(ri, ii, rj, ij, bw, bh, c, s, Rr, Ri, r_out, dx, dy);
    free(buf);
    return f;
}

double or_phase_weight(float a, float b)
{
    float wgt;
    if (a == b)
        wgt = 1.0f;
    else
        wgt = 1.0f - (2.0f * (fabsf(a - b) / (fabsf(a) + fabsf(b))));
    return (double)wgt;
}

void or_phase_map(const float *img, const float *means, const float *stds, int w, int h, int bw, int bh,
                  double eps, const int32_t *rows, int64_t nrows, double *out, int nthreads)
{
    int brw = bw / 2, brh = bh / 2, n = bw * bh;
    int iw = w - 2 * brw, ih = h - 2 * brh;
    int64_t npix = (int64_t)w * h;
    for (int64_t p = 0; p < npix; p++) out[p] = 0.0;
    if (iw <= 0 || ih <= 0) return;
    int64_t ni = (int64_t)iw * ih;
    double *c = (double *)malloc(sizeof(double) * n), *s = (double *)malloc(sizeof(double) * n);
    phase_twiddles(bw, bh, c, s);
    /* spectra of every interior block */
    double *sre = (double *)malloc(sizeof(double) * n * ni), *sim = (double *)malloc(sizeof(double) * n * ni);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        double *b = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (int64_t j = 0; j < ni; j++) {
            int x0 = brw + (int)(j % iw), y0 = brh + (int)(j / iw);
            int64_t p = (int64_t)y0 * w + x0;
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++)
                    b[y * bw + x] = ((double)img[(int64_t)(y0 - brh + y) * w + (x0 - brw + x)] - (double)means[p]) /
                                    ((double)stds[p] + eps);
            phase_dft(b, bw, bh, eps, c, s, sre + j * n, sim + j * n);
        }
        free(b);
    }
    int64_t R = rows ? nrows : npix;
#pragma omp parallel num_threads(nt)
    {
        double *Rr = (double *)malloc(sizeof(double) * n), *Ri = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 4)
        for (int64_t r = 0; r < R; r++) {
            int64_t p = rows ? rows[r] : r;
            if (p < 0 || p >= npix) continue;
            int x0 = (int)(p % w), y0 = (int)(p / w);
            if (x0 < brw || x0 >= w - brw || y0 < brh || y0 >= h - brh) continue;
            int64_t i = (int64_t)(y0 - brh) * iw + (x0 - brw);
            double acc = 0.0;
            for (int64_t j = 0; j < ni; j++) {
                int dx, dy;
                const int64_t lo = j < i ? j : i, hi = j < i ? i : j;  /* lower interior index first */
                phase_peak(sre + lo * n, sim + lo * n, sre + hi * n, sim + hi * n, bw, bh, c, s, Rr, Ri, NULL, &dx,
                           &dy);
                int64_t q = (int64_t)(brh + j / iw) * w + (brw + j % iw);
                acc += sqrt((double)(dx * dx + dy * dy)) * or_phase_weight(stds[p], stds[q]);
            }
            out[p] = acc;
        }
        free(Rr);
        free(Ri);
    }
    free(sre); free(sim); free(c); free(s);
}

/* ------------------------------------------------------------------------ */
/* Variance stabilisation: GAT and its simplex fit                          */
/* ------------------------------------------------------------------------ */

/* VarianceStabilisingTransform2D_.getGAT, VarianceStabilisingTransform2D_.java:136-151:
 * ref = 3/8 g^2 + s^2 - g o; v <- (g v + ref < 0) ? 0 : (2/g) sqrt(g v + ref),
 * in double, rounded to float, in place. */
void or_gat(float *px, int64_t n, double gain, double sigma, double offset)
{
    double ref = (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset;
    for (int64_t i = 0; i < n; i++) {
        double v = (double)px[i];
        if ((gain * v + ref) < 0.0)
            v = 0.0;
        else
            v = (2.0 / gain) * sqrt(gain * v + ref);
        px[i] = (float)v;
    }
}

/* The per-pixel transform of GATMinimizer2D.applyGAT (GATMinimizer2D.java:140-155),
 * whose special case is v <= -ref/gain. */
static float gat_fit_px(float p, double gain, double ref)
{
    double v = (double)p;
    if (v <= -ref / gain)
        v = 0.0;
    else
        v = (2.0 / gain) * sqrt(gain * v + ref);
    return (float)v;
}

/* Sum of a[0..63] as a 64-lane xor butterfly: for s = 32, 16, ..., 1 every
 * lane l takes a[l] + a[l ^ s]; the result is lane 0's value. */
static double lane_tree64(double *a)
{
    double t[64];
    for (int s = 32; s >= 1; s >>= 1) {
        for (int l = 0; l < 64; l++) t[l] = a[l] + a[l ^ s];
        memcpy(a, t, sizeof t);
    }
    return a[0];
}

/* GATMinimizer2D.userFunction (GATMinimizer2D.java:64-110): the image is cut
 * into nbx x nby blocks of (w/6) x (h/6) pixels (nbx = w / (w/6)); each
 * block's mean and variance of the transformed pixels (sum v and sum of the
 * float squares v*v, in double; var = sq/n - mean^2, :115-137) gives
 * error = sum over blocks of (var - 1)^2 / nBlocks.
 * Summation order (the reference adds all pixels of a block in one sequence
 * and the blocks in raster order; this order is the one the GPU reduction
 * reproduces exactly): a block row's pixels go to 64 lanes (pixel i to lane
 * i mod 64, each lane adding its pixels in order) and the lanes are combined
 * by lane_tree64; a block's row sums likewise (row r of the block to lane
 * r mod 64); the block terms likewise (block b, raster index, to lane
 * b mod 64).
 * Returns NaN for images narrower or shorter than 6 pixels (the reference
 * divides by zero there). */
double or_gat_objective(const float *px, int w, int h, double gain, double sigma, double offset)
{
    int bw = w / 6, bh = h / 6;
    if (bw <= 0 || bh <= 0) return NAN;
    int nbx = w / bw, nby = h / bh, nb = nbx * nby;
    double ref = (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset;
    double be[64], bm[64], bs[64], lm[64], ls[64];
    for (int l = 0; l < 64; l++) be[l] = 0.0;
    for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
            for (int l = 0; l < 64; l++) bm[l] = bs[l] = 0.0;
            for (int r = 0; r < bh; r++) {
                const float *row = px + (int64_t)(by * bh + r) * w + (int64_t)bx * bw;
                for (int l = 0; l < 64; l++) lm[l] = ls[l] = 0.0;
                for (int i = 0; i < bw; i++) {
                    float v = gat_fit_px(row[i], gain, ref);
                    lm[i % 64] += v;
                    ls[i % 64] += v * v;  /* float product, as the reference's sq_sum += v * v */
                }
                bm[r % 64] += lane_tree64(lm);
                bs[r % 64] += lane_tree64(ls);
            }
            double mean = lane_tree64(bm), sq = lane_tree64(bs);
            int n = bw * bh;
            mean = mean / n;
            double var = sq / n - mean * mean;
            double delta = var - 1.0;
            int b = by * nbx + bx;
            be[b % 64] += (delta * delta) / (double)nb;
        }
    return lane_tree64(be);
}

/* Nelder-Mead simplex in 3 dimensions (the build's own minimiser in place of
 * ImageJ's Minimizer used by GATMinimizer2D.run, GATMinimizer2D.java:33-61;
 * ImageJ's restarts and random initialisation are not reproduced).
 * Lagarias et al. (1998) coefficients: reflection 1, expansion 2,
 * contraction 1/2, shrink 1/2.  Initial simplex x0, x0 + d_i e_i.  Each
 * iteration orders the vertices by f (stable: ties keep their order), then
 * tries the reflection, expansion, outside or inside contraction, or
 * shrinks.  Stops after max_iter iterations or when every vertex has the
 * same f (the reference's setMaxError(0)).  x receives the best vertex;
 * returns its f.  f(params, ctx) is the objective. */
typedef double (*or_fn3)(const double *x, void *ctx);

/* A NaN objective marks invalid parameters: it ranks worst (+infinity), as
 * ImageJ's Minimizer treats NaN results (GATMinimizer3D returns NaN for
 * gain <= 0 on purpose). */
static double nm_eval(or_fn3 f, const double *x, void *ctx)
{
    double r = f(x, ctx);
    return r != r ? INFINITY : r;
}

double or_nelder_mead3(or_fn3 f, void *ctx, const double *x0, const double *d, int max_iter, double *x)
{
    double v[4][3], fv[4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) v[i][k] = x0[k];
        if (i > 0) v[i][i - 1] += d[i - 1];
        fv[i] = nm_eval(f, v[i], ctx);
    }
    for (int it = 0; it < max_iter; it++) {
        /* stable insertion sort by f */
        for (int i = 1; i < 4; i++)
            for (int j = i; j > 0 && fv[j] < fv[j - 1]; j--) {
                double tf = fv[j]; fv[j] = fv[j - 1]; fv[j - 1] = tf;
                for (int k = 0; k < 3; k++) { double t = v[j][k]; v[j][k] = v[j - 1][k]; v[j - 1][k] = t; }
            }
        if (fv[3] == fv[0]) break;
        double c[3], xr[3], xe[3], xc[3];
        for (int k = 0; k < 3; k++) c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
        for (int k = 0; k < 3; k++) xr[k] = c[k] + (c[k] - v[3][k]);
        double fr = nm_eval(f, xr, ctx);
        if (fr < fv[0]) {
            for (int k = 0; k < 3; k++) xe[k] = c[k] + 2.0 * (c[k] - v[3][k]);
            double fe = nm_eval(f, xe, ctx);
            if (fe < fr) { for (int k = 0; k < 3; k++) v[3][k] = xe[k]; fv[3] = fe; }
            else { for (int k = 0; k < 3; k++) v[3][k] = xr[k]; fv[3] = fr; }
            continue;
        }
        if (fr < fv[2]) { for (int k = 0; k < 3; k++) v[3][k] = xr[k]; fv[3] = fr; continue; }
        int outside = fr < fv[3];
        for (int k = 0; k < 3; k++) xc[k] = outside ? c[k] + 0.5 * (xr[k] - c[k]) : c[k] + 0.5 * (v[3][k] - c[k]);
        double fc = nm_eval(f, xc, ctx);
        if (outside ? fc <= fr : fc < fv[3]) { for (int k = 0; k < 3; k++) v[3][k] = xc[k]; fv[3] = fc; continue; }
        for (int i = 1; i < 4; i++) {  /* shrink towards the best vertex */
            for (int k = 0; k < 3; k++) v[i][k] = v[0][k] + 0.5 * (v[i][k] - v[0][k]);
            fv[i] = nm_eval(f, v[i], ctx);
        }
    }
    int b = 0;
    for (int i = 1; i < 4; i++) if (fv[i] < fv[b]) b = i;
    for (int k = 0; k < 3; k++) x[k] = v[b][k];
    return fv[b];
}

typedef struct { const float *px; int w, h; } gat_ctx;
static double gat_obj3(const double *x, void *c)
{
    const gat_ctx *g = (const gat_ctx *)c;
    return or_gat_objective(g->px, g->w, g->h, x[0], x[1], x[2]);
}

/* GATMinimizer2D.run with the simplex above: initial (gain, sigma, offset)
 * and steps (1, 10, 100) (GATMinimizer2D.java:36-45).  params[3] receives
 * the fit; returns the objective there. */
double or_gat_minimize(const float *px, int w, int h, double gain0, double sigma0, double offset0, int max_iter,
                       double *params)
{
    gat_ctx c = {px, w, h};
    double x0[3] = {gain0, sigma0, offset0}, d[3] = {1.0, 10.0, 100.0};
    return or_nelder_mead3(gat_obj3, &c, x0, d, max_iter, params);
}

/* ------------------------------------------------------------------------ */
/* Quadtree GAT estimate (QuadTree_.java:135-427, Utils.java:686-703)        */
/* ------------------------------------------------------------------------ */

#define OR_JEPS 1e-7f /* Utils.EPSILON (Utils.java:20) */

/* java.lang.Math.max / min for floats (NaN wins; -0 < +0). */
static float jmaxf(float a, float b)
{
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && signbit(a)) return b;
    return a >= b ? a : b;
}
static float jminf(float a, float b)
{
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && signbit(b)) return b;
    return a <= b ? a : b;
}

/* Regularised incomplete beta I_x(a, b): x^a (1-x)^b / (a B(a,b)) / g with
 * g = 1 + d1/(1 + d2/(1 + ...)) the continued fraction of Abramowitz &
 * Stegun 26.5.8, evaluated by Lentz's method (or_beta_cf returns 1/g); the
 * symmetry I_x(a,b) = 1 - I_{1-x}(b,a) where the fraction converges slowly
 * (x beyond the mean (a+1)/(a+b+2)). */
static double or_beta_cf(double a, double b, double x)
{
    const double tiny = 1e-300;
    double f = 1.0, C = 1.0, D = 0.0;  /* f -> g */
    for (int m = 0; m < 2000000; m++) {
        /* terms d_{2m+1} and d_{2m+2} (d_1 = -(a+b)x/(a+1)) */
        for (int half = 0; half < 2; half++) {
            double dm;
            if (half == 0)
                dm = -(a + m) * (a + b + m) * x / ((a + 2.0 * m) * (a + 2.0 * m + 1.0));
            else
                dm = (m + 1.0) * (b - m - 1.0) * x / ((a + 2.0 * m + 1.0) * (a + 2.0 * m + 2.0));
            D = 1.0 + dm * D;
            if (fabs(D) < tiny) D = tiny;
            C = 1.0 + dm / C;
            if (fabs(C) < tiny) C = tiny;
            D = 1.0 / D;
            double delta = C * D;
            f *= delta;
            if (half == 1 && fabs(delta - 1.0) < 1e-16) return 1.0 / f;
        }
    }
    return 1.0 / f;
}

double or_beta_reg(double a, double b, double x)
{
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    double lfront = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x);
    if (x < (a + 1.0) / (a + b + 2.0)) return exp(lfront) * or_beta_cf(a, b, x) / a;
    return 1.0 - exp(lfront) * or_beta_cf(b, a, 1.0 - x) / b;
}

/* Inverse CDF of the F distribution with (d1, d2) degrees of freedom:
 * commons-math3 3.6.1 FDistribution.inverseCumulativeProbability (the
 * reference's dependency; its Brent solver stops at 1e-9 absolute): here
 * bisection on F(x) = I_{d1 x/(d1 x + d2)}(d1/2, d2/2) to adjacent doubles. */
double or_f_quantile(double d1, double d2, double p)
{
    if (!(p > 0.0)) return 0.0;
    if (!(p < 1.0)) return INFINITY;
    double lo = 0.0, hi = 1.0;
    while (or_beta_reg(0.5 * d1, 0.5 * d2, d1 * hi / (d1 * hi + d2)) < p) {
        lo = hi;
        hi *= 2.0;
    }
    for (int it = 0; it < 2100; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (or_beta_reg(0.5 * d1, 0.5 * d2, d1 * mid / (d1 * mid + d2)) < p)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* commons-math3 3.6.1 SimpleRegression with intercept (the reference's
 * dependency, QuadTree_.java:395-407): the updating sums of addData, then
 * getSlope = sumXY/sumXX (NaN below 2 points or for |sumXX| < 10 Double.MIN_VALUE)
 * and getIntercept = (sumY - slope sumX)/n.  Returns the number of points
 * (regress() throws below 3). */
int or_simple_regression(const double *x, const double *y, int64_t n, double *slope, double *intercept)
{
    double sx = 0.0, sy = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0, xbar = 0.0, ybar = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (i == 0) {
            xbar = x[i];
            ybar = y[i];
        } else {
            double fact1 = 1.0 + (double)i, fact2 = (double)i / (1.0 + (double)i);
            double dx = x[i] - xbar, dy = y[i] - ybar;
            sxx += dx * dx * fact2;
            syy += dy * dy * fact2;
            sxy += dx * dy * fact2;
            xbar += dx / fact1;
            ybar += dy / fact1;
        }
        sx += x[i];
        sy += y[i];
    }
    (void)syy;
    double s = NAN;
    if (n >= 2 && !(fabs(sxx) < 10.0 * 4.9406564584124654e-324)) s = sxy / sxx;
    *slope = s;
    *intercept = (sy - s * sx) / (double)n;
    return (int)(n < INT32_MAX ? n : INT32_MAX);
}

/* QuadTree_.computeCriterion (QuadTree_.java:170-183): the node's variance
 * (Utils.getMeanAndVarBlock2D, Utils.java:854-898: float sums in raster
 * order) and the variance of the pseudo-residuals (5 z - neighbours) /
 * sqrt(30) (calculateNoiseVariance, :185-232; neighbours outside the image
 * omitted, neighbours outside the node included). */
void or_qt_criterion(const float *img, int w, int h, int x0, int y0, int nw, int nh, float *block_var,
                     float *noise_var)
{
    int n = nw * nh;
    float mean = 0.0f, sq = 0.0f;
    for (int y = y0; y < y0 + nh; y++)
        for (int x = x0; x < x0 + nw; x++) {
            float v = img[(int64_t)y * w + x];
            mean += v;
            sq += v * v;
        }
    mean = mean / (float)n;
    *block_var = sq / (float)n - mean * mean;

    const float l = 5.0f;
    const float scale = (float)(1.0 / sqrt((double)(l * l + l)));
    float sum = 0.0f;
    for (int pass = 0; pass < 2; pass++) {
        float mr = sum / (float)n, ss = 0.0f;
        for (int y = y0; y < y0 + nh; y++)
            for (int x = x0; x < x0 + nw; x++) {
                const float *p = img + (int64_t)y * w + x;
                float lap = l * p[0];
                if (x > 0) lap -= p[-1];
                if (x < w - 1) lap -= p[1];
                if (y > 0) lap -= p[-w];
                if (y < h - 1) lap -= p[w];
                float r = lap * scale;
                if (pass == 0) {
                    sum += r;
                } else {
                    double d = (double)r - (double)mr;
                    ss += (float)(d * d);  /* Math.pow(d, 2.0) */
                }
            }
        if (pass == 1) *noise_var = ss / ((float)n - 1.0f);
    }
}

/* QuadTree_.calculateRobustMean (:264-312): a Leclerc-weighted mean from the
 * one-pass estimate (whose variance uses the running sum in place of the
 * mean, as written), iterated while the mean moves by more than 1e-5. */
float or_qt_robust_mean(const float *img, int w, int x0, int y0, int nw, int nh, int max_iter)
{
    float mean = 0.0f, var = 0.0f;
    int n = 0;
    for (int y = y0; y < y0 + nh; y++)
        for (int x = x0; x < x0 + nw; x++) {
            float v = img[(int64_t)y * w + x];
            mean += v;
            var += (v - mean) * (v - mean);
            n++;
        }
    mean = mean / (float)n;
    var = var / (float)(n - 1);
    var = jmaxf(var, OR_JEPS);
    float prev;
    int it = 0;
    do {
        prev = mean;
        float ws = 0.0f, wv = 0.0f;
        for (int y = y0; y < y0 + nh; y++)
            for (int x = x0; x < x0 + nw; x++) {
                float v = img[(int64_t)y * w + x];
                float wt = (float)exp(-((double)(v - mean) * (double)(v - mean)) / (2.0 * (double)var));
                ws += wt;
                wv += wt * v;
            }
        mean = wv / jmaxf(ws, OR_JEPS);
        it++;
    } while ((double)fabsf(mean - prev) > 1e-5 && it < max_iter);
    return mean;
}

static int cmp_float_asc(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* QuadTree_.calculateLTSVariance (:326-356): the squared residuals about the
 * robust mean sorted ascending, the floor(frac n) smallest summed in float. */
float or_qt_lts_variance(const float *img, int w, int x0, int y0, int nw, int nh, float mean, float frac)
{
    int n = nw * nh;
    float *r = (float *)malloc(sizeof(float) * (size_t)n);
    int k = 0;
    for (int y = y0; y < y0 + nh; y++)
        for (int x = x0; x < x0 + nw; x++) {
            float z = img[(int64_t)y * w + x];
            r[k++] = (z - mean) * (z - mean);
        }
    qsort(r, (size_t)n, sizeof(float), cmp_float_asc);
    int hh = (int)floorf(frac * (float)n);
    float s = 0.0f;
    for (int i = 0; i < hh; i++) s += r[i];
    free(r);
    return s / ((float)hh + OR_JEPS);
}

typedef struct {
    const float *img;
    int w, h, min_size;
    float alpha;
    int n, cap, bad;
    int32_t *rects;
    int qn[64];
    float qthr[64];
} qt_build;

static float qt_threshold(qt_build *b, int region)
{
    for (int i = 0; i < b->n && i < 64; i++)
        if (b->qn[i] == region) return b->qthr[i];
    double d = (double)(region - 1);
    float t = (float)or_f_quantile(d, d, (double)b->alpha);
    if (b->n < 64) {
        b->qn[b->n] = region;
        b->qthr[b->n] = t;
        b->n++;
    }
    return t;
}

static int qt_leaves;

static void qt_visit(qt_build *b, int x, int y, int nw, int nh)
{
    if (b->bad) return;
    float bv, nv;
    or_qt_criterion(b->img, b->w, b->h, x, y, nw, nh, &bv, &nv);
    if (bv != bv) {  /* getMeanAndVarBlock2D: IllegalStateException */
        b->bad = -5;
        return;
    }
    /* performFisherTest (:236-249) */
    float ratio = jminf(bv, nv) / jmaxf(bv, nv);
    int divide = ratio < qt_threshold(b, nw * nh);
    if (nw <= b->min_size || nh <= b->min_size || !divide) {
        if (qt_leaves < b->cap) {
            int32_t *r = b->rects + 4 * qt_leaves;
            r[0] = x, r[1] = y, r[2] = nw, r[3] = nh;
        }
        qt_leaves++;
        return;
    }
    int hw = nw / 2, hh = nh / 2;
    qt_visit(b, x, y, hw, hh);
    qt_visit(b, x + hw, y, hw, hh);
    qt_visit(b, x, y + hh, hw, hh);
    qt_visit(b, x + hw, y + hh, hw, hh);
}

/* Utils.getInputImage2D's "Quad/Octree" branch (Utils.java:686-703):
 * QuadTree_(w, h, min_size, alpha).buildTree, calculateRobustMeans(max_iter),
 * calculateLTSVariances(frac), collectMeanVariancePairs (leaves in pre-order:
 * top-left, top-right, bottom-left, bottom-right) and performLinearRegression.
 * rects (x, y, w, h), means and vars receive up to cap leaves; *nleaves the
 * count.  Returns 0, -5 for a non-finite pixel or a NaN node variance
 * (Utils.java:881-895), -1 for a one-pixel root (FDistribution(0, 0)) or fewer
 * than 3 leaves (SimpleRegression.regress), -2 if cap is too small. */
int or_quadtree_gat(const float *img, int w, int h, int min_size, float alpha, int max_iter, float frac, int cap,
                    int32_t *rects, float *means, float *vars, int *nleaves, double *g0, double *edc)
{
    for (int64_t i = 0; i < (int64_t)w * h; i++)
        if (!isfinite(img[i])) return -5;
    if ((int64_t)w * h <= 1) return -1;
    qt_build b;
    memset(&b, 0, sizeof b);
    b.img = img, b.w = w, b.h = h, b.min_size = min_size, b.alpha = alpha, b.cap = cap, b.rects = rects;
    qt_leaves = 0;
    qt_visit(&b, 0, 0, w, h);
    *nleaves = qt_leaves;
    if (b.bad) return b.bad;
    if (qt_leaves > cap) return -2;
    for (int i = 0; i < qt_leaves; i++) {
        const int32_t *r = rects + 4 * i;
        means[i] = or_qt_robust_mean(img, w, r[0], r[1], r[2], r[3], max_iter);
        vars[i] = or_qt_lts_variance(img, w, r[0], r[1], r[2], r[3], means[i], frac);
    }
    if (qt_leaves < 3) return -1;
    double *x = (double *)malloc(sizeof(double) * 2 * (size_t)qt_leaves), *y = x + qt_leaves;
    for (int i = 0; i < qt_leaves; i++) {
        x[i] = (double)means[i];
        y[i] = (double)vars[i];
    }
    or_simple_regression(x, y, qt_leaves, g0, edc);
    free(x);
    return 0;
}

/* QuadTree_.applyGATtree (:409-425) in place: ref = 3/8 g0^2 + eDC. */
void or_gat_tree(float *px, int64_t n, double g0, double edc)
{
    double ref = (3.0 / 8.0) * g0 * g0 + edc;
    for (int64_t i = 0; i < n; i++) {
        double v = (double)px[i];
        if ((g0 * v + ref) < 0.0)
            v = 0.0;
        else
            v = (2.0 / g0) * sqrt(g0 * v + ref);
        px[i] = (float)v;
    }
}

/* ------------------------------------------------------------------------ */
/* Relevance-mask tools (RelevanceMask2D_, OptimiseRelevanceMask_,           */
/* CalculateLocalMeans_)                                                     */
/* ------------------------------------------------------------------------ */

/* The tools' shared input step (RelevanceMask2D_.java:133-164,
 * OptimiseRelevanceMask_.java:124-154, CalculateLocalMeans_.java:163-190):
 * GATMinimizer2D from (0, 100, 0), getGAT with the fitted doubles, min-max
 * normalisation in double, rounded to float.  In place; params receives the
 * fit. */
void or_tool_prep(float *img, int w, int h, int max_iter, double *params)
{
    double p[3];
    or_gat_minimize(img, w, h, 0.0, 100.0, 0.0, max_iter, p);
    int64_t n = (int64_t)w * h;
    or_gat(img, n, p[0], p[1], p[2]);
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int64_t i = 0; i < n; i++) {
        double v = (double)img[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    for (int64_t i = 0; i < n; i++) img[i] = (float)(((double)img[i] - mn) / (mx - mn + (double)1e-7f));
    if (params) memcpy(params, p, sizeof p);
}

/* RelevanceMask2D_'s noise variance (:203-233): 8x8 tiles up to CIF size,
 * 16x16 above; per tile the double sums of v and of the float squares v*v
 * (getMeanAndVarBlock, :401-419), the variance rounded to float; the lowest
 * 3% averaged in float and scaled as Utils.getMeanNoiseVar2D does. */
float or_noise_var_double(const float *img, int w, int h)
{
    int tb = ((int64_t)w * h <= 352 * 288) ? 8 : 16;
    int nbx = w / tb, nby = h / tb, nb = nbx * nby;
    if (nb <= 0) return NAN;
    float *vars = (float *)malloc(sizeof(float) * (size_t)nb);
    for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
            double mean = 0.0, sq = 0.0;
            for (int y = by * tb; y < (by + 1) * tb; y++)
                for (int x = bx * tb; x < (bx + 1) * tb; x++) {
                    float v = img[(int64_t)y * w + x];
                    mean += v;
                    sq += v * v;
                }
            mean = mean / (tb * tb);
            double var = sq / (tb * tb) - mean * mean;
            vars[by * nbx + bx] = (float)var;
        }
    qsort(vars, (size_t)nb, sizeof(float), cmp_float_asc);
    int nv_n = (int)(0.03f * (float)nb + 1.0f);
    float nv = 0.0f;
    for (int i = 0; i < nv_n; i++) nv += vars[i];
    nv = fabsf(nv / (float)nv_n);
    nv = (1.0f + 0.001f * (nv - 40.0f)) * nv;
    free(vars);
    return nv;
}

/* RelevanceMask2D_.run after the input step (:165-250), for odd square
 * blocks: local standard deviations as kernelGetLocalMeans.cl (the float
 * arithmetic of kernelGetLocalStatistics2D.cl), the noise variance above,
 * threshold = noiseVar * c (c itself when noiseVar is 0), and mask 0 where
 * std*std < threshold or std*std == 0, 1 elsewhere inside the borders.
 * Returns -1 for blocks that are not odd and square (the reference swaps the
 * radii and overruns its patch array otherwise). */
int or_relevance_mask_tool(const float *img, int w, int h, int bw, int bh, float c, float *mask, float *noise_var,
                           float *threshold)
{
    if (bw != bh || bw % 2 == 0 || bw < 3) return -1;
    int64_t n = (int64_t)w * h;
    float *means = (float *)malloc(sizeof(float) * (size_t)n), *stds = (float *)malloc(sizeof(float) * (size_t)n);
    or_local_stats_f32(img, w, h, bw, bh, means, stds);
    float nv = or_noise_var_double(img, w, h);
    float thr = (nv == 0.0f) ? c : nv * c;
    int br = bw / 2;
    for (int64_t i = 0; i < n; i++) mask[i] = 0.0f;
    for (int y = br; y < h - br; y++)
        for (int x = br; x < w - br; x++) {
            float s = stds[(int64_t)y * w + x];
            float v = s * s;
            mask[(int64_t)y * w + x] = (v < thr || v == 0.0f) ? 0.0f : 1.0f;
        }
    *noise_var = nv;
    *threshold = thr;
    free(means);
    free(stds);
    return 0;
}

/* kernelGetRelevanceMask2D.cl: the local variance (no square root) of each
 * inner pixel, the mean accumulated while gathering the ellipse. */
void or_local_var_f32(const float *img, int w, int h, int bw, int bh, float *var)
{
    int brw = bw / 2, brh = bh / 2;
    int K = or_footprint(bw, bh, NULL, NULL);
    int *dx = malloc(sizeof(int) * K), *dy = malloc(sizeof(int) * K);
    or_footprint(bw, bh, dx, dy);
    for (int gy = 0; gy < h; gy++)
        for (int gx = 0; gx < w; gx++) {
            int64_t o = (int64_t)gy * w + gx;
            if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) { var[o] = 0.0f; continue; }
            float m = 0.0f;
            for (int i = 0; i < K; i++) m += img[(int64_t)(gy + dy[i]) * w + gx + dx[i]];
            m /= (float)K;
            float v = 0.0f;
            for (int i = 0; i < K; i++) {
                float p = img[(int64_t)(gy + dy[i]) * w + gx + dx[i]];
                v += (p - m) * (p - m);
            }
            var[o] = v / (float)(K - 1);
        }
    free(dx); free(dy);
}

/* OptimiseRelevanceMask_.run after the input step (:155-287), literally:
 * the "standard deviations" are kernelGetRelevanceMask2D's variances, so the
 * mask compares their squares; the filter constant grows from 0 by step (a
 * float sum) and each value's mask is built (noise variance as
 * Utils.getMeanNoiseVar2D, threshold noiseVar * c, or c when noiseVar is 0);
 * *all_ones = the last constant whose mask is all ones inside the borders,
 * *all_zeros = the first whose mask is all zeros.  Returns -1 for blocks that
 * are not odd and square or step <= 0, -2 if no mask is all zeros within
 * max_steps constants (the reference loops for ever). */
int or_optimise_relevance_mask(const float *img, int w, int h, int bw, int bh, float step, int max_steps,
                               float *all_ones, float *all_zeros)
{
    if (bw != bh || bw % 2 == 0 || bw < 3 || !(step > 0.0f)) return -1;
    int64_t n = (int64_t)w * h;
    float *var = (float *)malloc(sizeof(float) * (size_t)n);
    or_local_var_f32(img, w, h, bw, bh, var);
    float nv = or_noise_var(img, w, h);
    int br = bw / 2;
    float c = 0.0f, ones = 0.0f;
    int rc = -2;
    for (int it = 0; it < max_steps; it++) {
        float thr = (nv == 0.0f) ? c : nv * c;
        int every_one = 1, every_zero = 1;
        for (int y = br; y < h - br; y++)
            for (int x = br; x < w - br; x++) {
                float s = var[(int64_t)y * w + x];
                float v = s * s;
                float m = (v < thr || v == 0.0f) ? 0.0f : 1.0f;
                if (m != 1.0f) every_one = 0;
                if (m != 0.0f) every_zero = 0;
            }
        if (every_one) ones = c;
        if (every_zero) {
            *all_zeros = c;
            rc = 0;
            break;
        }
        c += step;
    }
    *all_ones = ones;
    free(var);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Volumes: GlobalRepetition3D_ with its four metrics, as intended            */
/* ------------------------------------------------------------------------ */

/* Ellipsoid footprint (GlobalRepetition3D_.java:101-112, the kernels' z, y,
 * x loops and float predicate). */
int or_footprint3d(int bw, int bh, int bd, int *dx, int *dy, int *dz)
{
    int brw = bw / 2, brh = bh / 2, brd = bd / 2, k = 0;
    for (int z = 0; z < bd; z++)
        for (int y = 0; y < bh; y++)
            for (int x = 0; x < bw; x++) {
                float fx = (float)(x - brw), fy = (float)(y - brh), fz = (float)(z - brd);
                if (((fx * fx) / (float)(brw * brw)) + ((fy * fy) / (float)(brh * brh)) +
                        ((fz * fz) / (float)(brd * brd)) <= 1.0f) {
                    if (dx) { dx[k] = x - brw; dy[k] = y - brh; dz[k] = z - brd; }
                    k++;
                }
            }
    return k;
}

static int64_t *vol_offsets(int w, int h, int bw, int bh, int bd, int *K)
{
    int k = or_footprint3d(bw, bh, bd, NULL, NULL, NULL);
    int *dx = malloc(sizeof(int) * 3 * k), *dy = dx + k, *dz = dy + k;
    or_footprint3d(bw, bh, bd, dx, dy, dz);
    int64_t *off = malloc(sizeof(int64_t) * k);
    for (int i = 0; i < k; i++) off[i] = ((int64_t)dz[i] * h + dy[i]) * w + dx[i];
    free(dx);
    *K = k;
    return off;
}

static int vol_inner(int64_t o, int w, int h, int d, int brw, int brh, int brd)
{
    int x = (int)(o % w), y = (int)((o / w) % h), z = (int)(o / ((int64_t)w * h));
    return x >= brw && x < w - brw && y >= brh && y < h - brh && z >= brd && z < d - brd;
}

/* kernelGetLocalStatistics3D.cl: float mean and unbiased std over the
 * ellipsoid in footprint order; borders 0. */
void or_local_stats3d_f32(const float *img, int w, int h, int d, int bw, int bh, int bd, float *means, float *stds)
{
    int K;
    int64_t *off = vol_offsets(w, h, bw, bh, bd, &K);
    int64_t n = (int64_t)w * h * d;
#pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < n; o++) {
        if (!vol_inner(o, w, h, d, bw / 2, bh / 2, bd / 2)) { means[o] = 0.0f; stds[o] = 0.0f; continue; }
        float m = 0.0f;
        for (int i = 0; i < K; i++) m += img[o + off[i]];
        m /= (float)K;
        float v = 0.0f;
        for (int i = 0; i < K; i++) v += (img[o + off[i]] - m) * (img[o + off[i]] - m);
        means[o] = m;
        stds[o] = sqrtf(v / (float)(K - 1));
    }
    free(off);
}

/* Utils.getMeanNoiseVar3D(..., localStds) (Utils.java:1463-1499): float sum
 * of std^2 over the inner voxels in raster order over a float count. */
float or_vol_noise_var(const float *stds, int w, int h, int d, int brw, int brh, int brd)
{
    float s = 0.0f, cnt = 0.0f;
    for (int z = brd; z < d - brd; z++)
        for (int y = brh; y < h - brh; y++)
            for (int x = brw; x < w - brw; x++) {
                float v = stds[((int64_t)z * h + y) * w + x];
                s += v * v;
                cnt += 1.0f;
            }
    return s / cnt;
}

/* GlobalRepetition3D_.run -> CLUtils.calculateGlobalRepetitionMap3D (:2793-2815)
 * with the four metrics as intended (the statistics with their declared
 * argument order, Utils.getRelevanceMask3D, the reference block over the
 * full ellipsoid), in float (fp64 = 0: the kernels' float operations and
 * sequential sums) or in double (fp64 = 1: double block values, products,
 * weights and sums; validity and flatness still from the float statistics).
 * out (double) receives the map; S, W, mask, *noise_var, *n_pixels (each
 * may be NULL) the intermediates of the float form.  Per pair
 * (kernelGetGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}3D.cl): a reference voxel
 * below the threshold or with std 0 scores 0 (line 37 of the first three,
 * AbsDiffStds3D.cl:36); test voxels below the threshold are skipped, and
 * AbsDiffStds3D.cl:53 also skips flat test voxels; the weight is the 2D one;
 * NRMSE adds sqrt(mse) * weight (Nrmse3D.cl:111; the 2D kernel adds
 * sqrt(mse)); SSIM is the 2D form (Ssim3D.cl:107-124; its no-op line 76 on an
 * undeclared pearson_map is dropped); AbsDiffStds adds fmax(0, |d|) * weight
 * with the weight's exp in double (AbsDiffStds3D.cl:55-60).  Finalize
 * (CLUtils.java:2309-2330, 2435-2467, 2569-2618, 2720-2769): S / (W nPixels
 * + eps) on the inner voxels, NRMSE and AbsDiffStds inverted where the mask
 * is set, the mask applied when c > 0, normalised over the mask. */
int or_global_repetition3d(int metric, const float *img, int w, int h, int d, int bw, int bh, int bd, float c,
                           int normalize, int fp64, double *out, float *S_out, float *W_out, float *mask_out,
                           float *noise_var, float *n_pixels)
{
    if (metric < OR_PEARSON || metric > OR_ABSDIFF) return -1;
    const int brw = bw / 2, brh = bh / 2, brd = bd / 2;
    int K;
    int64_t *off = vol_offsets(w, h, bw, bh, bd, &K);
    if (K > 4096) { free(off); return -1; }  /* per-thread block buffers below */
    const int64_t n = (int64_t)w * h * d;
    float *means = malloc(sizeof(float) * n), *stds = malloc(sizeof(float) * n), *mask = malloc(sizeof(float) * n);
    double *S = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n);
    or_local_stats3d_f32(img, w, h, d, bw, bh, bd, means, stds);
    float nv = or_vol_noise_var(stds, w, h, d, brw, brh, brd);
    float thr = nv * c;
    float np = 0.0f;
    for (int64_t o = 0; o < n; o++) {
        mask[o] = 0.0f;
        if (!vol_inner(o, w, h, d, brw, brh, brd)) continue;
        mask[o] = (stds[o] * stds[o] <= thr) ? 0.0f : 1.0f;
        if (mask[o] > 0.0f) np += 1.0f;
    }
    const double fp_d = or_java_float_literal(nv), thr_d = or_java_float_literal(thr), eps_d = or_java_float_literal(EPS_F);
    const float c1 = 0.0001f, c2 = 0.0009f;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t o = 0; o < n; o++) {
        S[o] = 0.0; W[o] = 0.0;
        if (!vol_inner(o, w, h, d, brw, brh, brd)) continue;
        const float rs = stds[o], rv = rs * rs;
        if ((double)rv < thr_d || rv == 0.0f) continue;
        float rb[4096], tb[4096];
        double rbd[4096], tbd[4096];
        double rmd = 0.0;
        const float rm = means[o];
        if (fp64) {
            for (int i = 0; i < K; i++) rmd += img[o + off[i]];
            rmd /= K;
            for (int i = 0; i < K; i++) rbd[i] = (double)img[o + off[i]] - rmd;
        } else {
            for (int i = 0; i < K; i++) rb[i] = img[o + off[i]] - rm;
        }
        double rsd = 0.0;
        if (fp64) {
            for (int i = 0; i < K; i++) rsd += rbd[i] * rbd[i];
            rsd = sqrt(rsd / (K - 1));
        }
        float acc = 0.0f, wacc = 0.0f;
        double accd = 0.0, waccd = 0.0;
        for (int64_t t = 0; t < n; t++) {
            if (!vol_inner(t, w, h, d, brw, brh, brd)) continue;
            const float ts = stds[t], tv = ts * ts;
            if ((double)tv < thr_d) continue;
            if (metric == OR_ABSDIFF && tv == 0.0f) continue;
            if (fp64) {
                double tmd = 0.0, tsd = 0.0;
                for (int i = 0; i < K; i++) tmd += img[t + off[i]];
                tmd /= K;
                for (int i = 0; i < K; i++) { tbd[i] = (double)img[t + off[i]] - tmd; tsd += tbd[i] * tbd[i]; }
                tsd = sqrt(tsd / (K - 1));
                double wt = exp(-((rsd - tsd) * (rsd - tsd)) / (fp_d + eps_d));
                waccd += wt;
                if (metric == OR_ABSDIFF) { accd += fabs(rsd - tsd) * wt; continue; }
                if (metric == OR_NRMSE) {
                    double mse = 0.0;
                    for (int i = 0; i < K; i++) { double e = rbd[i] - tbd[i]; mse += e * e; }
                    accd += sqrt(mse / K) * wt;
                    continue;
                }
                if (ts == 0.0f) continue;  /* reference non-flat, test flat: 0 */
                double cov = 0.0;
                for (int i = 0; i < K; i++) cov += rbd[i] * tbd[i];
                cov /= (K - 1);
                if (metric == OR_PEARSON) {
                    double r = cov / (rsd * tsd + eps_d);
                    accd += (r > 0.0 ? r : 0.0) * wt;
                } else {
                    double ss = (2.0 * rmd * tmd + c1) * (2.0 * cov + c2) / ((rmd * rmd) * (tmd * tmd) + c1) /
                                (rsd * rsd + tsd * tsd + c2) * wt;
                    accd += ss > 0.0 ? ss : 0.0;
                }
            } else {
                float dd = rs - ts;
                if (metric == OR_ABSDIFF) {
                    float wt = (float)exp(-1.0 * ((double)(dd * dd) / (fp_d + eps_d)));
                    wacc += wt;
                    acc += fmaxf(0.0f, fabsf(dd)) * wt;
                    continue;
                }
                const float tm = means[t];
                for (int i = 0; i < K; i++) tb[i] = img[t + off[i]] - tm;
                float q = (float)((double)(dd * dd) / (fp_d + eps_d));
                float wt = (float)expf((float)(-1.0f) * q);
                wacc += wt;
                if (metric == OR_NRMSE) {
                    float mse = 0.0f;
                    for (int i = 0; i < K; i++) mse += (rb[i] - tb[i]) * (rb[i] - tb[i]);
                    mse /= (float)K;
                    acc += sqrtf(mse) * wt;
                    continue;
                }
                if (ts == 0.0f) continue;
                float cov = 0.0f;
                for (int i = 0; i < K; i++) cov += rb[i] * tb[i];
                cov /= (float)(K - 1);
                if (metric == OR_PEARSON) {
                    float r = (float)((double)cov / ((double)(rs * ts) + eps_d));
                    acc += fmaxf(0.0f, r) * wt;
                } else {
                    float num = (2.0f * rm * tm + c1) * (2.0f * cov + c2);
                    float den = ((rm * rm) * (tm * tm) + c1) * ((rs * rs) + (ts * ts) + c2);
                    acc += fmaxf(0.0f, num / den * wt);
                }
            }
        }
        S[o] = fp64 ? accd : (double)acc;
        W[o] = fp64 ? waccd : (double)wacc;
    }
    for (int64_t o = 0; o < n; o++) {
        double r = 0.0;
        if (vol_inner(o, w, h, d, brw, brh, brd)) {
            if (fp64) r = S[o] / (W[o] * (double)np + eps_d);
            else {
                float den = (float)W[o] * np;
                den = den + EPS_F;
                r = (double)((float)S[o] / den);
            }
            if ((metric == OR_NRMSE || metric == OR_ABSDIFF) && mask[o] > 0.0f)
                r = (r == 0.0) ? 1.0 : (fp64 ? 1.0 / r : (double)(1.0f / (float)r));
        }
        if (c > 0.0f) r = fp64 ? r * mask[o] : (double)((float)r * mask[o]);
        out[o] = r;
    }
    if (normalize) {
        double mn = DBL_MAX, mx = -DBL_MAX;
        for (int64_t o = 0; o < n; o++)
            if (vol_inner(o, w, h, d, brw, brh, brd) && mask[o] > 0.0f) {
                mn = out[o] < mn ? out[o] : mn;
                mx = out[o] > mx ? out[o] : mx;
            }
        for (int64_t o = 0; o < n; o++)
            if (vol_inner(o, w, h, d, brw, brh, brd) && mask[o] > 0.0f) {
                if (fp64) out[o] = (out[o] - mn) / (mx - mn + eps_d);
                else out[o] = (double)(((float)out[o] - (float)mn) / ((float)mx - (float)mn + EPS_F));
            }
    }
    if (S_out) for (int64_t o = 0; o < n; o++) S_out[o] = (float)S[o];
    if (W_out) for (int64_t o = 0; o < n; o++) W_out[o] = (float)W[o];
    if (mask_out) memcpy(mask_out, mask, sizeof(float) * n);
    if (noise_var) *noise_var = nv;
    if (n_pixels) *n_pixels = np;
    free(off); free(means); free(stds); free(mask); free(S); free(W);
    return 0;
}

int or_global_pearson3d(const float *img, int w, int h, int d, int bw, int bh, int bd, float c, int normalize,
                        int fp64, double *out, float *S_out, float *W_out, float *mask_out, float *noise_var,
                        float *n_pixels)
{
    return or_global_repetition3d(OR_PEARSON, img, w, h, d, bw, bh, bd, c, normalize, fp64, out, S_out, W_out,
                                  mask_out, noise_var, n_pixels);
}

/* ------------------------------------------------------------------------ */
/* Volumes: BlockRepetition3D_, as intended                                 */
/* ------------------------------------------------------------------------ */

static void java_normalize_array(float *a, int n)
{   /* Utils.normalizeArray (Utils.java:1642-1660), float */
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int i = 0; i < n; i++) { mn = fminf(a[i], mn); mx = fmaxf(a[i], mx); }
    for (int i = 0; i < n; i++) a[i] = (a[i] - mn) / (mx - mn + EPS_F);
}

/* Utils.getReferenceBlock3D (Utils.java:1047-1155): the bw x bh x bd block
 * (slice z at z*bw*bh) restricted to the ellipsoid in z, y, x order,
 * normalised, mean-subtracted, normalised again; mean and unbiased std of
 * the result (float sums in order).  Returns K, or -1 for even sides or a
 * depth below 3 (:1066-1076; sides below 3 give an empty ellipsoid). */
int or_reference_block3d(const float *blk, int bw, int bh, int bd, float *px, float *mean, float *std_)
{
    if (bw % 2 == 0 || bh % 2 == 0 || bd % 2 == 0 || bd < 3 || bw < 3 || bh < 3) return -1;
    int K;
    int64_t *off = vol_offsets(bw, bh, bw, bh, bd, &K);
    const int64_t c = ((int64_t)(bd / 2) * bh + bh / 2) * bw + bw / 2;
    for (int i = 0; i < K; i++) px[i] = blk[c + off[i]];
    free(off);
    java_normalize_array(px, K);
    float m = 0.0f;
    for (int i = 0; i < K; i++) m += px[i];
    m /= (float)K;
    for (int i = 0; i < K; i++) px[i] = px[i] - m;
    java_normalize_array(px, K);
    m = 0.0f;
    for (int i = 0; i < K; i++) m += px[i];
    m /= (float)K;
    float s = 0.0f;
    for (int i = 0; i < K; i++) s += (px[i] - m) * (px[i] - m);
    s = (float)sqrt((double)(s / (float)(K - 1)));
    *mean = m;
    *std_ = s;
    return K;
}

/* Raw per-voxel output of kernelGetBlock{Pearson,Ssim,Nrmse,AbsDiffStds}3D.cl
 * with the block radii the kernels mean (getBlockPearson3D passes the block
 * depth as bRZ, CLUtils.java:1149; getBlockSsim3D the block width as bRW,
 * :1310).  The macro literals ref_mean, ref_std and EPSILON are double, so
 * the remaps divide in double, Pearson's quotient is double and SSIM's
 * numerator is rounded to float before its double division
 * (kernelGetBlockSsim3D.cl:101).  Borders 0. */
int or_block_raw3d(int metric, const float *img, const float *means, const float *stds, int w, int h, int d,
                   const float *tmpl, int bw, int bh, int bd, float tmpl_mean, float tmpl_std, float *out)
{
    if (metric < OR_PEARSON || metric > OR_ABSDIFF) return -1;
    const int brw = bw / 2, brh = bh / 2, brd = bd / 2;
    int K;
    int64_t *off = vol_offsets(w, h, bw, bh, bd, &K);
    if (K > 8192) { free(off); return -1; }
    const double eps_d = or_java_float_literal(EPS_F);
    const double std_d = or_java_float_literal(tmpl_std);
    const double mean_d = or_java_float_literal(tmpl_mean);
    const int64_t n = (int64_t)w * h * d;
#pragma omp parallel for schedule(static)
    for (int64_t zy = 0; zy < (int64_t)d * h; zy++) {
        float cb[8192];
        for (int x = 0; x < w; x++) {
            const int64_t o = zy * w + x;
            if (!vol_inner(o, w, h, d, brw, brh, brd)) { out[o] = 0.0f; continue; }
            const float cs = stds[o];
            if (metric == OR_ABSDIFF) { out[o] = fabsf((float)std_d - cs); continue; }  /* AbsDiffStds3D.cl:28 */
            for (int i = 0; i < K; i++) cb[i] = img[o + off[i]];
            const float cm = means[o];
            if (metric == OR_PEARSON) {  /* kernelGetBlockPearson3D.cl:58-112 */
                float mn = FLT_MAX, mx = -FLT_MAX;
                for (int i = 0; i < K; i++) { mn = fminf(mn, cb[i]); mx = fmaxf(mx, cb[i]); }
                for (int i = 0; i < K; i++) cb[i] = (float)((double)(cb[i] - mn) / ((double)(mx - mn) + eps_d));
                for (int i = 0; i < K; i++) cb[i] = cb[i] - cm;
                mn = FLT_MAX; mx = -FLT_MAX;
                for (int i = 0; i < K; i++) { mn = fminf(mn, cb[i]); mx = fmaxf(mx, cb[i]); }
                for (int i = 0; i < K; i++) cb[i] = (float)((double)(cb[i] - mn) / ((double)(mx - mn) + eps_d));
                float cov = 0.0f;
                for (int i = 0; i < K; i++) cov += tmpl[i] * cb[i];
                cov /= (float)(K - 1);
                if (std_d == 0.0 && cs == 0.0f) out[o] = 1.0f;
                else if (std_d == 0.0 || cs == 0.0f) out[o] = 0.0f;
                else out[o] = fmaxf(0.0f, (float)((double)cov / (std_d * (double)cs + eps_d)));
                continue;
            }
            for (int i = 0; i < K; i++) cb[i] = cb[i] - cm;
            if (metric == OR_NRMSE) {  /* kernelGetBlockNrmse3D.cl:61-75 */
                float mse = 0.0f;
                for (int i = 0; i < K; i++) mse += (tmpl[i] - cb[i]) * (tmpl[i] - cb[i]);
                mse /= (float)K;
                out[o] = sqrtf(mse);
                continue;
            }
            /* kernelGetBlockSsim3D.cl:66-103 */
            float cov = 0.0f;
            for (int i = 0; i < K; i++) cov += tmpl[i] * cb[i];
            cov /= (float)(K - 1);
            const float c1 = 0.0001f, c2 = 0.0009f;
            if (std_d == 0.0 && cs == 0.0f) { out[o] = 1.0f; continue; }
            if (std_d == 0.0 || cs == 0.0f) { out[o] = 0.0f; continue; }
            const float num = (float)((2.0 * mean_d * (double)cm + (double)c1) * (double)(2.0f * cov + c2));
            const double den = ((mean_d * mean_d) + (double)(cm * cm) + (double)c1) *
                               ((std_d * std_d) + (double)(cs * cs) + (double)c2);
            out[o] = fmaxf(0.0f, (float)((double)num / den));
        }
    }
    (void)n;
    free(off);
    return 0;
}

static void vol_normalize_map(float *map, int w, int h, int d, int brw, int brh, int brd, const float *mask)
{   /* Utils.normalizeImage3D (Utils.java:1572-1637) */
    const int64_t n = (int64_t)w * h * d;
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int64_t o = 0; o < n; o++)
        if (vol_inner(o, w, h, d, brw, brh, brd) && (!mask || mask[o] > 0.0f)) {
            mn = fminf(map[o], mn);
            mx = fmaxf(map[o], mx);
        }
    for (int64_t o = 0; o < n; o++)
        if (vol_inner(o, w, h, d, brw, brh, brd) && (!mask || mask[o] > 0.0f))
            map[o] = (map[o] - mn) / (mx - mn + EPS_F);
}

/* CLUtils.calculateBlockRepetitionMap3D (CLUtils.java:1639-1653) on a
 * normalised volume: the statistics, the raw kernel output, then the host
 * steps of getBlockPearson3D / getBlockSsim3D (:1208-1233, :1335-1360),
 * getBlockNrmse3D (:1447-1496: 1/rmse, 1 where rmse == 0) and
 * getBlockAbsDiffStds3D (:1585-1620: 1 - d), each followed by the 3D mask
 * when c > 0 (Utils.getRelevanceMask3D, Utils.applyMask3D) and
 * Utils.normalizeImage3D over the mask (all inner voxels when c <= 0). */
int or_block_map3d(int metric, const float *img, int w, int h, int d, const float *tmpl, int bw, int bh, int bd,
                   float tmpl_mean, float tmpl_std, float c, int normalize, float *out)
{
    const int brw = bw / 2, brh = bh / 2, brd = bd / 2;
    const int64_t n = (int64_t)w * h * d;
    float *means = malloc(sizeof(float) * n), *stds = malloc(sizeof(float) * n), *mask = NULL;
    or_local_stats3d_f32(img, w, h, d, bw, bh, bd, means, stds);
    int rc = or_block_raw3d(metric, img, means, stds, w, h, d, tmpl, bw, bh, bd, tmpl_mean, tmpl_std, out);
    if (rc) goto done;
    for (int64_t o = 0; o < n; o++) {
        if (!vol_inner(o, w, h, d, brw, brh, brd)) continue;
        if (metric == OR_NRMSE) out[o] = (out[o] == 0.0f) ? 1.0f : 1.0f / out[o];
        else if (metric == OR_ABSDIFF) out[o] = 1.0f - out[o];
    }
    if (c > 0.0f) {
        mask = malloc(sizeof(float) * n);
        const float thr = or_vol_noise_var(stds, w, h, d, brw, brh, brd) * c;
        for (int64_t o = 0; o < n; o++) {
            mask[o] = 0.0f;
            if (vol_inner(o, w, h, d, brw, brh, brd)) mask[o] = (stds[o] * stds[o] <= thr) ? 0.0f : 1.0f;
            out[o] *= mask[o];
        }
    }
    if (normalize) vol_normalize_map(out, w, h, d, brw, brh, brd, mask);
done:
    free(means); free(stds); free(mask);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Volumes: OcTree_ GAT estimate                                             */
/* ------------------------------------------------------------------------ */

/* OcTree_.computeCriterion (OcTree_.java:216-228): the node's variance from
 * Utils.getMeanAndVarBlock3D (Utils.java:1356-1380: double sums of v and v*v
 * in z, y, x order, var = sq/n - mean^2, cast to float) and
 * calculateNoiseVariance3D (:230-280): pseudo-residuals (7 v - the six
 * neighbours inside the volume) / sqrt(56), their float mean and the float
 * sum of (float)((r - mean)^2) over n - 1. */
void or_oc_criterion(const float *img, int w, int h, int d, int x0, int y0, int z0, int nw, int nh, int nd,
                     float *block_var, float *noise_var)
{
    const int64_t n = (int64_t)nw * nh * nd, plane = (int64_t)w * h;
    double mean = 0.0, sq = 0.0;
    for (int z = z0; z < z0 + nd; z++)
        for (int y = y0; y < y0 + nh; y++)
            for (int x = x0; x < x0 + nw; x++) {
                double v = (double)img[z * plane + (int64_t)y * w + x];
                mean += v;
                sq += v * v;
            }
    mean = mean / (double)n;
    *block_var = (float)(sq / (double)n - mean * mean);

    const float l = 7.0f;
    const float scale = (float)(1.0 / sqrt((double)(l * l + l)));
    float sum = 0.0f, mr = 0.0f, ss = 0.0f;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) mr = sum / (float)n;
        for (int z = z0; z < z0 + nd; z++)
            for (int y = y0; y < y0 + nh; y++)
                for (int x = x0; x < x0 + nw; x++) {
                    const float *p = img + z * plane + (int64_t)y * w + x;
                    float lap = l * p[0];
                    if (x > 0) lap -= p[-1];
                    if (x < w - 1) lap -= p[1];
                    if (y > 0) lap -= p[-w];
                    if (y < h - 1) lap -= p[w];
                    if (z > 0) lap -= p[-plane];
                    if (z < d - 1) lap -= p[plane];
                    float r = lap * scale;
                    if (pass == 0) {
                        sum += r;
                    } else {
                        double dd = (double)r - (double)mr;
                        ss += (float)(dd * dd);  /* Math.pow(dd, 2.0) */
                    }
                }
    }
    *noise_var = ss / ((float)n - 1.0f);
}

/* OcTree_.calculateRobustMean (:316-366): as the quadtree's, over the box. */
float or_oc_robust_mean(const float *img, int w, int h, int x0, int y0, int z0, int nw, int nh, int nd,
                        int max_iter)
{
    const int64_t plane = (int64_t)w * h;
    float mean = 0.0f, var = 0.0f;
    int n = 0;
    for (int z = z0; z < z0 + nd; z++)
        for (int y = y0; y < y0 + nh; y++)
            for (int x = x0; x < x0 + nw; x++) {
                float v = img[z * plane + (int64_t)y * w + x];
                mean += v;
                var += (v - mean) * (v - mean);
                n++;
            }
    mean = mean / (float)n;
    var = var / (float)(n - 1);
    var = jmaxf(var, OR_JEPS);
    float prev;
    int it = 0;
    do {
        prev = mean;
        float ws = 0.0f, wv = 0.0f;
        for (int z = z0; z < z0 + nd; z++)
            for (int y = y0; y < y0 + nh; y++)
                for (int x = x0; x < x0 + nw; x++) {
                    float v = img[z * plane + (int64_t)y * w + x];
                    float wt = (float)exp(-((double)(v - mean) * (double)(v - mean)) / (2.0 * (double)var));
                    ws += wt;
                    wv += wt * v;
                }
        mean = wv / jmaxf(ws, OR_JEPS);
        it++;
    } while ((double)fabsf(mean - prev) > 1e-5 && it < max_iter);
    return mean;
}

/* OcTree_.calculateLTSVariance (:382-420). */
float or_oc_lts_variance(const float *img, int w, int h, int x0, int y0, int z0, int nw, int nh, int nd, float mean,
                         float frac)
{
    const int64_t plane = (int64_t)w * h;
    int n = nw * nh * nd, k = 0;
    float *r = (float *)malloc(sizeof(float) * (size_t)n);
    for (int z = z0; z < z0 + nd; z++)
        for (int y = y0; y < y0 + nh; y++)
            for (int x = x0; x < x0 + nw; x++) {
                float v = img[z * plane + (int64_t)y * w + x];
                r[k++] = (v - mean) * (v - mean);
            }
    qsort(r, (size_t)n, sizeof(float), cmp_float_asc);
    int hh = (int)floorf(frac * (float)n);
    float s = 0.0f;
    for (int i = 0; i < hh; i++) s += r[i];
    free(r);
    return s / ((float)hh + OR_JEPS);
}

typedef struct {
    qt_build q;     /* thresholds cache, alpha, cap */
    int d, bad, leaves;
    int32_t *boxes;
} oc_build;

static void oc_visit(oc_build *b, int x, int y, int z, int nw, int nh, int nd)
{
    if (b->bad) return;
    if (nw * nh <= 1) {  /* FDistribution(0, 0): NotStrictlyPositiveException */
        b->bad = -1;
        return;
    }
    float bv, nv;
    or_oc_criterion(b->q.img, b->q.w, b->q.h, b->d, x, y, z, nw, nh, nd, &bv, &nv);
    /* performFisherTest (:282-296) with regionSize = width * height (:225) */
    float ratio = jminf(bv, nv) / jmaxf(bv, nv);
    int divide = ratio < qt_threshold(&b->q, nw * nh);
    if (nw <= b->q.min_size || nh <= b->q.min_size || nd <= b->q.min_size || !divide) {
        if (b->leaves < b->q.cap) {
            int32_t *r = b->boxes + 6 * b->leaves;
            r[0] = x, r[1] = y, r[2] = z, r[3] = nw, r[4] = nh, r[5] = nd;
        }
        b->leaves++;
        return;
    }
    int hw = nw / 2, hh = nh / 2, hd = nd / 2;  /* splitNode (:197-213): children 0..7, x fastest */
    for (int c = 0; c < 8; c++)
        oc_visit(b, x + (c & 1) * hw, y + ((c >> 1) & 1) * hh, z + (c >> 2) * hd, hw, hh, hd);
}

/* The "Quad/Octree" branch of Utils.getInputImage3D (Utils.java:1234-1254),
 * the same steps as the OcTree_ plugin (OcTree_.java:82-96): OcTree_(w, h,
 * d, min_size, alpha).buildTree (leaf when a side is <= min_size or the
 * F test says keep), calculateRobustMeans(max_iter),
 * calculateLTSVariances(frac), collectMeanVariancePairs (leaves in
 * pre-order) and performLinearRegression, whose slope and intercept are cast
 * to float (:458-469).  boxes (x, y, z, w, h, d), means and vars receive up
 * to cap leaves; *nleaves the count.  Returns 0, -1 for a visited node of
 * one pixel per slice (FDistribution(0, 0)) or fewer than 3 leaves, -2 if
 * cap is too small. */
int or_octree_gat(const float *img, int w, int h, int d, int min_size, float alpha, int max_iter, float frac,
                  int cap, int32_t *boxes, float *means, float *vars, int *nleaves, double *g0, double *edc)
{
    oc_build b;
    memset(&b, 0, sizeof b);
    b.q.img = img, b.q.w = w, b.q.h = h, b.q.min_size = min_size, b.q.alpha = alpha, b.q.cap = cap;
    b.d = d, b.boxes = boxes;
    oc_visit(&b, 0, 0, 0, w, h, d);
    *nleaves = b.leaves;
    if (b.bad) return b.bad;
    if (b.leaves > cap) return -2;
    for (int i = 0; i < b.leaves; i++) {
        const int32_t *r = boxes + 6 * i;
        means[i] = or_oc_robust_mean(img, w, h, r[0], r[1], r[2], r[3], r[4], r[5], max_iter);
        vars[i] = or_oc_lts_variance(img, w, h, r[0], r[1], r[2], r[3], r[4], r[5], means[i], frac);
    }
    if (b.leaves < 3) return -1;
    double *x = (double *)malloc(sizeof(double) * 2 * (size_t)b.leaves), *y = x + b.leaves;
    for (int i = 0; i < b.leaves; i++) {
        x[i] = (double)means[i];
        y[i] = (double)vars[i];
    }
    double s, c;
    or_simple_regression(x, y, b.leaves, &s, &c);
    *g0 = (double)(float)s;
    *edc = (double)(float)c;
    free(x);
    return 0;
}

/* OcTree_.applyGATtree (:471-485) in place with float g0, eDC:
 * ref = 3/8 g0^2 + eDC; 0 where v < -ref / g0. */
void or_gat_tree3d(float *px, int64_t n, float g0, float edc)
{
    const double g = (double)g0;
    const double ref = 3.0 / 8.0 * g * g + (double)edc;
    for (int64_t i = 0; i < n; i++) {
        double v = (double)px[i];
        if (v < -ref / g)
            v = 0.0;
        else
            v = (2.0 / g) * sqrt(g * v + ref);
        px[i] = (float)v;
    }
}

/* ------------------------------------------------------------------------ */
/* Volumes: GATMinimizer3D                                                   */
/* ------------------------------------------------------------------------ */

/* GATMinimizer3D's block grid (GATMinimizer3D.java:108-127): sides
 * max(w/6, 1), max(h/6, 1), max(d/6, 1), halved (not below 1) along each
 * axis with fewer than 2 blocks until every axis has 2 or all sides are 1.
 * Returns 0, or -1 where the reference's loop never ends: an axis with one
 * block at side 1 (a 1-voxel axis) while another side is above 1. */
int or_gat_blocks3d(int w, int h, int d, int *bw, int *bh, int *bd, int *nbx, int *nby, int *nbz)
{
    int a = w / 6 > 1 ? w / 6 : 1, b = h / 6 > 1 ? h / 6 : 1, c = d / 6 > 1 ? d / 6 : 1;
    int nx = w / a, ny = h / b, nz = d / c;
    while (nx < 2 || ny < 2 || nz < 2) {
        const int a0 = a, b0 = b, c0 = c;
        if (nx < 2 && a > 1) a = a / 2 > 1 ? a / 2 : 1;
        if (ny < 2 && b > 1) b = b / 2 > 1 ? b / 2 : 1;
        if (nz < 2 && c > 1) c = c / 2 > 1 ? c / 2 : 1;
        nx = w / a, ny = h / b, nz = d / c;
        if (a == 1 && b == 1 && c == 1) break;
        if (a == a0 && b == b0 && c == c0) return -1;
    }
    *bw = a, *bh = b, *bd = c, *nbx = nx, *nby = ny, *nbz = nz;
    return 0;
}

/* GATMinimizer3D.userFunction (:79-157): NaN for gain <= 0, sigma < 0 or
 * offset < 0; else the volume transformed (GATMinimizer2D.applyGAT: 0 where
 * v <= -ref/gain), per block Utils.getMeanAndVarBlock3D (double v, double
 * v*v; var = sq/n - mean^2) and the mean of (var - 1)^2 over the blocks.
 * The sums run in the device's fixed order, as or_gat_objective's: a row's
 * element i to lane i mod 64, a block's row r (z, then y) to lane r mod 64,
 * block b (z, y, x order) to lane b mod 64, each combined by lane_tree64. */
double or_gat_objective3d(const float *px, int w, int h, int d, double gain, double sigma, double offset)
{
    if (gain <= 0.0 || sigma < 0.0 || offset < 0.0) return NAN;
    int bw, bh, bd, nbx, nby, nbz;
    if (or_gat_blocks3d(w, h, d, &bw, &bh, &bd, &nbx, &nby, &nbz)) return NAN;
    const int nb = nbx * nby * nbz;
    const int64_t plane = (int64_t)w * h;
    double ref = (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset;
    double be[64], bm[64], bs[64], lm[64], ls[64];
    for (int l = 0; l < 64; l++) be[l] = 0.0;
    for (int bz = 0; bz < nbz; bz++)
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                for (int l = 0; l < 64; l++) bm[l] = bs[l] = 0.0;
                for (int r = 0; r < bh * bd; r++) {
                    const int zz = r / bh, yy = r - zz * bh;
                    const float *row = px + (bz * bd + zz) * plane + (int64_t)(by * bh + yy) * w + (int64_t)bx * bw;
                    for (int l = 0; l < 64; l++) lm[l] = ls[l] = 0.0;
                    for (int i = 0; i < bw; i++) {
                        double v = (double)gat_fit_px(row[i], gain, ref);
                        lm[i % 64] += v;
                        ls[i % 64] += v * v;
                    }
                    bm[r % 64] += lane_tree64(lm);
                    bs[r % 64] += lane_tree64(ls);
                }
                double mean = lane_tree64(bm), sq = lane_tree64(bs);
                int64_t n = (int64_t)bw * bh * bd;
                mean = mean / (double)n;
                double var = sq / (double)n - mean * mean;
                double delta = var - 1.0;
                int b = (bz * nby + by) * nbx + bx;
                be[b % 64] += (delta * delta) / (double)nb;
            }
    return lane_tree64(be);
}

typedef struct { const float *px; int w, h, d; } gat3_ctx;
static double gat3_obj(const double *x, void *c)
{
    const gat3_ctx *g = (const gat3_ctx *)c;
    return or_gat_objective3d(g->px, g->w, g->h, g->d, x[0], x[1], x[2]);
}

/* GATMinimizer3D.run (:40-70): the same simplex as the 2D fit (steps 1, 10,
 * 100) over the 3D objective; NaN ranks worst. */
double or_gat_minimize3d(const float *px, int w, int h, int d, double gain0, double sigma0, double offset0,
                         int max_iter, double *params)
{
    int g[6];
    if (or_gat_blocks3d(w, h, d, g, g + 1, g + 2, g + 3, g + 4, g + 5)) return NAN;  /* the reference never ends */
    gat3_ctx c = {px, w, h, d};
    double x0[3] = {gain0, sigma0, offset0}, dd[3] = {1.0, 10.0, 100.0};
    return or_nelder_mead3(gat3_obj, &c, x0, dd, max_iter, params);
}

/* RelevanceMaskStack2D_ after the tools' input step (RelevanceMaskStack2D_
 * .java:140-285): kernelGetRelevanceMask2D.cl's local variances, then for
 * the constants c_0 = lower, c_{i+1} = c_i + step (a float sum), n =
 * (int) ceil((upper - lower) / step) + 1 of them (:262-263), the mask of
 * OptimiseRelevanceMask_.getRelevanceMap (OptimiseRelevanceMask_.java:
 * 297-380: float tile-sum noise variance, threshold nv * c or c when nv is
 * 0, 0 where var^2 < threshold or var^2 == 0) into masks[i].  Odd square
 * blocks only (the plugin swaps the radii).  Returns 0, -1 for another
 * block, a step <= 0 or NaN or fewer than 1 constant, -2 if more than cap
 * (*n holds the count either way). */
int or_relevance_mask_stack(const float *img, int w, int h, int bw, int bh, float lower, float upper, float step,
                            int cap, float *masks, float *constants, int *n)
{
    *n = 0;
    if (bw != bh || bw % 2 == 0 || bw < 3 || !(step > 0.0f)) return -1;
    double q = ceil((double)((upper - lower) / step));
    if (!(q >= 0.0) || q > 1e8) return -1;
    int count = (int)q + 1;
    *n = count;
    if (count > cap) return -2;
    int64_t np = (int64_t)w * h;
    float *var = (float *)malloc(sizeof(float) * (size_t)np);
    or_local_var_f32(img, w, h, bw, bh, var);
    float nv = or_noise_var(img, w, h);
    int br = bw / 2;
    float c = lower;
    for (int i = 0; i < count; i++) {
        float thr = (nv == 0.0f) ? c : nv * c;
        float *m = masks + (size_t)i * np;
        memset(m, 0, sizeof(float) * (size_t)np);
        for (int y = br; y < h - br; y++)
            for (int x = br; x < w - br; x++) {
                float s = var[(int64_t)y * w + x];
                float v = s * s;
                m[(int64_t)y * w + x] = (v < thr || v == 0.0f) ? 0.0f : 1.0f;
            }
        constants[i] = c;
        c += step;
    }
    free(var);
    return 0;
}
