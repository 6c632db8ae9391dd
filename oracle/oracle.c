/*
 * oracle.c -- CPU restatement of RecView's FBP path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Built with
 * -ffp-contract=off so that every float expression rounds once per operation,
 * in the order the reference source writes it (MSVC x64 /fp:precise does not
 * contract a*b+c into an FMA either).
 *
 * Parity pinning of each function is listed in the test files that check it
 * (tests/test_oracle_*.py): closed forms, brute-force DFTs and invariants.
 * orc_fft1rev/orc_fft1 reproduce CFft's operation order; their rounding is
 * pinned only against the closed form (to FP32 tolerance), because cfft.cpp
 * cannot be compiled here (it needs MFC headers).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>

/* constants.h:3,6-8,39-41,53 */
#define RV_DEG_TO_RAD 0.0174532925199
#define RV_PI 3.1415926535897
#define RV_LOG2 0.69314718055994530941723212145818
#define RV_BACKPROJ_SCALE 100
#define RV_GINTP 16

int orc_geometry(int ixdim, int iInterpolation, double center, int float_center, orc_geom* g)
{
    /* cudaReconstHost.cpp:636-641 (CDLGRECONST_OPT_ZOOMING_NONE = 2, constants.h:64) */
    const int izoom = (iInterpolation > 2) ? (iInterpolation - 2) : 0;
    const int z = (int)pow(2.0, izoom);
    const int ndimp = (int)(log((double)ixdim) / RV_LOG2) + 1 + izoom;
    const int ndim = (int)pow(2.0, ndimp);
    const int N = ixdim * z;
    const int h = N / 2;
    const float fC = (float)center;
    int ic;
    float fc16;
    if (float_center) {
        /* cudaReconstHost.cpp:640,655; cudaReconst.cu:220-221 */
        ic = (int)fC;
        fc16 = (z == 1) ? (h + fC - (int)fC) * RV_GINTP
                        : (h + (fC - (int)fC) * z) * RV_GINTP;
    } else {
        /* general.cpp:2027, 2034-2035, 2109 */
        ic = (int)center;
        fc16 = (z == 1) ? (h + (float)center - (int)center) * RV_GINTP
                        : (h + ((float)center - (int)center) * z) * RV_GINTP;
    }
    g->ixdim = ixdim; g->izoom = izoom; g->z = z; g->ndimp = ndimp; g->ndim = ndim;
    g->N = N; g->h = h; g->ic = ic; g->fcenter16 = fc16;
    g->ihoffset = ndim / 2 - 1 - h;              /* cudaReconstHost.cpp:654 */
    return (ndimp < 3) ? -1 : 0;                 /* CFft::Init1 rejects np < 3 (cfft.cpp:375) */
}

void orc_set_filter(int ndim, int ifilter, double cutoff, double order, float* f)
{
    /* gazoDoc.cpp:3761-3809 */
    const int ndimh = (int)(ndim * cutoff);
    const float forder = (float)order;
    const float revndim = 1.f / ndim;
    int i;
    switch (ifilter) {
    case 0:
        for (i = 0; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = 0.5f * revndim * i * (float)(1 + cos(RV_PI * frac));
        }
        for (i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 1:
        for (i = 0; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * (float)(0.54 + 0.46 * cos(RV_PI * frac));
        }
        for (i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 2:
        for (i = 0; i < ndimh; i++) f[i] = revndim * i;
        for (i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 3:
        for (i = 0; i < ndimh / 2; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * (1 - 6 * frac * frac * (1 - frac));
        }
        for (i = ndimh / 2; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * 2 * (1 - frac) * (1 - frac) * (1 - frac);
        }
        for (i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 4:
        /* float overloads of pow/sqrt (C++ <cmath>), as the reference's C++ resolves them */
        for (i = 0; i < ndim; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i / (float)sqrtf(1 + powf(frac, forder));
        }
        break;
    case 5:
        /* Shepp-Logan: no counterpart in the reference.  |k|/n * sinc(pi k / (2 ndimh)). */
        for (i = 0; i < ndimh; i++) {
            const double x = RV_PI * (double)i / (2.0 * ndimh);
            const double s = (i == 0) ? 1.0 : sin(x) / x;
            f[i] = (float)((double)i / ndim * s);
        }
        for (i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    default:
        for (i = 0; i < ndim; i++) f[i] = 1;
        break;
    }
}

/* ---------------------------------------------------------------- CFft -- */

void orc_cfft_twiddles(int np, float* wr, float* wi)
{
    /* cfft.cpp:403-409 with min0 < 0: p0 = -(float)__PI2.  `a` is a float and
     * cfft.cpp is C++, so cos(a) / sin(a) resolve to the float overloads. */
    const int n = (int)pow(2.0, np);
    float p0 = (float)6.283185307178;
    int i;
    p0 *= -1;
    for (i = 0; i < n / 2; i++) {
        const float a = p0 * ((float)i / (float)n);
        wr[i] = cosf(a);
        wi[i] = sinf(a);
    }
}

/* One radix-2 Stockham stage shared by FFTrev and FFT (cfft.cpp:228-247,
 * 312-341): butterfly b = j*al + m reads x0[b] and x0[b + n/2], twiddle
 * w[m*bt], writes x1[2*j*al + m] (sum) and x1[2*j*al + m + al] (difference). */
static void cfft_stage(int n, int al, int bt, const float* wr, const float* wi,
                       const float* x0r, const float* x0i, float* x1r, float* x1i)
{
    const int half = n / 2;
    int b;
    for (b = 0; b < half; b++) {
        const int j = b / al, m = b - (b / al) * al;
        const float ar = x0r[b], ai = x0i[b];
        const float br = x0r[b + half], bi = x0i[b + half];
        const float w_r = wr[m * bt], w_i = wi[m * bt];
        const float sr = br * w_r - bi * w_i;
        const float si = br * w_i + bi * w_r;
        const int o = 2 * j * al + m;
        x1r[o + al] = ar - sr; x1i[o + al] = ai - si;
        x1r[o] = ar + sr; x1i[o] = ai + si;
    }
}

void orc_fft1rev(int np, const float* wr, const float* wi, float* fr, float* fi, float* scratch)
{
    const int n = 1 << np, half = n / 2;
    float* x0r = scratch;         float* x0i = scratch + n;
    float* x1r = scratch + 2 * n; float* x1i = scratch + 3 * n;
    int j, k, l, al, bt;
    /* l = 0 with nmin < 0: input read in reverse, cfft.cpp:199-211 */
    for (j = 0; j < half; j++) {
        const int idx = half - 1 - j;
        const float sr = fr[idx + half], si = fi[idx + half];
        x1r[2 * j] = fr[idx] + sr;     x1i[2 * j] = fi[idx] + si;
        x1r[2 * j + 1] = fr[idx] - sr; x1i[2 * j + 1] = fi[idx] - si;
    }
    al = 2; bt = n / 4;
    for (l = 1; l < np - 1; l++) {
        float* t;
        t = x0r; x0r = x1r; x1r = t;
        t = x0i; x0i = x1i; x1i = t;
        cfft_stage(n, al, bt, wr, wi, x0r, x0i, x1r, x1i);
        al *= 2; bt /= 2;
    }
    /* l = np-1, output reversed, cfft.cpp:249-268 (no twiddle multiply for k = 0) */
    {
        const float sr = x1r[half], si = x1i[half];
        fr[0] = x1r[0] + sr;    fi[0] = x1i[0] + si;
        fr[half] = x1r[0] - sr; fi[half] = x1i[0] - si;
    }
    for (k = 1; k < half; k++) {
        const float br = x1r[half + k], bi = x1i[half + k];
        const float sr = br * wr[k] - bi * wi[k];
        const float si = br * wi[k] + bi * wr[k];
        fr[n - k] = x1r[k] + sr;    fi[n - k] = x1i[k] + si;
        fr[half - k] = x1r[k] - sr; fi[half - k] = x1i[k] - si;
    }
}

void orc_fft1(int np, const float* wr, const float* wi, float* fr, float* fi, float* scratch)
{
    const int n = 1 << np, half = n / 2;
    float* x0r = scratch;         float* x0i = scratch + n;
    float* x1r = scratch + 2 * n; float* x1i = scratch + 3 * n;
    int i, k, l, al, bt;
    /* l = 0, cfft.cpp:284-294 */
    for (i = 0; i < half; i++) {
        const float sr = fr[i + half], si = fi[i + half];
        x1r[2 * i] = fr[i] + sr;     x1i[2 * i] = fi[i] + si;
        x1r[2 * i + 1] = fr[i] - sr; x1i[2 * i + 1] = fi[i] - si;
    }
    al = 2; bt = n / 4;
    for (l = 1; l < np - 1; l++) {
        float* t;
        t = x0r; x0r = x1r; x1r = t;
        t = x0i; x0i = x1i; x1i = t;
        cfft_stage(n, al, bt, wr, wi, x0r, x0i, x1r, x1i);
        al *= 2; bt /= 2;
    }
    /* l = np-1 with nmin < 0, cfft.cpp:357-367 */
    for (k = 0; k < half; k++) {
        const float br = x1r[half + k], bi = x1i[half + k];
        const float sr = br * wr[k] - bi * wi[k];
        const float si = br * wi[k] + bi * wr[k];
        fr[k] = x1r[k] - sr;        fi[k] = x1i[k] - si;
        fr[k + half] = x1r[k] + sr; fi[k + half] = x1i[k] + si;
    }
    /* FFT1 tail, cfft.cpp:419-425: reverse and scale by 1/n (a power of two: exact) */
    {
        const float a = 1.f / n;
        for (i = 0; i < half; i++) {
            const int idx = n - 1 - i;
            const float sr = fr[i], si = fi[i];
            fr[i] = fr[idx] * a; fi[i] = fi[idx] * a;
            fr[idx] = sr * a;    fi[idx] = si * a;
        }
    }
}

/* ------------------------------------------------------------ row prep -- */

void orc_row_prep_i16(const int16_t* s, const orc_geom* g, float* re)
{
    /* cudaReconstHost.cpp:734-751 (general.cpp:2108-2126).  The two fill loops
     * are clamped to [0, n): the reference writes outside p[] when the centre
     * lies outside the detector. */
    const int n = g->ndim, W = g->ixdim, z = g->z, ic = g->ic;
    const int idx0 = (0 - ic) * z + (n / 2 - 1);
    const int idx1 = (W - 1 - ic) * z + (n / 2 - 1) + 1;
    int m, k, j;
    memset(re, 0, sizeof(float) * n);
    for (m = 0; m < idx0 && m < n; m++) re[m] = s[0];
    for (m = (idx1 > 0 ? idx1 : 0); m < n; m++) re[m] = s[W - 1];
    for (k = 0; k < W; k++) {
        const int idx = (k - ic) * z + (n / 2 - 1);
        if (idx < 0) continue;
        if (idx >= n) break;
        re[idx] = s[k];
        if (k == W - 1) break;
        for (j = 1; j < z; j++) {
            if (idx + j >= n) break;
            re[idx + j] = (float)(s[k] * (z - j) / z + s[k + 1] * j / z);
        }
    }
}

void orc_row_prep_f32(const float* s, const orc_geom* g, float* re)
{
    const int n = g->ndim, W = g->ixdim, z = g->z, ic = g->ic;
    const int idx0 = (0 - ic) * z + (n / 2 - 1);
    const int idx1 = (W - 1 - ic) * z + (n / 2 - 1) + 1;
    int m, k, j;
    memset(re, 0, sizeof(float) * n);
    for (m = 0; m < idx0 && m < n; m++) re[m] = s[0];
    for (m = (idx1 > 0 ? idx1 : 0); m < n; m++) re[m] = s[W - 1];
    for (k = 0; k < W; k++) {
        const int idx = (k - ic) * z + (n / 2 - 1);
        if (idx < 0) continue;
        if (idx >= n) break;
        re[idx] = s[k];
        if (k == W - 1) break;
        for (j = 1; j < z; j++) {
            if (idx + j >= n) break;
            re[idx + j] = (s[k] * (float)(z - j) + s[k + 1] * (float)j) / (float)z;
        }
    }
}

void orc_row_prep_cudafft(const int16_t* s, const orc_geom* g, float* re)
{
    /* cudaReconst.cu:124-134 after cudaMemset (cu:174) */
    const int n = g->ndim, W = g->ixdim, z = g->z, ic = g->ic;
    int ix, iy;
    memset(re, 0, sizeof(float) * n);
    for (ix = 0; ix < W; ix++) {
        for (iy = 0; iy < z; iy++) {
            const int idx = (ix - ic) * z + (n / 2) + iy;
            if (idx < 0 || idx >= n) continue;
            if (iy == 0) { re[idx] = s[ix]; continue; }
            if (ix == W - 1) continue;
            re[idx] = (float)(s[ix] * (z - iy) + s[ix + 1] * iy) / z;
        }
    }
}

void orc_filter_rows(const void* sino, int kind, int nrow, int row_stride,
                     const orc_geom* g, const float* filt, float* px)
{
    const int n = g->ndim, N = g->N;
    const int crop = (kind == 2) ? (n / 2 - g->h) : g->ihoffset;
    int r;
    float* wr = (float*)malloc(sizeof(float) * n / 2);
    float* wi = (float*)malloc(sizeof(float) * n / 2);
    orc_cfft_twiddles(g->ndimp, wr, wi);
    #pragma omp parallel for schedule(dynamic, 4)
    for (r = 0; r < nrow; r++) {
        float* re = (float*)malloc(sizeof(float) * n * 6);
        float* im = re + n;
        float* scratch = re + 2 * n;
        int k;
        if (kind == 1) orc_row_prep_f32((const float*)sino + (size_t)r * row_stride, g, re);
        else if (kind == 2) orc_row_prep_cudafft((const int16_t*)sino + (size_t)r * row_stride, g, re);
        else orc_row_prep_i16((const int16_t*)sino + (size_t)r * row_stride, g, re);
        memset(im, 0, sizeof(float) * n);
        orc_fft1rev(g->ndimp, wr, wi, re, im, scratch);
        for (k = 0; k < n; k++) { re[k] *= filt[k]; im[k] *= filt[k]; }   /* cudaReconstHost.cpp:753 */
        orc_fft1(g->ndimp, wr, wi, re, im, scratch);
        for (k = 0; k < N; k++) px[(size_t)r * N + k] = re[k + crop];      /* :756 */
        free(re);
    }
    free(wr); free(wi);
}

/* ----------------------------------------------------------------- LUT -- */

void orc_lut16_i32(const float* px, int N, int formula, int32_t* igp)
{
    int j, k;
    for (j = 0; j < N; j++) {
        float p0, p1p0;
        if (formula == 0) {
            /* general.cpp:2131-2136 */
            p0 = px[j] * RV_BACKPROJ_SCALE;
            p1p0 = (j == N - 1) ? 0.0f : (px[j + 1] - px[j]) / RV_GINTP * RV_BACKPROJ_SCALE;
        } else {
            /* cudaReconst.cu:237-239 */
            p0 = px[j] * RV_BACKPROJ_SCALE;
            p1p0 = (j == N - 1) ? 0.0f : (px[j + 1] * RV_BACKPROJ_SCALE - p0) / RV_GINTP;
        }
        for (k = 0; k < RV_GINTP; k++) igp[j * RV_GINTP + k] = (int32_t)(p0 + p1p0 * k);
    }
}

void orc_lut16_f32(const float* px, int N, float* lut)
{
    int j, k;
    for (j = 0; j < N; j++) {
        const float p0 = px[j];
        const float p1p0 = (j == N - 1) ? 0.0f : (px[j + 1] - p0) / RV_GINTP;
        for (k = 0; k < RV_GINTP; k++) lut[j * RV_GINTP + k] = p0 + p1p0 * k;
    }
}

void orc_bp_params(float fdeg, float tilt, const orc_geom* g, int iCenterOffset,
                   float* fcos, float* fsin, float* foffset)
{
    /* cudaReconstHost.cpp:776 (theta), cudaReconst.cu:216-222 */
    const double th = (fdeg + tilt) * RV_DEG_TO_RAD;
    const float c = (float)(cos(th) * RV_GINTP);
    const float s = (float)(-sin(th) * RV_GINTP);
    *fcos = c;
    *fsin = s;
    *foffset = g->fcenter16 - g->h * (c + s) + RV_GINTP * iCenterOffset;
}

void orc_bp_lut_i32(int32_t* img, const int32_t* lut, int N, float fcos, float fsin,
                    float foffset, int rule, int iy0, int iy1)
{
    const int ng = N * RV_GINTP;
    const float fng = (float)ng;
    int iy;
    #pragma omp parallel for schedule(static)
    for (iy = iy0; iy < iy1; iy++) {
        const float fyoff = iy * fsin + foffset;               /* general.cpp:2150 */
        uint32_t* row = (uint32_t*)img + (size_t)iy * N;
        int ix;
        for (ix = 0; ix < N; ix++) {
            const float t = ix * fcos + fyoff;                  /* :2152 */
            if (rule == 0) {
                if (t >= 0.f && t < fng) row[ix] += (uint32_t)lut[(int)t];   /* projx64.asm:259-270 */
            } else {
                int i0;
                if (!(t > -2147483648.f && t < 2147483648.f)) continue;
                i0 = (int)t;
                if (i0 < 0 || i0 >= ng) continue;                          /* :2153-2156 */
                row[ix] += (uint32_t)lut[i0];
            }
        }
    }
}

void orc_bp_lut_f64(double* img, const float* lut, int N, float fcos, float fsin,
                    float foffset, int rule)
{
    const int ng = N * RV_GINTP;
    const float fng = (float)ng;
    int iy;
    #pragma omp parallel for schedule(static)
    for (iy = 0; iy < N; iy++) {
        const float fyoff = iy * fsin + foffset;
        double* row = img + (size_t)iy * N;
        int ix;
        for (ix = 0; ix < N; ix++) {
            const float t = ix * fcos + fyoff;
            if (rule == 0) {
                if (t >= 0.f && t < fng) row[ix] += lut[(int)t];
            } else {
                int i0;
                if (!(t > -2147483648.f && t < 2147483648.f)) continue;
                i0 = (int)t;
                if (i0 < 0 || i0 >= ng) continue;
                row[ix] += lut[i0];
            }
        }
    }
}

void orc_recon_rint(const int16_t* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, int formula, int rule,
                    int32_t* img)
{
    const int N = g->N, W = g->ixdim;
    float* px = (float*)malloc(sizeof(float) * (size_t)P * N);
    int32_t* lut = (int32_t*)malloc(sizeof(int32_t) * (size_t)N * RV_GINTP);
    int i;
    orc_filter_rows(sino, 0, P, W, g, filt, px);
    memset(img, 0, sizeof(int32_t) * (size_t)N * N);
    for (i = 0; i < P; i++) {
        float c, s, o;
        orc_lut16_i32(px + (size_t)i * N, N, formula, lut);
        orc_bp_params(fdeg[i], tilt, g, 0, &c, &s, &o);
        orc_bp_lut_i32(img, lut, N, c, s, o, rule, 0, N);
    }
    free(lut); free(px);
}

void orc_recon_rf32(const float* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, int rule, double* img)
{
    const int N = g->N, W = g->ixdim;
    float* px = (float*)malloc(sizeof(float) * (size_t)P * N);
    float* lut = (float*)malloc(sizeof(float) * (size_t)N * RV_GINTP);
    int i;
    orc_filter_rows(sino, 1, P, W, g, filt, px);
    memset(img, 0, sizeof(double) * (size_t)N * N);
    for (i = 0; i < P; i++) {
        float c, s, o;
        orc_lut16_f32(px + (size_t)i * N, N, lut);
        orc_bp_params(fdeg[i], tilt, g, 0, &c, &s, &o);
        orc_bp_lut_f64(img, lut, N, c, s, o, rule);
    }
    free(lut); free(px);
}

/* --------------------------------------------------------------- O-f64 -- */

void orc_fft_f64(int n, double* re, double* im, int inverse)
{
    /* Iterative radix-2 decimation-in-time FFT, X[k] = sum x[j] e^{-+2 pi i jk/n};
     * the inverse includes the 1/n factor. */
    int i, j, len;
    for (i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (len = 2; len <= n; len <<= 1) {
        const int hl = len / 2;
        const double sgn = inverse ? 1.0 : -1.0;
        int s;
        for (s = 0; s < n; s += len) {
            int k;
            for (k = 0; k < hl; k++) {
                const double ang = sgn * 2.0 * M_PI * (double)k / (double)len;
                const double wr = cos(ang), wi = sin(ang);
                const double br = re[s + k + hl], bi = im[s + k + hl];
                const double tr = br * wr - bi * wi, ti = br * wi + bi * wr;
                re[s + k + hl] = re[s + k] - tr; im[s + k + hl] = im[s + k] - ti;
                re[s + k] += tr; im[s + k] += ti;
            }
        }
    }
    if (inverse) {
        for (i = 0; i < n; i++) { re[i] /= n; im[i] /= n; }
    }
}

void orc_filter_rows_f64(const float* sino, int nrow, int row_stride, const orc_geom* g,
                         const float* filt, double* px)
{
    /* Closed form of FFT1Rev / x H / FFT1 / .re (cfft.cpp): circular convolution
     * y = f (*) h_re, h_re[d] = (1/n) sum_k H[k] cos(2 pi k d / n), evaluated as
     * Re(ifft(fft(f) H)) in double. */
    const int n = g->ndim, N = g->N;
    int r;
    #pragma omp parallel for schedule(dynamic, 4)
    for (r = 0; r < nrow; r++) {
        float* ref = (float*)malloc(sizeof(float) * n);
        double* re = (double*)malloc(sizeof(double) * n * 2);
        double* im = re + n;
        int k;
        orc_row_prep_f32(sino + (size_t)r * row_stride, g, ref);
        for (k = 0; k < n; k++) { re[k] = ref[k]; im[k] = 0.0; }
        orc_fft_f64(n, re, im, 0);
        for (k = 0; k < n; k++) { re[k] *= (double)filt[k]; im[k] *= (double)filt[k]; }
        orc_fft_f64(n, re, im, 1);
        for (k = 0; k < N; k++) px[(size_t)r * N + k] = re[k + g->ihoffset];
        free(re); free(ref);
    }
}

void orc_bp_linear_f64(double* img, const double* px, int N, double th, double fc_minus_h,
                       int y0, int y1)
{
    const double c = cos(th), s = -sin(th);
    const double h = N / 2;
    int iy;
    #pragma omp parallel for schedule(static)
    for (iy = y0; iy < y1; iy++) {
        double* row = img + (size_t)iy * N;
        int ix;
        for (ix = 0; ix < N; ix++) {
            const double u = (ix - h) * c + (iy - h) * s + h + fc_minus_h;
            double k;
            int ik;
            if (u < 0.0 || u >= N) continue;
            k = floor(u);
            ik = (int)k;
            if (ik <= N - 2) row[ix] += px[ik] + (u - k) * (px[ik + 1] - px[ik]);
            else row[ix] += px[N - 1];
        }
    }
}

void orc_recon_of64(const float* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, double* img)
{
    const int N = g->N;
    double* px = (double*)malloc(sizeof(double) * (size_t)P * N);
    int i;
    orc_filter_rows_f64(sino, P, g->ixdim, g, filt, px);
    memset(img, 0, sizeof(double) * (size_t)N * N);
    for (i = 0; i < P; i++) {
        const double th = (fdeg[i] + tilt) * RV_DEG_TO_RAD;
        orc_bp_linear_f64(img, px + (size_t)i * N, N, th, (double)g->fcenter16 / 16.0 - g->h, 0, N);
    }
    free(px);
}

void orc_bp_rows_linear_f64(const float* px, int P, const float* fdeg, float tilt,
                            const orc_geom* g, int iCenterOffset, double* img)
{
    const int N = g->N;
    double* row = (double*)malloc(sizeof(double) * N);
    int i, k;
    memset(img, 0, sizeof(double) * (size_t)N * N);
    for (i = 0; i < P; i++) {
        const double th = (fdeg[i] + tilt) * RV_DEG_TO_RAD;
        for (k = 0; k < N; k++) row[k] = px[(size_t)i * N + k];
        orc_bp_linear_f64(img, row, N, th, (double)g->fcenter16 / 16.0 - g->h + iCenterOffset, 0, N);
    }
    free(row);
}

void orc_bp_points_linear_f64(const double* px, int P, const float* fdeg, float tilt,
                              const orc_geom* g, int iCenterOffset, const double* xs,
                              const double* ys, int npts, double* out)
{
    const int N = g->N;
    const double h = N / 2;
    const double off = (double)g->fcenter16 / 16.0 - h + iCenterOffset;
    int q;
    #pragma omp parallel for schedule(static)
    for (q = 0; q < npts; q++) {
        double acc = 0.0;
        int i;
        for (i = 0; i < P; i++) {
            const double th = (fdeg[i] + tilt) * RV_DEG_TO_RAD;
            const double u = (xs[q] - h) * cos(th) + (ys[q] - h) * -sin(th) + h + off;
            const double* row = px + (size_t)i * N;
            double k;
            int ik;
            if (u < 0.0 || u >= N) continue;
            k = floor(u);
            ik = (int)k;
            if (ik <= N - 2) acc += row[ik] + (u - k) * (row[ik + 1] - row[ik]);
            else acc += row[N - 1];
        }
        out[q] = acc;
    }
}

int orc_sinogram(uint16_t* frames, int nframes, int mult, int W, const char* binc,
                 const float* fexp, const float* fdeg, unsigned flags, int drStart, int drEnd,
                 double drX, double drY, int drOmit, const int* piDrift, int is8bit, int start,
                 int step, int binning)
{
    /* general.cpp:2394-2520; constants.h:39-48, StdAfx.h:29-31, DlgQueue.h:20-22 */
    const int pixmin = is8bit ? 3 : 20;
    double* inc = (double*)malloc(sizeof(double) * (size_t)W);
    int16_t* tmp = (int16_t*)malloc(sizeof(int16_t) * (size_t)W);
    int i0 = -1, i1 = -1, i, j, k, status = 0;
    for (i = 0; i < nframes - 1; i++) {
        if (binc[i] & 2) continue;
        if (!(binc[i] & 1)) { i0 = i; break; }
    }
    for (i = 0; i < nframes - 1; i++) {
        float t0;
        double cth, sth;
        int idrift = 0, sh;
        if ((binc[i] & 2) && !(binc[i] & 1)) continue;
        if (!(binc[i] & 1)) { i0 = i; i1 = -1; continue; }
        if (i1 < 0) {
            for (j = i; j < nframes - 1; j++) {
                if (binc[j] & 1) continue;
                i1 = j;
                break;
            }
            if (i1 < 0) {
                if (i0 >= 0) i1 = i0;
                else { status = -1; break; }
            }
        }
        if (i0 < 0 || i1 < 0) { status = -1; break; }
        if (i % step != start) continue;
        t0 = fexp[i0];
        {
            const float t1 = fexp[i1], ti = fexp[i];
            if (t0 == t1) t0 = 0;
            else t0 = (ti - t0) / (t1 - t0);
        }
        cth = cos(fdeg[i] * RV_DEG_TO_RAD);
        sth = sin(fdeg[i] * RV_DEG_TO_RAD);
        if (drEnd > drStart && (flags & 8)) {
            if (drStart < i && i < drEnd && drOmit) {
                for (k = 0; k < mult; k++) memset(frames + ((size_t)i * mult + k) * W, 0, sizeof(uint16_t) * W);
                continue;
            }
            if (i >= drEnd) idrift = (int)(drX * cth - drY * sth);
            else if (i > drStart) idrift = (int)((drX * cth - drY * sth) * (i - drStart) / (drEnd - drStart));
        }
        if (binc[i] & 4) {
            for (k = 0; k < mult; k++) memset(frames + ((size_t)i * mult + k) * W, 0, sizeof(uint16_t) * W);
            continue;
        }
        if ((flags & 4) && piDrift) idrift -= piDrift[i];
        sh = idrift / binning;
        for (k = 0; k < mult; k++) {
            const uint16_t* dark = frames + ((size_t)(nframes - 1) * mult + k) * W;
            const uint16_t* a0 = frames + ((size_t)i0 * mult + k) * W;
            const uint16_t* a1 = frames + ((size_t)i1 * mult + k) * W;
            uint16_t* row = frames + ((size_t)i * mult + k) * W;
            for (j = 0; j < W; j++) {
                /* unsigned short + float -> float arithmetic, stored as double */
                const float fa = (float)((int)a1[j] - (int)a0[j]) * t0;
                const float fb = (float)a0[j] + fa;
                inc[j] = (double)(fb - (float)dark[j]);
            }
            for (j = 0; j < W; j++) {
                int isample;
                if (inc[j] <= 0) { tmp[j] = 0; continue; }
                isample = (int)row[j] - (int)dark[j];
                if (isample < pixmin) { tmp[j] = 0; continue; }
                if (flags & 2) tmp[j] = (int16_t)(int)(inc[j] * 1000 / isample + 0.5);
                else tmp[j] = (int16_t)(int)(log(inc[j] / isample) * 3000 + 0.5);
            }
            /* drift shift, general.cpp:2501-2513 */
            for (j = 0; j < W; j++) {
                const int src = j - sh;
                row[j] = (uint16_t)((src >= 0 && src < W) ? tmp[src] : 0);
            }
        }
    }
    free(inc);
    free(tmp);
    return status;
}

void orc_output_stage(int32_t* img, int N, int z, double center, int bin, int* pimin, int* pimax)
{
    /* gazoDoc.cpp:3993-3999, 4007-4021 */
    const size_t n2 = (size_t)N * N;
    int imax = 0, imin = INT_MAX, irad2, i, j;
    const int ixdimh = N / 2;
    size_t q;
    for (q = 0; q < n2; q++) {
        imax = img[q] > imax ? img[q] : imax;
        imin = img[q] < imin ? img[q] : imin;
    }
    if (imin == imax) imax++;
    irad2 = ((z * center / bin) < (N - z * center / bin)) ? (int)(z * center / bin)
                                                          : (int)(N - z * center / bin);
    irad2 *= irad2;
    for (i = 0; i < N; i++) {
        int i2 = i - ixdimh;
        i2 *= i2;
        for (j = 0; j < N; j++) {
            const int j2 = j - ixdimh;
            const size_t idx = (size_t)i * N + j;
            if (i2 + j2 * j2 > irad2) img[idx] = 0;
            else img[idx] = (int)((float)(img[idx] - imin) * (65535 - 1) / (imax - imin));
        }
    }
    *pimin = imin; *pimax = imax;
}

/* ------------------------------------------------------ offset-CT stitch -- */

int orc_offset_stitch(const int16_t* sino, int nframes, int W, const float* fdeg,
                      const char* binc, double center, double deltac, int bin, int trim,
                      int out_width, int16_t* ofs)
{
    /* gazoDoc.cpp:2967-2978: first frame beyond 180 deg (nframes = iLenSinogr - 1) */
    int iRevIdx = -1, i, k;
    const int ixFrm = W;
    const int ixlen = W / bin;
    const int iNativeXdim = W * 2;
    for (i = 0; i < nframes; i++) {
        if (fdeg[i] > 180.) { iRevIdx = i; break; }
    }
    if (iRevIdx < 0) return -21051;
    {
        const int nout = iRevIdx;
        memset(ofs, 0, sizeof(int16_t) * (size_t)nout * out_width);
        for (i = 0; i < nout; i++) {
            const double th = fdeg[i];
            const int16_t* s1;
            const int16_t* s2;
            int16_t* o;
            if (!(binc[i] & 1)) continue;
            while ((fdeg[iRevIdx] - th < 180) | ((binc[iRevIdx] & 1) == 0)) {   /* :3566-3569 */
                iRevIdx++;
                if (iRevIdx >= nframes) return -21052;
            }
            s1 = sino + (size_t)i * ixlen;           /* rows are already binned: W / bin */
            s2 = sino + (size_t)iRevIdx * ixlen;
            o = ofs + (size_t)i * out_width;
            if (center > ixFrm / 2.) {                                           /* :3575-3616 */
                const double fc = (center + deltac + trim) / bin;
                const int iCent = (int)fc;
                const double dCent = fc - iCent;
                const int g = ixFrm / bin - 1 - iCent;
                short lastDens = 0;
                for (k = trim / bin; k <= iCent - g; k++) o[k - trim / bin] = s1[k];
                for (k = iCent - g + 1; k < (iNativeXdim - trim) / bin; k++) {
                    const int idx = iCent - (k - iCent);
                    if (dCent < 0.5) {
                        if (idx >= 0 && idx < ixlen - 1) {
                            const short ir0 = s2[idx + 1], ir1 = s2[idx];
                            lastDens = (short)(ir0 + (ir1 - ir0) * (1 - 2 * dCent));
                            if (k <= iCent + g) {
                                const double ratio = (double)(k - (iCent - g)) / (g * 2 + 1);
                                lastDens = (short)(ratio * lastDens + (1. - ratio) * s1[k]);
                            }
                        }
                    } else {
                        if (idx >= 1 && idx < ixlen) {
                            const short ir0 = s2[idx], ir1 = s2[idx - 1];
                            lastDens = (short)(ir1 + (ir0 - ir1) * (2 - 2 * dCent));
                            if (k <= iCent + g) {
                                const double ratio = (double)(k - (iCent - g)) / (g * 2 + 1);
                                lastDens = (short)(ratio * lastDens + (1. - ratio) * s1[k]);
                            }
                        }
                    }
                    o[k - trim / bin] = lastDens;
                }
            } else {                                                             /* :3617-3660 */
                const double fc = (center + deltac) / bin;
                const int iCent = (int)(fc + 0.5);
                const double dCent = fc - (int)fc;
                const int g = iCent;
                const int shift = (ixFrm - trim) / bin;
                short lastDens = 0;
                for (k = (ixFrm - 1 - trim) / bin; k >= iCent + g; k--) o[k + shift] = s1[k];
                if (dCent < 0.5) {
                    for (k = iCent + g - 1; k >= iCent * 2 + 1 - ixFrm / bin; k--) {
                        const int idx = iCent + ((iCent - 1) - k) + 1;
                        if (idx >= 0 && idx < ixlen - 1) {
                            const short ir1 = s2[idx], ir2 = s2[idx + 1];
                            lastDens = (short)(ir1 + (ir2 - ir1) * (2 * dCent));
                            if (k >= iCent - g) {
                                const double ratio = (double)((iCent + g) - k) / (g * 2 + 1);
                                lastDens = (short)(ratio * lastDens + (1. - ratio) * s1[k]);
                            }
                        }
                        o[k + shift] = lastDens;
                    }
                } else {
                    for (k = iCent + g - 1; k >= iCent * 2 - ixFrm / bin; k--) {
                        const int idx = iCent + ((iCent - 1) - k);
                        if (idx >= 0 && idx < ixlen - 1) {
                            const short ir0 = s2[idx], ir1 = s2[idx + 1];
                            lastDens = (short)(ir0 + (ir1 - ir0) * (2 * dCent - 1));
                            if (k >= iCent - g) {
                                const double ratio = (double)((iCent + g) - k) / (g * 2 + 1);
                                lastDens = (short)(ratio * lastDens + (1. - ratio) * s1[k]);
                            }
                        }
                        o[k + shift] = lastDens;
                    }
                }
            }
        }
        return nout;
    }
}
