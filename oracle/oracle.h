/*
 * oracle.h -- CPU restatement of RecView's filtered back-projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product path (recview_amd, librecview_hip.so) never links or calls it.
 *
 * Three reconstructions are restated here (names used throughout the tests):
 *   R-int : reference semantics end to end -- int16 sinogram, CFft-exact FP32
 *           filter, int32 LUT with 16 sub-samples, int32 accumulation.
 *   R-f32 : the same with a float sinogram and a float LUT (no x3000 / x100
 *           truncation), accumulated in double.
 *   O-f64 : an FP64 FBP: double FFT convolution with the same filter, true
 *           linear detector interpolation, double accumulation.
 * Every function cites the reference file:line it follows.
 */
#ifndef RV_ORACLE_H
#define RV_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Geometry derived from (ixdim, iInterpolation, center) exactly as
 * cudaReconstHost.cpp:635-658 and general.cpp:2021-2035. */
typedef struct {
    int ixdim;        /* W: detector width after binning */
    int izoom;        /* zoom exponent: 0..3 */
    int z;            /* iIntpDim = 2^izoom */
    int ndimp;        /* log2(n) */
    int ndim;         /* n: FFT length */
    int N;            /* ixdimp = W*z, image is N x N */
    int h;            /* ixdimh = N/2 */
    int ic;           /* intcenter = (int)center */
    float fcenter16;  /* fcenter in 1/16-sample units */
    int ihoffset;     /* crop offset n/2-1-h (host path) */
} orc_geom;

int  orc_geometry(int ixdim, int iInterpolation, double center, int float_center, orc_geom* g);

/* Filter design, gazoDoc.cpp:3747-3812.  ifilter: 0 Hann, 1 Hamming, 2 Ramp,
 * 3 Parzen, 4 Butterworth, other = all ones (DlgReconOpt.h:4-8). */
void orc_set_filter(int ndim, int ifilter, double cutoff, double order, float* out);

/* CFft (cfft.cpp:374-434, 193-369): twiddles of Init1(np,-1), FFT1Rev, FFT1.
 * Arrays are split re/im, length n; scratch needs 4n floats. */
void orc_cfft_twiddles(int np, float* wr, float* wi);
void orc_fft1rev(int np, const float* wr, const float* wi, float* re, float* im, float* scratch);
void orc_fft1(int np, const float* wr, const float* wi, float* re, float* im, float* scratch);

/* Row preparation into the FFT input (real part), cudaReconstHost.cpp:733-751. */
void orc_row_prep_i16(const int16_t* strip, const orc_geom* g, float* re);
/* The same placement and replicate padding for a float strip; zoom positions use
 * float interpolation (s[k]*(z-j) + s[k+1]*j)/z. */
void orc_row_prep_f32(const float* strip, const orc_geom* g, float* re);
/* CUDAFFT variant, cudaReconst.cu:123-156, 174: zero padding, float zoom, placement n/2. */
void orc_row_prep_cudafft(const int16_t* strip, const orc_geom* g, float* re);

/* Filter nrow rows end to end (prep, FFT1Rev, x filter, FFT1, crop), producing
 * px[nrow][N] (cudaReconstHost.cpp:731-756).  kind: 0 = int16 input host path,
 * 1 = float input host path, 2 = int16 input CUDAFFT placement. */
void orc_filter_rows(const void* sino, int kind, int nrow, int row_stride,
                     const orc_geom* g, const float* filt, float* px);

/* 16-sub-sample LUT.  formula 0: CPU (general.cpp:2130-2137),
 * formula 1: CUDA px2igp (cudaReconst.cu:230-240).  Int32, x BACKPROJ_SCALE. */
void orc_lut16_i32(const float* px, int N, int formula, int32_t* igp);
/* Float LUT for R-f32: p0 = px[j], slope (px[j+1]-p0)/16, value p0 + slope*s. */
void orc_lut16_f32(const float* px, int N, float* lut);

/* Per-projection BP parameters, cudaReconst.cu:216-222 / general.cpp:2139-2142. */
void orc_bp_params(float fdeg, float tilt, const orc_geom* g, int iCenterOffset,
                   float* fcos, float* fsin, float* foffset);

/* LUT back-projection of one projection over image rows [iy0, iy1).
 * rule 0: float mask 0 <= t < 16N (projx64.asm:258-270);
 * rule 1: trunc then integer range check (general.cpp:2152-2156, cudaReconst.cu:36-38). */
void orc_bp_lut_i32(int32_t* img, const int32_t* lut, int N, float fcos, float fsin,
                    float foffset, int rule, int iy0, int iy1);
void orc_bp_lut_f64(double* img, const float* lut, int N, float fcos, float fsin,
                    float foffset, int rule);

/* Whole-slice R-int / R-f32 reconstructions (selection of all rows, nSinogr = P). */
void orc_recon_rint(const int16_t* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, int formula, int rule,
                    int32_t* img);
void orc_recon_rf32(const float* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, int rule, double* img);

/* O-f64 helpers. */
void orc_fft_f64(int n, double* re, double* im, int inverse);
/* Double filter of float rows (host placement and replicate padding, crop). */
void orc_filter_rows_f64(const float* sino, int nrow, int row_stride, const orc_geom* g,
                         const float* filt, double* px);
/* Linear-interpolation BP in double.  Detector coordinate of pixel (ix, iy):
 *   u = (ix - h) cos(th) - (iy - h) sin(th) + h + frac*z + iCenterOffset
 * value 0 outside [0, N); p[k] + (u-k)(p[k+1]-p[k]) on [0, N-1); p[N-1] on [N-1, N). */
void orc_bp_linear_f64(double* img, const double* px, int N, double th, double fc_minus_h,
                       int y0, int y1);
void orc_recon_of64(const float* sino, int P, const float* fdeg, float tilt,
                    const orc_geom* g, const float* filt, double* img);
/* Linear BP of already filtered float rows in double (used to isolate the BP stage). */
void orc_bp_rows_linear_f64(const float* px, int P, const float* fdeg, float tilt,
                            const orc_geom* g, int iCenterOffset, double* img);

/* FP64 linear-interpolation back-projection of FP64 rows px[P][N] at arbitrary
 * points (x, y) of the N x N reconstruction grid (fractional for zoomed
 * regions): out[q] = sum_i lerp(px_i, u_i), u_i = (x - h) cos th_i - (y - h) sin th_i
 * + fcenter16/16 + iCenterOffset, the same detector coordinate and edge rules
 * as orc_bp_linear_f64 (which it equals on integer points). */
void orc_bp_points_linear_f64(const double* px, int P, const float* fdeg, float tilt,
                              const orc_geom* g, int iCenterOffset, const double* xs,
                              const double* ys, int npts, double* out);

/* Sinogram generation, the production body of GenerateSinogramThread
 * (general.cpp:2377-2532): frames are uint16 rows [nframes][mult][W] with the
 * dark frame last; every sample frame i (< nframes-1) taken by
 * i % step == start becomes int16 -ln(I/I0)*3000 (or the Zernike ratio), with
 * I0 interpolated in exposure time between the enclosing incident frames, the
 * pixel-count floor, SKIP / drift-omit zeroing and the drift shift, in place.
 * Returns 0, or -1 where the reference sets RECONST_INFO_ERROR. */
int orc_sinogram(uint16_t* frames, int nframes, int mult, int W, const char* binc,
                 const float* fexp, const float* fdeg, unsigned flags, int drStart, int drEnd,
                 double drX, double drY, int drOmit, const int* piDrift, int is8bit, int start,
                 int step, int binning);

/* Caller-side output stage, gazoDoc.cpp:3993-4021: min/max over the square,
 * normalise to 0..65534 inside the circle of radius min(z*c, N-z*c). */
void orc_output_stage(int32_t* img, int N, int z, double center_user_plus_half, int bin,
                      int* imin, int* imax);

/* Offset-CT stitch (gazoDoc.cpp:3536-3665) for one multiplex row: pairs every
 * sampled frame i < iRevIdx0 (the first frame beyond 180 deg, gazoDoc.cpp:2967-2978)
 * with the first later sampled frame at >= 180 deg further, and writes the
 * double-width row of frame i into ofs[i] (rows not written stay zero).
 * center = user centre + 0.5 (gazoDoc.cpp:3560), deltac = deltaCent*j*bin.
 * Returns the number of output rows, or a negative error (21051/21052). */
int orc_offset_stitch(const int16_t* sino, int nframes, int W, const float* fdeg,
                      const char* binc, double center, double deltac, int bin, int trim,
                      int out_width, int16_t* ofs);

#ifdef __cplusplus
}
#endif
#endif
