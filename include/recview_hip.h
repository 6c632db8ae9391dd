/*
 * recview_hip.h -- C ABI of librecview_hip.so, the MI355X filtered
 * back-projection engine that drops into RecView's reconstruction path.
 *
 * Part 1 replaces cudaReconst.h + reconstinfo.h of the reference: the same
 * entry-point names, argument meanings and error behaviour, with CUDA types
 * replaced by HIP-neutral ones (no CUDA or HIP header is needed to include
 * this file).  Part 2 is the native batched API (no counterpart in the
 * reference) used for multi-slice, mixed-precision and zoomed work.
 *
 * Every Part-1 declaration cites the reference declaration it replaces.
 */
#ifndef RECVIEW_HIP_H
#define RECVIEW_HIP_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types -- */

/* cudaStream_t (reconstinfo.h:73) -> opaque handle (a hipStream_t inside). */
typedef struct rv_stream_opaque* rv_stream_t;
/* cufftHandle (reconstinfo.h:72) -> opaque; the engine has no FFT plan object. */
typedef void* rv_fft_plan_t;
/* float2 (reconstinfo.h:60) */
typedef struct { float x, y; } rv_float2;
/* cudaError_t (cudaReconst.h:58-59) -> int, 0 = success */
typedef int rv_error_t;
typedef int RV_BOOL;               /* BOOL of reconstinfo.h:41 */

/* Device-query error flags, cudaReconst.h:8-13 (same values). */
#define CUDA_ERROR_DEVICEINFO_MASK 0x0000ffff
#define CUDA_ERROR_INSUFFICIENT_DRIVER 0x00010000
#define CUDA_ERROR_INSUFFICIENT_COMPUTE_CAPABILITY 0x00020000
#define CUDA_ERROR_VIRTUAL_DEVICE_DETECTED 0x00040000
#define CUDA_ERROR_DEVICE_GETPROPERTY 0x00080000
#define CUDA_ERROR_DEVICE_GETCOUNT 0x00100000

/* constants.h:43-45 */
#define RECONST_INFO_IDLE 0
#define RECONST_INFO_BUSY 1
#define RECONST_INFO_ERROR 2
/* DlgQueue.h:23-26, StdAfx.h:29 */
#define RQFLAGS_ZERNIKE 2          /* DlgQueue.h:20-22 */
#define RQFLAGS_DRIFTLIST 4
#define RQFLAGS_DRIFTPARAMS 8
#define RQFLAGS_USEONLYEVENFRAMES 16
#define RQFLAGS_USEONLYODDFRAMES 32
#define RQFLAGS_SINOGRAMKEPT 128
#define CGAZODOC_BINC_SAMPLE 1
#define CGAZODOC_BINC_NOUSE 2      /* StdAfx.h:30-31 */
#define CGAZODOC_BINC_SKIP 4
#define READTIF8bit 2              /* general.h:45, RECONST_INFO.iFlag */

/* RECONST_INFO, reconstinfo.h:21-101: same field names and order, so that the
 * caller's assignments compile unchanged.  Device buffers are owned by the
 * library, grow only, and persist across calls until CudaReconstMemFree /
 * CudaReconstResourceFree.  The engine keeps the whole filtered sinogram
 * (d_px) instead of a 16x int LUT: d_igp stays NULL on the native path and is
 * only allocated by the CudaReconstHostFFT / kernel-wrapper compatibility
 * path.  d_strip holds all selected int16 rows of the slice. */
struct RECONST_INFO {
    unsigned char iStatus;
    int ixdim;
    int iInterpolation;
    double center;
    int iSinoCenter;
    int iLenSinogr;
    int* iReconst;
    int iStartSino;
    int iStepSino;
    unsigned int* pDoc;
    int* nSinogr;
    char dataName[61];
    char* bInc;
    float* fdeg;
    float* fexp;
    int iMultiplex;
    int iOffset;
    int maxSinogrLen;
    short** iSinogr;
    float* fFilter;
    unsigned int dReconFlags;
    float fTiltAngle;
    int drStart;
    int drEnd;
    double drX;
    double drY;
    RV_BOOL drOmit;
    unsigned int threadID;
    uintptr_t hThread;
    unsigned int max_d_ifp;
    unsigned int max_d_igp;
    int* d_ifp;
    int* d_igp;
    float* d_filt;
    unsigned int max_d_filt;
    short* d_strip;
    unsigned int max_d_strip;
    rv_float2* d_p;
    unsigned int max_d_p;
    float* d_px;
    unsigned int max_d_px;
    rv_fft_plan_t fftplan;
    unsigned int ifftdim;
    rv_stream_t stream1, stream2;
    uint64_t i64result;
    uint64_t i64sum;
    short** ppRef;
    short** ppQry;
    unsigned int uiMaxRef;
    unsigned int uiMaxQry;
    void* pcxyz1;
    void* pcxyz2;
    int iLsqBin;
    int iLsqBxref;
    int* piDrift;
    int iFlag;
};
typedef struct RECONST_INFO RECONST_INFO;

/* ------------------------------------------------- Part 1: host routines -- */

/* cudaReconst.h:16 (impl cudaReconstHost.cpp:629-813): reconstructs one slice
 * (int16 rows ri->iSinogr selected by bInc/fdeg) into the int32 image
 * ri->iReconst, with the reference's LUT16 integer semantics.  Sets
 * ri->iStatus = RECONST_INFO_ERROR on allocation failure; never throws.
 * A pinned iReconst (CUDA_MALLOC_HOST_INT, as RecView allocates it for the GPU
 * path) is written by the back-projection kernel itself, so the download
 * overlaps it; d_ifp then does not receive the image. */
void CudaReconstHost(RECONST_INFO* ri, int idev, bool bReport, bool bEnStream);
/* cudaReconst.h:17 (cudaReconstHost.cpp:142-203) */
void CudaReconstMemFree(RECONST_INFO* ri);
/* cudaReconst.h:18 (cudaReconstHost.cpp:205-214) */
void CudaReconstResourceFree(RECONST_INFO* ri, bool bCudaDeviceReset);
/* cudaReconst.h:19 (cudaReconstHost.cpp:221-291): the CUDAFFT-build variant:
 * zero padding, float zoom interpolation, LUT written per projection. */
void CudaReconstHostFFT(RECONST_INFO* ri, int idev, bool bReport);
/* cudaReconst.h:20-26 (cudaReconst.cu:449-538) */
int GetCudaDeviceCount(int iMinComputeCapability);
int GetCudaMaxThreadsPerBlock(int iDevice);
int GetCudaWarpSize(int iDevice);
int GetCudaDeviceName(int iDevice, char* pcName, int iszcName);
int GetCudaDeviceComputingCapability(int iDevice, int* piMajor, int* piMinor);
int GetCudaNumberOfCores(int iDevice, int* piCores, int* piProcessors);
int GetCudaClockRate(int iDevice, int* piClockRate, int* piMemRate);
/* cudaReconst.h:28: sinogram generation on device idev, in place on
 * ri->iSinogr (maxSinogrLen frames x iMultiplex rows of ixdim uint16 samples,
 * dark frame last), with the production semantics of GenerateSinogramThread
 * (general.cpp:2377-2532): incident interpolated in exposure time (fexp), bInc
 * SAMPLE / NOUSE / SKIP, iStartSino / iStepSino, drift parameters and list,
 * Zernike mode, 8- / 16-bit pixel floors.  Sets ri->iStatus to IDLE, or to
 * ERROR where the reference does (no incident frame). */
void CudaSinogramHost(RECONST_INFO* ri, int idev, bool bReport);
/* The same conversion on device-resident frames (no counterpart in the
 * reference): d_frames = uint16 [nframes][mult][W] (dark frame last); binc,
 * fexp, fdeg (and piDrift, may be NULL) are host arrays of nframes entries.
 * Converted rows are written to d_out in the frames' layout (row i*mult + k),
 * as int16 (x3000, the reference's values) or, with out_float = 1, as float
 * line integrals -ln(I/I0) without scaling or rounding, ready for an rvPlan.
 * Other rows of d_out are not written.  Returns the number of converted rows,
 * -1 on bad arguments / device errors, -2 where the reference sets
 * RECONST_INFO_ERROR (no incident frame; rows before it are converted). */
int RecviewSinogramDevice(const unsigned short* d_frames, int nframes, int mult, int W, const char* binc,
                          const float* fexp, const float* fdeg, unsigned flags, int drStart, int drEnd,
                          double drX, double drY, int drOmit, const int* piDrift, int is8bit, int start,
                          int step, int binning, int out_float, void* d_out, rv_stream_t stream);

/* ------------------------------------------ Part 1: kernel wrappers -- */

/* cudaReconst.h:31-32 (cudaReconst.cu:158-192): filters one int16 strip
 * (CUDAFFT semantics) and writes its 16N-entry int LUT to d_igp.  d_p
 * (>= ndim rv_float2) is scratch; fftplan is ignored (may be NULL). */
void CudaDeconv(int ixdim, int iIntpDim, int ndim, float center, float* d_filt,
                short* d_strip, int* d_igp, rv_float2* d_p, rv_fft_plan_t* fftplan);
/* cudaReconst.h:33-34 (cudaReconst.cu:195-228): adds one projection's LUT
 * d_igp[16*ixdimp] into the int32 image d_ifp[ixdimp^2]. */
void CudaBackProjStream(int ixdimp, float center, int iCenterOffset, int iIntpDim, double theta,
                        int* d_ifp, int* d_igp, rv_stream_t stream);
/* cudaReconst.h:35 (cudaReconst.cu:230-246): LUT build from a filtered row. */
void CudaSinoPx2igpStream(int ixdimp, int* d_igp, float* d_px, rv_stream_t stream);

/* cudaReconst.h:58-59 (cudaReconstHost.cpp:216-218): pinned host memory. */
rv_error_t CUDA_MALLOC_HOST_INT(int** ptr, size_t size);
rv_error_t CUDA_FREE_HOST(void* ptr);
/* Device memory and synchronous copies for hosts that include only this
 * header (no counterpart in the reference, which calls the CUDA runtime
 * directly): hipMalloc / hipFree / hipMemcpy on the current device. */
rv_error_t RecviewDeviceAlloc(void** ptr, size_t bytes);
rv_error_t RecviewDeviceFree(void* ptr);
rv_error_t RecviewCopyToHost(void* dst, const void* src, size_t bytes);
rv_error_t RecviewCopyToDevice(void* dst, const void* src, size_t bytes);

/* cudaReconst.cu:19: block size of the compatibility kernels, written by the
 * application (gazo.cpp:549-550).  The tiled kernels use their own shapes. */
extern int blocksize;

/* cudaReconst.h:40: progress/abort callback imported from the application
 * (general.cpp:2352).  Resolved as a weak symbol; a host without it (Python,
 * tests) installs one with RecviewSetDlgCtrl.  Returning true aborts. */
typedef bool (*rv_dlgctrl_fn)(RECONST_INFO* ri, int iProgStep, int iSino, int* pCurrStep);
void RecviewSetDlgCtrl(rv_dlgctrl_fn fn);

/* Error ring (replaces CError, cerror.cpp:35-43): codes 288xx / 300xx as the
 * reference logs them.  Returns the number copied (most recent last). */
int RecviewGetErrors(int* codes, int maxn);
void RecviewClearErrors(void);
/* Last HIP/library error message (thread-local). */
const char* RecviewLastMessage(void);

/* Options of the Part-1 host routines (no counterpart: the reference fixes
 * them at compile time).  lut_formula 1 = cudaReconst.cu:237-239 (default),
 * 0 = general.cpp:2131-2136.  index_rule 1 = trunc then range check
 * (cudaReconst.cu:36-38, default), 0 = float mask (projx64.asm:261-266). */
void RecviewSetCompatOptions(int lut_formula, int index_rule);

/* ---------------------------------------------- Part 2: native batched API -- */

#define RV_INTERP_LINEAR 0     /* production: true linear detector interpolation, FP32 */
#define RV_INTERP_LUT16_F32 1  /* reference 1/16 sub-sample LUT, float values/accumulation */
#define RV_INTERP_LUT16_I32 2  /* reference LUT16 with x3000 int16 input, x100 int LUT, int32 sum */

#define RV_STORE_F32 0         /* filtered sinogram kept in FP32 */
#define RV_STORE_F16 1         /* FP16 with per-row power-of-two scale */
#define RV_STORE_BF16 2        /* BF16 with per-row power-of-two scale */

#define RV_SINO_I16 0          /* input rows int16 (x LOG_SCALE) */
#define RV_SINO_F32 1          /* input rows float line integrals */

#define RV_PAD_REPLICATE 0     /* host path: edge replicate, placement n/2-1 */
#define RV_PAD_ZERO 1          /* CUDAFFT path: zeros, placement n/2 */

typedef struct {
    int W;               /* detector width (after binning / stitching) */
    int P;               /* number of projections (rows) */
    int iInterpolation;  /* 2 = none, 3/4/5 = zoom x2/x4/x8 of the whole FOV (reference) */
    double center;       /* RECONST_INFO.center semantics: (c_user + 0.5) / bin */
    float tilt;          /* degrees added to each angle */
    const float* fdeg;   /* host array [P] of angles in degrees */
    int filter;          /* 0 Hann 1 Hamming 2 Ramp 3 Parzen 4 Butterworth 5 Shepp-Logan */
    double cutoff;       /* filter cutoff (0.5 default) */
    double order;        /* Butterworth order */
    int sino_dtype;      /* RV_SINO_I16 / RV_SINO_F32 */
    int pad_mode;        /* RV_PAD_REPLICATE / RV_PAD_ZERO */
    /* Output grid.  out_n = 0: the reference grid (N = W*zoom square).
     * Otherwise an out_n x out_n grid of pitch 1/pix_zoom detector samples
     * whose pixel (0,0) centre sits at (x0, y0) of the reference z=1 grid. */
    int out_n;
    double pix_zoom;
    double x0, y0;
    /* Tile configuration of this slice: -1 = the plan's (RvOptions.tile);
     * 0..4 = 64x64, 64x128, 32x32, 128x64, 32x64 (chosen per shape by an autotuner).
     * Slices with different tiles run as separate launches of one execution. */
    int tile;
    /* 1: the input is a 360-degree scan of P frames with the axis off centre
     * (fdeg over 0..360, optional frame flags binc); the plan stitches it on the
     * GPU to P' rows of 2W samples (RecviewOffsetCTStitch) and reconstructs
     * those, with the centre remap of gazoDoc.cpp:3832-3837. */
    int offset_ct;
    const char* binc;
    /* Angles per LDS chunk of this slice's back-projection (an autotuned value):
     * 0 = the plan's (RvOptions.angles_per_chunk, else the mode's default).  The
     * slices of one launch (same tile and mode) take the first one's. */
    int chunk;
    /* 0: the slice filters its own sinogram.  s + 1: the slice back-projects the
     * filtered rows of the earlier slice s with its own geometry (same detector,
     * zoom, projections and filter; no offset CT): the integer part of the centre
     * difference shifts the rows as the reference's kept-sinogram iCenterOffset
     * does (cudaReconstHost.cpp:775-777), so a centre sweep filters once.  The
     * slice's sinogram pointer is then not read. */
    int rows_of;
} RvSliceSpec;

typedef struct {
    int interp;          /* RV_INTERP_* */
    int store;           /* RV_STORE_* (LINEAR only) */
    int lut_formula;     /* LUT16 modes: 0 CPU formula, 1 CUDA formula */
    int index_rule;      /* LUT16 modes: 0 float mask (floor), 1 trunc */
    int persistent;      /* 1: persistent work-queue back-projection */
    int tile;            /* 0 = autotuned, else tile config id */
    int angles_per_chunk;/* 0 = default */
    int lds_half;        /* FP16/BF16 stores: 0 = FP32 pairs in LDS (default), 1 = FP16 pairs */
    int normalize;       /* float modes: 1 multiplies each image by 2*pi / (projections used), i.e.
                          * attenuation per detector sample (the caller-side scale of
                          * gazoDoc.cpp:4000-4001 without the pixel-width factor); 0 = raw sums */
    int exact_filter;    /* LINEAR modes: 1 keeps CFft's operation order (filtered rows bit-identical to
                          * the reference CPU filter); 0 (default) filters two rows per transform with
                          * the symmetrised filter, the same operator at half the cost */
    int slices_per_item; /* LINEAR modes: slices of identical geometry back-projected together, sharing
                          * each pixel's coordinate and LDS reads: 0 (default) = four per work item
                          * with FP32 values in LDS (results identical to one per item; a class's
                          * leftovers as a pair and a single), sixteen with FP16 pairs (lds_half:
                          * one FP16 dot product per update and one power-of-two scale per work
                          * item; eight on the 64x64 tile, which has no sixteen-slice kernel); or
                          * 1, 2, 4, 8, 16, 32 (at most 4 with FP32 values in LDS; 8 on the 64x64,
                          * 32x32 and 32x64 tiles, 16 on the 32x32 and 32x64 tiles, else fewer;
                          * 32, opt-in: the matrix-core kernel, which
                          * runs 32x32 tiles whatever the slice's tile, for every 32 slices of one
                          * geometry, the rest of a class as FP16 pairs).  With FP16 pairs and more
                          * than two per item, the pairs and singles left over by a class run the
                          * FP32-pair kernels on the 16-bit rows (faster for them) */
} RvOptions;

typedef struct rv_plan_opaque* rv_plan_t;

/* Builds a plan for nslices slices on device `device`: geometry, filter
 * tables, angle tables, filtered-sinogram workspace, the packed rows of
 * multi-slice work items (FP32: 16 B per sample and slice pair; FP16 pairs:
 * 4 B per sample and slice) and the tile work list.
 * Returns 0 or a negative error (message in RecviewLastMessage). */
int rvPlanCreate(rv_plan_t* plan, int device, const RvSliceSpec* specs, int nslices,
                 const RvOptions* opt);
/* Reconstructs every slice: d_sino[s] points to device rows [P][W] of the
 * slice's sino_dtype, d_out[s] to a device image (float, or int32 in the
 * LUT16_I32 mode) of out_n^2 (or N^2) elements.  Asynchronous on `stream`. */
int rvPlanExecute(rv_plan_t plan, const void* const* d_sino, void* const* d_out, rv_stream_t stream);
/* Captures rvPlanExecute with these pointers into a hipGraph once and
 * replays it on every later call with the same pointers. */
int rvPlanExecuteGraph(rv_plan_t plan, const void* const* d_sino, void* const* d_out, rv_stream_t stream);
/* Stage-level access for tests and benchmarks.  rvPlanBackproject packs the
 * grouped slices' filtered rows and back-projects them; rvPlanPack and
 * rvPlanBackprojectPacked are its two halves (the second reads the rows the
 * last pack wrote, so that the back-projection kernels can be timed alone). */
int rvPlanFilter(rv_plan_t plan, const void* const* d_sino, rv_stream_t stream);
int rvPlanBackproject(rv_plan_t plan, void* const* d_out, rv_stream_t stream);
int rvPlanPack(rv_plan_t plan, rv_stream_t stream);
int rvPlanBackprojectPacked(rv_plan_t plan, void* const* d_out, rv_stream_t stream);
/* Device pointer of slice s's filtered rows (FP32 [P][N], or the FP16/BF16
 * rows with their per-row scales), and their row length. */
int rvPlanFilteredRows(rv_plan_t plan, int s, void** rows, float** scales, int* row_len);
/* Uploads filtered FP32 rows [P][row_len] for slice s (BP-only testing). */
int rvPlanSetFilteredRows(rv_plan_t plan, int s, const float* d_rows, rv_stream_t stream);
/* Plan facts: image side, transform length, number of tiles, tile shape. */
int rvPlanInfo(rv_plan_t plan, int s, int* out_n, int* n_fft, int* tiles, int* tile_w, int* tile_h);
/* Rows and width that slice s reconstructs (after offset-CT stitching). */
int rvPlanRows(rv_plan_t plan, int s, int* nrows, int* width);
void rvPlanDestroy(rv_plan_t plan);

/* Offset CT (gazoDoc.cpp:3536-3665, 2967-2978).  RecviewOffsetCTPairs pairs
 * every sampled frame below 180 deg with its partner >= 180 deg further (host);
 * returns the number of stitched rows or -21051 / -21052 as the reference.
 * RecviewOffsetCTStitch writes the double-width rows [nout][out_width] on the
 * device (int16 rows with the reference's truncations, or float rows);
 * center = user centre + 0.5, deltac = deltaCent * j * binning. */
int RecviewOffsetCTPairs(const float* fdeg, const char* binc, int nframes, int* pair);
int RecviewOffsetCTStitch(const void* d_sino, int is_float, int W, const int* d_pair, int nout,
                          double center, double deltac, int bin, int trim, int out_width,
                          void* d_out, rv_stream_t stream);

/* ------------------------------------------------------------------------
 * Part 3: learned restoration of sparse-view slices (no counterpart in the
 * reference).  A residual U-Net (levels, base channels; 4 and 32 by default)
 * maps an FBP slice reconstructed from 1/8 of the angles to an estimate of the
 * full-angle FBP slice.  3x3 convolutions run on BF16 MFMA with FP32
 * accumulation; parameters are one FP32 vector (FP32 master weights, Adam).
 * Image pointers are device float32 [B][H][W]; H and W must be divisible by
 * 2^(levels-1).  `scale` multiplies the inputs before the network (outputs are
 * divided by it).  Entry points return 0 on success, -1 on error
 * (RecviewLastMessage() says why).
 * ------------------------------------------------------------------------ */
typedef struct rv_unet_opaque* rv_unet_t;
int rvUnetCreate(rv_unet_t* net, int device, int base_channels, int levels, unsigned long long seed);
long long rvUnetParamCount(rv_unet_t net);
int rvUnetGetParams(rv_unet_t net, float* host, long long n);
int rvUnetSetParams(rv_unet_t net, const float* host, long long n);
/* out = in + net(in * scale) / scale */
int rvUnetForward(rv_unet_t net, const float* d_in, float* d_out, int B, int H, int W, float scale,
                  rv_stream_t stream);
/* Mean squared error of the normalised output against target * scale and its
 * gradient w.r.t. every parameter (copied to host_grads when non-null; loss
 * when non-null, which synchronises the stream). */
int rvUnetLossGrad(rv_unet_t net, const float* d_in, const float* d_target, int B, int H, int W,
                   float scale, float* loss, float* host_grads, rv_stream_t stream);
/* rvUnetLossGrad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8). */
int rvUnetTrainStep(rv_unet_t net, const float* d_in, const float* d_target, int B, int H, int W,
                    float scale, float lr, float* loss, rv_stream_t stream);
void rvUnetDestroy(rv_unet_t net);
/* Single 3x3 convolution on NHWC bf16 tensors (uint16 storage), FP32 weights
 * W[Cout][3][3][Cin] packed to bf16 fragments per call.  data_gradient = 1
 * computes dX (Cin channels, into d_out) from dY (Cout channels, d_in) with the
 * same W; d_mask zeroes outputs whose mask value is not > 0. */
int rvConv3x3Bf16(const unsigned short* d_in, int in_stride, int in_off, int Cin, const float* d_w,
                  const float* d_bias, unsigned short* d_out, int out_stride, int out_off, int Cout,
                  const unsigned short* d_mask, int mask_stride, int mask_off, int relu,
                  int data_gradient, int B, int H, int W, rv_stream_t stream);
/* Weight (FP32 [Cout][3][3][Cin]) and bias (may be null) gradients of that
 * convolution from dY and X. */
int rvConv3x3WgradBf16(const unsigned short* d_dy, int dy_stride, int dy_off, int Cout,
                       const unsigned short* d_x, int x_stride, int x_off, int Cin, float* d_dw,
                       float* d_db, int B, int H, int W, rv_stream_t stream);

/* ------------------------------------------------------------------------
 * Output side of a slice (gazoDoc.cpp:3993-4021, 1986-1989; general.cpp:1313-1470).
 * ------------------------------------------------------------------------ */
/* Caller-side normalisation on the device: min / max over the N x N image
 * (int32, or float with is_float = 1; the maximum starts at 0 as in the
 * reference), then (int)((float)(v - min) * 65534 / (max - min)) inside the
 * circle of radius min(z*c/bin, N - z*c/bin) and 0 outside, written as uint16.
 * center is the value the caller passes at gazoDoc.cpp:4008; zoom and binning
 * follow iInterpolation.  The minimum and maximum are returned (max bumped by
 * one when equal to min). */
int RecviewOutputStage(const void* d_img, int is_float, int N, double center, int iInterpolation,
                       unsigned short* d_out, double* vmin, double* vmax, rv_stream_t stream);
/* RecviewOutputStage without any allocation, for callers that cannot allocate
 * device memory (code under hipGraph stream capture, serving loops, harnesses
 * that forbid hipMalloc); it has no counterpart in the reference.  The same
 * kernels and results, asynchronous on `stream`: the (min, max) pair is written
 * to the caller's device buffer d_range (two values of the image type) instead
 * of being copied to the host.  The per-block partial minima and maxima live in
 * a library-held 8 KiB device array per GPU (held until the process ends), so
 * calls that overlap on one GPU race; the partials are capped at 1024, which
 * serves every image side up to the limit N <= 65536 (larger N: error). */
#define RV_OUTPUT_NOALLOC_MAX_SIDE 65536
int RecviewOutputStageNoAlloc(const void* d_img, int is_float, int N, double center, int iInterpolation,
                              unsigned short* d_out, void* d_range, rv_stream_t stream);
/* The ImageDescription string RecView stores in every slice file
 * (gazoDoc.cpp:1986-1988): sinogram count, pixel width, pixelDiv, pixelBase,
 * centre, filter id + 1 and two constants.  Returns the length (snprintf). */
int RecviewImageDescription(char* buf, int bufsize, int numSinogr, double pixelWidth, double pixelDiv,
                            double pixelBase, double center, int ifilter);
/* WriteTifMonochrome16 (general.cpp:1313-1470): the same bytes, from a host
 * int32 image (negative values written as 0).  desc / artist may be NULL (the
 * reference's defaults).  Returns 0, or 18100 / 18102 as the reference. */
int RecviewWriteTif16(const char* path, const int* buffer, int height, int width, const char* desc,
                      const char* artist);

/* The filter design of gazoDoc.cpp:3747-3812 (plus 5 = Shepp-Logan, new). */
void RecviewSetFilter(int ndim, int ifilter, double cutoff, double order, float* out);
/* Geometry of cudaReconstHost.cpp:635-658: returns n (ndim) and writes N, zoom. */
int RecviewGeometry(int ixdim, int iInterpolation, int* N, int* zoom);

#ifdef __cplusplus
}
#endif
#endif
