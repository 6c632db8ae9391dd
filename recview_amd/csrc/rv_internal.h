// rv_internal.h -- structures shared by the host orchestration (rv_host.cpp,
// rv_plan.cpp) and the CDNA4 kernels (rv_filter.hip, rv_bp.hip).  Plain C++,
// no HIP types in the layouts, so host and device agree on every field.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rv {

// ------------------------------------------------------------- filter stage
// One work item = one sinogram row.  Row prep follows cudaReconstHost.cpp:733-751
// (pad_mode 0) or cudaReconst.cu:123-156 (pad_mode 1); the transform is the
// CFft FFT1Rev / filter / FFT1 sequence (cfft.cpp:193-434), crop at `crop`.
struct FilterJob {
    const void* sino;      // input rows (int16 or float)
    long long in_stride;   // elements between input rows
    const int* row_in;     // [nrows] input row index per work row (nullptr: identity)
    const int* row_out;    // [nrows] output row index per work row (nullptr: identity)
    int nrows;
    int in_kind;           // 0 int16, 1 float
    int W, z, ic, n, np, N;
    int place;             // padded index of sample ic: n/2-1 (host) or n/2 (CUDAFFT)
    int pad_mode;          // 0 replicate + integer zoom division, 1 zero + float zoom
    int crop;              // first padded index of the N cropped outputs
    const float* filt;     // [n]
    const float* tw;       // [n] interleaved (re, im) twiddles, n/2 entries
    float* out;            // [rows][out_stride] FP32 rows (nullptr when store != 0)
    long long out_stride;
    int store;             // 0 f32, 1 f16, 2 bf16
    void* out_lp;          // [rows][out_stride] 16-bit rows
    float* out_scale;      // [rows] per-row scale (power of two)
    int pair;              // 1: two rows per transform (filter_pair_kernel): filt = symmetrised
                           //    filter, tw = full table exp(-2 pi i k / n), k < n
};

// Both return hipErrorInvalidValue for a job they cannot run (no launch).
hipError_t launch_filter(const FilterJob& job, hipStream_t stream);
// Several slices of identical filter geometry in one launch (grid y = slice):
// the per-slice pointers come from the batch, every other field from the job.
constexpr int kFilterBatch = 16;
struct FilterBatch {
    const void* sino[kFilterBatch];
    float* out[kFilterBatch];
    void* out_lp[kFilterBatch];
    float* out_scale[kFilterBatch];
};
hipError_t launch_filter_batch(const FilterJob& job, const FilterBatch& batch, int count, hipStream_t stream);
void prepare_filter(int n);   // kernel attributes for an n-point transform (not capturable)
// Transforms longer than 2^kMaxLdsFftLog2 points (the complex row no longer fits
// in LDS): CFft's network in two passes per transform through a work buffer of
// 2 * work_rows * n float2 (rows processed work_rows at a time), bit-identical
// to the one-pass network.  job.pair must be 0; up to 2^kMaxFftLog2 points.
constexpr int kMaxLdsFftLog2 = 14;
constexpr int kMaxFftLog2 = 20;
hipError_t launch_filter_large(const FilterJob& job, void* work, size_t work_rows, hipStream_t stream);
// Rows per batch that fit a work buffer of budget_bytes (at most nrows, at least 1).
size_t filter_large_rows(int n, size_t budget_bytes, int nrows);

// LUT build (cudaReconst.cu:230-246) and single-projection LUT back-projection
// (cudaReconst.cu:31-39, 195-228) for the kernel-wrapper entry points.
void launch_px2igp(const float* d_px, int* d_igp, int N, int formula, float scale, int block,
                   hipStream_t stream);
void launch_bp_lut_global(int* d_ifp, const int* d_igp, int N, float fcos, float fsin,
                          float foffset, int rule, int block, hipStream_t stream);

// ------------------------------------------------------- sinogram generation
// One converted row: frame `frame`, multiplex line k, incident frames i0 / i1
// with weight t0, drift shift (samples), action 1 = convert, 2 = zero.
struct SinoJob {
    int frame, k, i0, i1;
    float t0;
    int shift, action, pad;
};
void launch_sinogram(const uint16_t* frames, int nframes, int mult, int W, const SinoJob* jobs,
                     int njobs, void* out, int pixmin, int zernike, int out_float, int out_full,
                     hipStream_t stream);
// The sequential frame walk of GenerateSinogramThread (general.cpp:2394-2466):
// incident frames, exposure weights, drift shifts, SKIP / drift-omit zeroing.
// Returns false where the reference sets RECONST_INFO_ERROR (jobs before the
// failing frame are kept, as the reference converts them).
struct SinoWalk {
    int nframes, mult;          // frames include the dark frame (last)
    const char* binc;
    const float* fexp;
    const float* fdeg;
    unsigned flags;             // RQFLAGS_*
    int drStart, drEnd, drOmit;
    double drX, drY;
    const int* piDrift;
    int start, step, binning;
};
bool sinogram_jobs(const SinoWalk& w, std::vector<SinoJob>& jobs);

// ------------------------------------------------------- output stage
// Min/max + normalisation to 0..65534 inside the circle of squared radius
// irad2 (gazoDoc.cpp:3993-4021).  scratch: >= 2050 elements of the image type;
// its last two hold (min, max) afterwards.
void launch_output_stage(const void* img, int is_float, int N, int irad2, uint16_t* out, void* scratch,
                         hipStream_t s);
// The same three launches with the 1024 partial minima and 1024 partial maxima at
// part (2048 values of the image type) and the (min, max) result at range.
void launch_output_stage_parts(const void* img, int is_float, int N, int irad2, uint16_t* out, void* part,
                               void* range, hipStream_t s);
constexpr int kOutputParts = 1024;
// Squared mask radius of gazoDoc.cpp:4008-4010: r = zoom * center / binning,
// min(r, N - r) truncated, for the zoom and binning of iInterpolation.
inline int output_irad2(int N, double center, int iInterpolation)
{
    const int ibin = iInterpolation == 0 ? 4 : (iInterpolation == 1 ? 2 : 1);
    const int ipintp = iInterpolation > 2 ? 1 << (iInterpolation - 2) : 1;
    const double r = ipintp * center / ibin;
    const int ir = r < (N - r) ? (int)r : (int)(N - r);
    return ir * ir;
}

// ------------------------------------------------------- back-projection
// BP_LINEAR2: BP_LINEAR over two slices of identical geometry per work item
// (BP_LINEAR4: four);
// BP_LINEAR_H2 / H4 / H8 / H16: BP_LINEAR_H over two / four / eight / sixteen;
// BP_LINEAR_M32: 32 FP16 slices per work item, the updates done as matrix
// products on the MFMA units (the others in BpSlice::rows_x / scales_x / out_x).
enum BpMode {
    BP_LINEAR = 0, BP_LUT_F32 = 1, BP_LUT_I32 = 2, BP_LINEAR_H = 3, BP_LINEAR2 = 4,
    BP_LINEAR_H2 = 5, BP_LINEAR_H4 = 6, BP_LINEAR_H8 = 7, BP_LINEAR_H16 = 8, BP_LINEAR_M32 = 9,
    BP_LINEAR4 = 10
};
constexpr int bp_slices_per_item(int mode)
{
    return (mode == 4 || mode == 5) ? 2 : (mode == 6 || mode == 10) ? 4 : mode == 7 ? 8 : mode == 8 ? 16
         : mode == 9 ? 32 : 1;
}
constexpr int BP_MAX_SLICES_PER_ITEM = 32;
// BP_LINEAR_M32 packed rows: per angle row, (N + 7) / 8 chunks of eight samples,
// each chunk the 32 slices' samples side by side (16 B per slice)
constexpr long long bp_m32_chunks(int N) { return (N + 7) / 8; }

// Per-angle table entry.  LINEAR: detector coordinate of pixel (ix, iy) of the
// output grid is u = u0 + ix*dux + iy*duy (all double).  LUT16: the reference
// float triple t = ix*fcos + (iy*fsin + foffset) in 1/16-sample units.
struct AngleEntry {
    double u0, dux, duy;   // LINEAR geometry (FP64, per angle)
    float fcos, fsin, foffset;
    int row;               // row index into the slice's filtered rows
};

struct BpSlice {
    const void* rows;       // filtered rows (float / half / bf16)
    const float* scales;    // per-row scale (16-bit stores), else nullptr
    long long row_stride;   // elements between rows (BP_LINEAR_M32: 8-sample chunks per row)
    int row_len;            // samples per row (N)
    int nang;               // angles
    const AngleEntry* ang;  // [nang]
    void* out;              // image out_n * out_n (float or int32)
    int out_n;              // image side
    int tiles_x, tiles_y;
    int tile_base;          // first global work item of this slice
    int accumulate;         // 1: add into existing image
    int lut_formula;        // LUT modes
    int store;              // 0 f32, 1 f16, 2 bf16
    float out_scale;        // multiplies the final image (float modes)
    const void* rows_x[BP_MAX_SLICES_PER_ITEM - 1];   // BP_LINEAR2: rows, scales and image
    const float* scales_x[BP_MAX_SLICES_PER_ITEM - 1]; // of the item's other slices
    void* out_x[BP_MAX_SLICES_PER_ITEM - 1];
    const float* gscale;    // FP16 multi-slice modes: the item scale (device, set by the pack pass)
    // BP_LINEAR2: per (tile, angle) the tile geometry (dux, duy, ut - s0, s0 as int
    // bits), [tiles][nang], built once per plan, read one angle ahead instead of
    // formed in LDS per chunk; nullptr: formed per chunk
    const float4* tgeom;
};

struct BpLaunch {
    const BpSlice* slices;  // device array
    int nslices;
    int total_items;        // sum of tiles
    const int* item_order;  // optional permutation of work items (cost-sorted)
    unsigned int* counter;  // persistent work: kXcds counters kCounterStride apart, zeroed
                            // before the launch (nullptr: one item per block)
    int run;                // persistent work: list positions dealt to the XCDs in runs of this length
    int mode;               // BpMode
    int rule;               // LUT: 0 floor (float mask), 1 trunc
    int chunk;              // angles per LDS chunk
    int tile;               // tile config id
    int angle_parts;        // LUT16 integer mode: a tile's angles split over this many work
                            // items, their integer sums added atomically into a zeroed
                            // image (order-free, so bit-exact); 0 or 1: one item per tile
    int ang_begin, ang_end; // ang_end > 0: only the slice's angles [ang_begin, ang_end)
                            // (a segment of a launch sequence); 0: all of them
    int tgeom;              // BP_LINEAR2: 1 = every slice has BpSlice::tgeom
};

// Persistent launches take work-list positions from kXcds counters: the list is
// dealt to the eight XCDs in runs of consecutive positions, run j belonging to XCD
// j % 8, whose blocks (blockIdx % 8 is the XCD a block is dispatched to) take its
// runs from its own counter, then the other XCDs' runs from theirs.  Neighbouring
// tiles of one slice read overlapping sinogram segments, so an XCD working through
// a run finds them in its own L2.  Kernels that do not gain from it
// (bp_deals_runs) get a single run of the whole list: one queue in list order.
constexpr int kXcds = 8;
constexpr int kCounterStride = 32;      // counters 128 B apart
constexpr int kMaxRun = 256;
inline int work_run(int total_items, bool deal)
{
    if (!deal) return total_items < 1 ? 1 : total_items;    // one run: one queue in list order
    const int r = (total_items + kXcds - 1) / kXcds;
    return r < 1 ? 1 : r > kMaxRun ? kMaxRun : r;
}

void launch_bp(const BpLaunch& L, int grid, hipStream_t stream);

// Pointer fields of device tables (the image pointers of BpSlice entries),
// written by a kernel in stream order: an execution's patches are ordered with
// the work queued before it on the caller's stream, and a captured graph
// carries its own pointers, so every replay restores them.
constexpr int kPtrPatch = 32;
struct PtrPatch {
    void** dst[kPtrPatch];
    void* val[kPtrPatch];
    int n;
};
void launch_ptr_patch(const PtrPatch& P, hipStream_t stream);

// Packed input of the multi-slice modes, built once per execution from the
// filtered rows so that the back-projection stages a sample with one 8- or
// 16-byte load:
//   BP_LINEAR2      float4 (p_k, d_k, p'_k, d'_k), d_k = p_{k+1} - p_k (0 at the
//                   last sample);
//   BP_LINEAR4      two such float4 per sample, slices 0-1 then 2-3 ([row][k][2]);
//   BP_LINEAR_H2/H4/H8/H16 two / four / eight / sixteen FP16 pairs (p_k, d_k), values multiplied by
//                   row scale / item scale, where the item scale is the largest
//                   row scale of the slices over their angle list (a power of two);
//   BP_LINEAR_M32   FP16 samples p_k times row scale / item scale, in chunks of
//                   eight samples x 32 slices ([row][k / 8][slice][8]), zero past N.
struct PackJob {
    const void* rows[BP_MAX_SLICES_PER_ITEM];      // filtered rows (float / half / bf16)
    const float* scales[BP_MAX_SLICES_PER_ITEM];   // per-row scales (16-bit stores), else nullptr
    const AngleEntry* ang;  // FP16 pairs: the angle list whose rows set the item scale
    int nang;
    int ns;                 // slices: 2, or 4 / 8 / 16 / 32 (FP16 only)
    int half;               // 1: FP16 pairs
    int store;              // 0 f32, 1 f16, 2 bf16
    int nrows, N;           // rows of N samples, row stride N
    void* out;              // [nrows][N] packed samples
    float* gscale;          // FP16 pairs: the item scale, written by the scale pass
};
// Item scales of FP16-pair jobs (jobs of other kinds are skipped), then the
// packing of njobs jobs of one kind (0: FP32 pairs; 3: FP32, four slices; 2 / 4 /
// 8 / 16: FP16 pairs of that many slices; 32: the BP_LINEAR_M32 chunks) and one
// store, in stream order.
void launch_pack_scales(const PackJob* d_jobs, int njobs, hipStream_t stream);
void launch_pack(const PackJob* d_jobs, int njobs, int kind, int store, int max_rows, hipStream_t stream);

// ------------------------------------------------------- offset-CT stitch
// gazoDoc.cpp:3536-3665 for one multiplex row of every frame below 180 deg.
struct StitchJob {
    const void* sino;      // [frames][W] int16 or float
    int W, bin, trim;
    double center;         // user centre + 0.5 (gazoDoc.cpp:3560)
    double deltac;         // deltaCent * j * bin
    const int* pair;       // [nout] partner frame of output row i, -1 = row left zero
    int nout, out_width;
    void* out;             // [nout][out_width]
    int is_float;
};
void launch_stitch(const StitchJob& J, hipStream_t stream);
int bp_threads(int mode, int tile);         // workgroup size of a mode's kernel on a tile
int bp_tile_side(int tile);                 // tile width in pixels for a config id
int bp_tile_h(int tile);                    // tile height in pixels
int bp_default_chunk(int mode, int tile);   // angles per chunk
int bp_blocks_per_cu_tg(int mode, int tile, int chunk, bool tgeom);   // resident workgroups per CU
bool bp_has_tgeom(int mode);                // the mode has a tile geometry table variant
bool bp_deals_runs(int mode);               // persistent work dealt to the XCDs in runs
int bp_max_chunk(int mode);                 // largest chunk a mode supports
size_t bp_lds_bytes(int mode, int tile, int chunk);
void launch_zero_u32(unsigned int* p, int n, hipStream_t stream);
// Tile geometry table of BpSlice::tgeom for tiles of tw x th pixels, ntiles of them
// in rows of tiles_x, over nang angle entries.
void launch_tile_geometry(const AngleEntry* ang, int nang, int tiles_x, int ntiles, int tw, int th, float4* out,
                          hipStream_t stream);

}  // namespace rv
