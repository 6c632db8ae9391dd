// rv_bp.hip -- tiled back-projection for CDNA4 (gfx950).
//
// Replaces projAtomicKernel9 / CudaBackProjStream (cudaReconst.cu:31-39,
// 195-228: one launch per projection, a global atomic read-modify-write of the
// whole image per projection) with one launch per batch of slices:
//
//  * a workgroup (4 wave64) owns a TW x TH pixel tile and keeps its sums in
//    VGPRs across all projections, so the image is written once, coalesced;
//  * projections are processed in chunks: for every angle of the chunk the
//    detector segment that the tile can see (<= SEG samples) is staged in LDS,
//    together with the per-angle geometry (FP64 on the host, reduced here to a
//    tile-relative FP32 origin so the FP32 coordinate never exceeds SEG);
//  * within a wave, the 64 lanes cover an 8x8 pixel block, so the LDS reads of
//    one instruction fall in a window of about 12 consecutive samples;
//  * work items (slice, tile) come from one grid or, persistently, from atomic
//    work counters over a cost-sorted list (uneven slices in one launch), dealt
//    to the XCDs in runs of neighbouring tiles so that the segments they share
//    are read through one L2.
//
// Modes:
//  BP_LINEAR   : production FP32.  LDS holds per-sample pairs (a_j, d_j) with
//                d_j = p[k+1]-p[k] and a_j = p[k] - j*d_j (j = sample index
//                within the segment), so value(u) = a_j + u*d_j for the
//                segment-relative coordinate u: one FMA, no fract.  8 B/update.
//  BP_LINEAR2  : BP_LINEAR for two slices of identical geometry at once: LDS
//                holds (a_j, d_j, a'_j, d'_j), so one coordinate and one
//                16-byte LDS read serve two updates.  Its rows come packed by
//                pack_kernel as (p_k, d_k, p'_k, d'_k), so staging a
//                sample is one 16-byte load and two FMAs.  On gfx950 ds_read_b128
//                returns ~1.7x the bytes per clock of ds_read_b64 (measured
//                ~0.2 vs ~0.23 instructions per clock per CU), and the shared
//                coordinate cuts VALU work from ~5.3 to ~3.6 per update.
//  BP_LINEAR_H : FP16/BF16 filtered sinogram.  LDS holds FP16 pairs (p_k, d_k)
//                rescaled to one power-of-two scale per chunk of angles, value =
//                p + fract(u)*d (mixed-precision FMA), FP32 sums.  4 B/update.
//  BP_LINEAR_H2/H4: BP_LINEAR_H for two / four slices of identical geometry:
//                LDS holds the slices' FP16 pairs side by side (8 / 16 B per
//                sample), rescaled by one power-of-two scale per work item (the
//                largest row scale of its slices), and each update is one
//                v_dot2_f32_f16: sum += (p, d) . (1, fract(u)) in FP32.  One
//                coordinate, one fract and one (1, f) pack serve 2 / 4 updates.
//                pack_kernel forms the rescaled FP16 pairs once per execution,
//                so staging a sample is one 8- / 16-byte load.
//  BP_LUT_F32  : the reference's 16-sub-sample LUT (cudaReconst.cu:230-240)
//                in float, float sums; index t = ix*fcos + (iy*fsin + foffset)
//                in the reference's FP32 operation order.
//  BP_LUT_I32  : the reference LUT in int32 (x BACKPROJ_SCALE), int32 sums --
//                bit-exact with the reference integer path.
//                Both LUT modes stage one pair (p0, p1p0) per sample (8 B) and
//                form entry s of sample j as p0 + p1p0 * s at use, the float
//                operations of the reference's LUT build.
// This file is compiled with -ffp-contract=off; FMAs are written explicitly.
#include "rv_internal.h"
#include <type_traits>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

namespace rv {

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
// the tile geometry table, read with global (vector-memory) loads: their counter
// is separate from the LDS reads', so a load in flight does not hold those up
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) f4v* gf4p;

// smallest integer >= sqrt(v)
constexpr int isqrt_ceil(int v)
{
    int r = 0;
    while (r * r < v) r++;
    return r;
}

template <int PX, int PY, int NW = 4>
struct TileCfg {
    static constexpr int WX = 8 * PX, WY = 8 * PY;   // wave block
    static constexpr int TW = 2 * WX, TH = (NW / 2) * WY;   // 2 x NW/2 waves
    static constexpr int T = 64 * NW;
    // detector footprint of a tile: (TW-1)|cos| + (TH-1)|sin| <= the tile diagonal D;
    // with s0 = floor(umin) - 1 every index used is <= D + 2, so SEG >= D + 3
    static constexpr int SEG = (isqrt_ceil((TW - 1) * (TW - 1) + (TH - 1) * (TH - 1)) + 7) / 4 * 4;
};

template <int MODE>
struct AccT { using type = float; };
template <>
struct AccT<BP_LUT_I32> { using type = int; };

constexpr bool is_half_multi(int mode)
{
    return mode == BP_LINEAR_H2 || mode == BP_LINEAR_H4 || mode == BP_LINEAR_H8 || mode == BP_LINEAR_H16;
}

constexpr bool is_linear(int mode)
{
    return mode == BP_LINEAR || mode == BP_LINEAR_H || mode == BP_LINEAR2 || mode == BP_LINEAR4 ||
           is_half_multi(mode);
}

// LDS bytes per staged angle
constexpr int entry_bytes(int mode, int seg)
{
    return (mode == BP_LINEAR_H || is_half_multi(mode)) ? seg * 4 * bp_slices_per_item(mode)
                                                         : seg * 8 * bp_slices_per_item(mode);
}

__device__ __forceinline__ float load_row_val(const void* rows, const float* scales, const BpSlice& S, int row,
                                              int k)
{
    const long long off = (long long)row * S.row_stride + k;
    if (S.store == 0) return ((const float*)rows)[off];
    const float sc = scales[row];
    if (S.store == 1) return __half2float(((const __half*)rows)[off]) * sc;
    return __bfloat162float(((const __hip_bfloat16*)rows)[off]) * sc;
}

// 16-bit row value without its row scale
__device__ __forceinline__ float raw_row_val(const void* rows, const BpSlice& S, int row, int k)
{
    const long long off = (long long)row * S.row_stride + k;
    if (S.store == 1) return __half2float(((const __half*)rows)[off]);
    return __bfloat162float(((const __hip_bfloat16*)rows)[off]);
}

__device__ __forceinline__ float load_row_val(const BpSlice& S, int row, int k)
{
    return load_row_val(S.rows, S.scales, S, row, k);
}

// (a_j, d_j) of the staged segment: value(u) = a + u*d for u in [j, j+1)
__device__ __forceinline__ float2 linear_pair(const void* rows, const float* scales, const BpSlice& S, int row,
                                              int k, int j)
{
    const int N = S.row_len;
    float av = 0.f, dv = 0.f;
    if (row >= 0 && k >= 0 && k < N) {
        const float p0 = load_row_val(rows, scales, S, row, k);
        if (k < N - 1) {
            const float p1 = load_row_val(rows, scales, S, row, k + 1);
            dv = p1 - p0;
            av = fmaf(-(float)j, dv, p0);
        } else {
            av = p0;
        }
    }
    return make_float2(av, dv);
}

// Next work-list position of a persistent block (one thread calls it): runs of
// L.run positions are dealt to the XCDs in turn, run j to XCD j % kXcds, and a
// block takes the runs of its own XCD first, then, when those are taken, the
// runs of the next XCDs in order (probe counts the counters already found
// empty).  Counter q maps its k-th hand-out to ((k / run) * kXcds + q) * run +
// k % run, which grows with k, so the first position past the list means every
// later one is too; L.total_items when all are taken.
__device__ __forceinline__ int next_work_item(const BpLaunch& L, int& probe)
{
    const int xcd = (int)(blockIdx.x % kXcds);
    while (probe < kXcds) {
        const int q = (xcd + probe) % kXcds;
        const unsigned int k = atomicAdd(L.counter + q * kCounterStride, 1u);
        const long long pos = ((long long)(k / (unsigned)L.run) * kXcds + q) * L.run + k % (unsigned)L.run;
        if (pos < L.total_items) return (int)pos;
        probe++;
    }
    return L.total_items;
}

// TG (BP_LINEAR2 only): the per-angle tile geometry comes from BpSlice::tgeom,
// built once per plan, so the LDS holds only the staged samples and a chunk
// needs one barrier less.
template <int MODE, int PX, int PY, int NW = 4, bool TG = false>
__global__ void __launch_bounds__(64 * NW, NW == 8 ? 4 : 1) bp_tiled_kernel(BpLaunch L)
{
    using C = TileCfg<PX, PY, NW>;
    using Acc = typename AccT<MODE>::type;
    constexpr int SEG = C::SEG;
    constexpr int NS = bp_slices_per_item(MODE);             // slices per work item
    constexpr int NQ = NS / 2;                                // float4 per staged sample (multi-slice)
    const int chunk = L.chunk;

    // LDS layout (pointer arithmetic only, so every access stays a ds_* op):
    // tab[chunk] float4 | s0s[chunk] int | rws[chunk] int | pad | data
    extern __shared__ float4 lds4[];
    float4* tab = lds4;
    int* s0s = (int*)(lds4 + chunk);
    int* rws = s0s + chunk;
    void* data = (void*)(lds4 + chunk + (2 * chunk + 3) / 4);
    __shared__ int item_sh;
    __shared__ float chunk_scale;

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wx0 = (wave & 1) * C::WX, wy0 = (wave >> 1) * C::WY;
    const int lx = lane & 7, ly = lane >> 3;
    // per-lane offset of the first pixel of the lane inside the tile
    const float dxl = (float)(wx0 + lx), dyl = (float)(wy0 + ly);

    int iter = 0, probe = 0;
    while (true) {
        int item;
        if (L.counter) {
            // only the kernels bp_deals_runs names deal the list to the XCDs; the
            // others take positions in list order from the first counter (the probe
            // loop costs the FP32 pair kernel three VGPRs and a wave per SIMD)
            if constexpr (MODE == BP_LINEAR_H16) {
                if (tid == 0) item_sh = next_work_item(L, probe);
            } else {
                if (tid == 0) item_sh = (int)atomicAdd(L.counter, 1u);
            }
            __syncthreads();
            item = __builtin_amdgcn_readfirstlane(item_sh);     // uniform: slice data in scalar registers
            __syncthreads();
        } else {
            item = (iter == 0) ? (int)blockIdx.x : L.total_items;
        }
        iter++;
        if (item >= L.total_items) break;
        if (L.item_order) item = L.item_order[item];
        const int parts = (MODE == BP_LUT_I32 && L.angle_parts > 1) ? L.angle_parts : 1;
        int part = 0;
        if (parts > 1) {
            part = item % parts;
            item /= parts;
        }

        // slice lookup: last slice with tile_base <= item
        int lo = 0, hi = L.nslices - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (L.slices[mid].tile_base <= item) lo = mid; else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);           // uniform: the slice's fields in scalar registers
        const BpSlice S = L.slices[lo];
        const int t = item - S.tile_base;
        const float4* tgv = TG ? S.tgeom + (size_t)t * S.nang : nullptr;   // this tile's row of the table
        const int X0 = (t % S.tiles_x) * C::TW;
        const int Y0 = (t / S.tiles_x) * C::TH;
        const int N = S.row_len;

        Acc acc[PY][PX], accx[NS > 1 ? NS - 1 : 1][PY][PX];
#pragma unroll
        for (int j = 0; j < PY; j++)
#pragma unroll
            for (int i = 0; i < PX; i++) {
                acc[j][i] = (Acc)0;
#pragma unroll
                for (int e = 0; e + 1 < NS; e++) accx[e][j][i] = (Acc)0;
            }

        // FP16 multi-slice modes: the item scale the packed pairs were divided by
        const float gscale = is_half_multi(MODE) ? *S.gscale : 1.f;

        // angles of this work item: all (or the launch's window of them), or one
        // chunk-aligned part of those
        const int w0 = L.ang_end > 0 ? min(L.ang_begin, S.nang) : 0;
        const int w1 = L.ang_end > 0 ? min(L.ang_end, S.nang) : S.nang;
        int a_begin = w0, a_end = w1;
        if (parts > 1) {
            const int per = ((w1 - w0 + parts - 1) / parts + chunk - 1) / chunk * chunk;
            a_begin = min(w1, w0 + part * per);
            a_end = min(w1, a_begin + per);
        }
        for (int a0 = a_begin; a0 < a_end; a0 += chunk) {
            const int na = TG ? min(chunk, a_end - a0) : chunk;     // angles of this chunk
            // ---- per-angle tile geometry
            if (!TG && tid < chunk) {
                const int ga = a0 + tid;
                if (ga < a_end) {
                    const AngleEntry e = S.ang[ga];
                    if (is_linear(MODE)) {
                        const double ut = e.u0 + (double)X0 * e.dux + (double)Y0 * e.duy;
                        const double umin = ut + fmin(0.0, (C::TW - 1) * e.dux) + fmin(0.0, (C::TH - 1) * e.duy);
                        const int s0 = (int)floor(umin) - 1;
                        const float sc = (MODE == BP_LINEAR_H) ? S.scales[e.row] : 0.f;
                        tab[tid] = make_float4((float)e.dux, (float)e.duy, (float)(ut - (double)s0), sc);
                        s0s[tid] = s0;
                    } else {
                        const double fc = e.fcos, fs = e.fsin;
                        const double tt = (double)X0 * fc + (double)Y0 * fs + (double)e.foffset;
                        const double tmin = tt + fmin(0.0, (C::TW - 1) * fc) + fmin(0.0, (C::TH - 1) * fs);
                        const int s0 = (int)floor(tmin / 16.0) - 1;
                        tab[tid] = make_float4(e.fcos, e.fsin, e.foffset, 0.f);
                        s0s[tid] = s0;
                    }
                    rws[tid] = e.row;
                } else {
                    tab[tid] = is_linear(MODE) ? make_float4(0.f, 0.f, 1.f, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
                    s0s[tid] = is_linear(MODE) ? -(1 << 29) : 0;
                    rws[tid] = -1;
                }
            }
            if (MODE == BP_LINEAR_H) {
                // one power-of-two scale for the chunk: the largest row scale (chunk <= 64)
                __syncthreads();
                if (tid < 64) {
                    float m = (tid < chunk) ? tab[tid].w : 0.f;
                    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
                    if (tid == 0) chunk_scale = (m > 0.f) ? m : 1.f;
                }
            }
            if (!TG) __syncthreads();

            // ---- stage detector segments into LDS
            // row and segment origin of angle a of the chunk
            auto origin = [&](int a, int& row, int& s0) {
                if (TG) {
                    const bool ok = a < na;
                    row = ok ? S.ang[a0 + a].row : -1;
                    s0 = ok ? __float_as_int(tgv[a0 + a].w) : 0;
                } else {
                    row = rws[a];
                    s0 = s0s[a];
                }
            };
            if (MODE == BP_LINEAR) {
                float2* pairs = (float2*)data;
                for (int idx = tid; idx < na * SEG; idx += C::T) {
                    const int a = idx / SEG, j = idx - a * SEG;
                    int row, s0;
                    origin(a, row, s0);
                    pairs[idx] = linear_pair(S.rows, S.scales, S, row, s0 + j, j);
                }
            } else if (MODE == BP_LINEAR2 || MODE == BP_LINEAR4) {
                // a_j = p_k - j d_k as linear_pair forms it (d = 0 at the last sample);
                // NQ float4 planes per sample (two slices each), plane q at q * SEG
                float4* quads = (float4*)data;
                const float4* packed = (const float4*)S.rows;
                auto stage = [&](int a, int j, int row, int s0) {
                    const int k = s0 + j;
                    const float fj = -(float)j;
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (row >= 0 && k >= 0 && k < N) v = packed[((long long)row * S.row_stride + k) * NQ + q];
                        quads[(a * NQ + q) * SEG + j] = make_float4(fmaf(fj, v.y, v.x), v.y, fmaf(fj, v.w, v.z), v.w);
                    }
                };
                if (C::T % chunk == 0) {
                    // threads of one angle: its row and segment origin read once per
                    // thread (C2, FP32 pairs: back-projection 40.85 -> 39.87 ms)
                    const int tpa = C::T / chunk;
                    const int a = tid / tpa, j0 = tid - a * tpa;
                    int row, s0;
                    origin(a, row, s0);
                    if (!TG || a < na)
                        for (int j = j0; j < SEG; j += tpa) stage(a, j, row, s0);
                } else {
                    for (int idx = tid; idx < na * SEG; idx += C::T) {
                        const int a = idx / SEG, j = idx - a * SEG;
                        int row, s0;
                        origin(a, row, s0);
                        stage(a, j, row, s0);
                    }
                }
            } else if (MODE == BP_LINEAR_H) {
                __half2* pairs = (__half2*)data;
                const float inv = 1.f / chunk_scale;              // power of two: exact
                for (int idx = tid; idx < chunk * SEG; idx += C::T) {
                    const int a = idx / SEG, j = idx - a * SEG;
                    const int row = rws[a];
                    const int k = s0s[a] + j;
                    float pv = 0.f, dv = 0.f;
                    if (row >= 0 && k >= 0 && k < N) {
                        pv = __half2float(__float2half_rn(load_row_val(S, row, k) * inv));
                        if (k < N - 1) {
                            const float p1 = __half2float(__float2half_rn(load_row_val(S, row, k + 1) * inv));
                            dv = p1 - pv;
                        }
                    }
                    pairs[idx] = __halves2half2(__float2half_rn(pv), __float2half_rn(dv));
                }
            } else if (is_half_multi(MODE)) {
                // packed FP16 pairs of the item's slices: one 8- / 16-byte load per sample
                for (int idx = tid; idx < na * SEG; idx += C::T) {
                    const int a = idx / SEG, j = idx - a * SEG;
                    int row, s0;
                    origin(a, row, s0);
                    const int k = s0 + j;
                    const bool in = row >= 0 && k >= 0 && k < N;
                    const long long off = (long long)row * S.row_stride + k;
                    if (NS >= 8) {
                        // NS / 4 16-byte planes per angle: slices 0-3, then 4-7, ...
                        constexpr int NPL = NS >= 8 ? NS / 4 : 1;
                        uint4 v[NPL];
#pragma unroll
                        for (int pl = 0; pl < NPL; pl++) {
                            v[pl] = make_uint4(0u, 0u, 0u, 0u);
                            if (in) v[pl] = ((const uint4*)S.rows)[NPL * off + pl];
                        }
#pragma unroll
                        for (int pl = 0; pl < NPL; pl++) ((uint4*)data)[(NPL * a + pl) * SEG + j] = v[pl];
                    } else if (NS == 4) {
                        uint4 v = make_uint4(0u, 0u, 0u, 0u);
                        if (in) v = ((const uint4*)S.rows)[off];
                        ((uint4*)data)[idx] = v;
                    } else {
                        uint2 v = make_uint2(0u, 0u);
                        if (in) v = ((const uint2*)S.rows)[off];
                        ((uint2*)data)[idx] = v;
                    }
                }
            } else {
                // LUT modes: per sample the pair (p0, p1p0) of the reference LUT
                // igp[16 j + s] = p0 + p1p0 * s; the entry is formed at use, with
                // the same float operations, instead of staging 16 entries.
                float2* lut = (float2*)data;
                for (int idx = tid; idx < chunk * SEG; idx += C::T) {
                    const int a = idx / SEG, j = idx - a * SEG;
                    const int row = rws[a];
                    const int k = s0s[a] + j;
                    float2 e = make_float2(0.f, 0.f);
                    if (row >= 0 && k >= 0 && k < N) {
                        const float q0 = load_row_val(S, row, k);
                        const float q1 = (k < N - 1) ? load_row_val(S, row, k + 1) : 0.f;
                        float p0, p1p0;
                        if (MODE == BP_LUT_I32) {
                            if (S.lut_formula == 0) {       // general.cpp:2131-2133
                                p0 = q0 * 100.f;
                                p1p0 = (k == N - 1) ? 0.0f : (q1 - q0) / 16.f * 100.f;
                            } else {                        // cudaReconst.cu:237-238
                                p0 = q0 * 100.f;
                                p1p0 = (k == N - 1) ? 0.0f : (q1 * 100.f - p0) / 16.f;
                            }
                        } else {
                            p0 = q0;
                            p1p0 = (k == N - 1) ? 0.0f : (q1 - p0) / 16.f;
                        }
                        e = make_float2(p0, p1p0);
                    }
                    lut[idx] = e;
                }
            }
            __syncthreads();

            // ---- accumulate
            // fn(a, tile geometry) for every angle of the chunk.  With the table, the
            // next angle's entry loads (vector memory, counted apart from the LDS
            // reads) while this angle's updates run.  For the FP32 pairs the compiler
            // schedules this form for six waves per SIMD (79 VGPRs, two LDS reads in
            // flight per wave); measured faster than forms scheduled for four (64-slice
            // C2 stack: 37.8 ms vs 39.4 ms per-chunk geometry, 41.2 ms table read at the
            // top of each angle)
            auto for_angles = [&](auto&& fn) {
                if constexpr (TG) {
                    const gf4p tgg = (gf4p)tgv;
                    f4v tn = tgg[a0];
                    for (int a = 0; a < na; a++) {
                        const float4 tb = make_float4(tn.x, tn.y, tn.z, tn.w);
                        if (a + 1 < na) tn = tgg[a0 + a + 1];
                        fn(a, tb);
                    }
                } else {
                    for (int a = 0; a < na; a++) fn(a, tab[a]);
                }
            };
            if (MODE == BP_LINEAR) {
                const float2* pairs = (const float2*)data;
                for_angles([&](int a, const float4 tb) {
                    const float2* seg = pairs + a * SEG;
                    // u(i, j) = base + 8 i cos + 8 j sin, base at the lane's first pixel
                    const float base = fmaf(dyl, tb.y, fmaf(dxl, tb.x, tb.z));
#pragma unroll
                    for (int j = 0; j < PY; j++) {
                        const float bj = (j == 0) ? base : fmaf((float)(8 * j), tb.y, base);
#pragma unroll
                        for (int i = 0; i < PX; i++) {
                            // u >= 1 by construction of s0, so truncation is the floor
                            const float u = (i == 0) ? bj : fmaf((float)(8 * i), tb.x, bj);
                            const float2 ad = seg[(int)u];
                            acc[j][i] += fmaf(u, ad.y, ad.x);
                        }
                    }
                });
            } else if (MODE == BP_LINEAR2 || MODE == BP_LINEAR4) {
                const float4* quads = (const float4*)data;
                auto angle = [&](int a, const float4 tb) {
                    const float4* seg = quads + a * SEG * NQ;
                    const float base = fmaf(dyl, tb.y, fmaf(dxl, tb.x, tb.z));
#pragma unroll
                    for (int j = 0; j < PY; j++) {
                        const float bj = (j == 0) ? base : fmaf((float)(8 * j), tb.y, base);
#pragma unroll
                        for (int i = 0; i < PX; i++) {
                            const float u = (i == 0) ? bj : fmaf((float)(8 * i), tb.x, bj);
                            const float4* e = seg + (int)u;
#pragma unroll
                            for (int q = 0; q < NQ; q++) {
                                const float4 v = e[q * SEG];
                                if (q == 0) acc[j][i] += fmaf(u, v.y, v.x);
                                else accx[2 * q - 1][j][i] += fmaf(u, v.y, v.x);
                                accx[2 * q][j][i] += fmaf(u, v.w, v.z);
                            }
                        }
                    }
                };
                for_angles(angle);
            } else if (is_half_multi(MODE)) {
                // FP16 pairs of 2 / 4 / 8 / 16 slices per sample (8, 16: two / four 16-byte planes).
                using QT = typename std::conditional<NS >= 4, uint4, uint2>::type;
                constexpr int NPL = NS >= 8 ? NS / 4 : 1;
                const QT* ents = (const QT*)data;
                // entries of one pixel row and their FP16 weights (1, fract(u)),
                // rounded toward zero
                auto load_row = [&](const float4& tb, float base, const QT* seg, int j, QT (&q)[PX][NPL],
                                    half2v (&w)[PX]) {
                    const float bj = (j == 0) ? base : fmaf((float)(8 * j), tb.y, base);
#pragma unroll
                    for (int i = 0; i < PX; i++) {
                        const float u = (i == 0) ? bj : fmaf((float)(8 * i), tb.x, bj);
#pragma unroll
                        for (int pl = 0; pl < NPL; pl++) q[i][pl] = seg[pl * SEG + (int)u];
                        w[i] = __builtin_bit_cast(half2v, __builtin_amdgcn_cvt_pkrtz(1.f, __builtin_amdgcn_fractf(u)));
                    }
                };
                auto dot_row = [&](int j, const QT (&q)[PX][NPL], const half2v (&w)[PX]) {
                    constexpr int PER = (NS >= 4) ? 4 : 2;
#pragma unroll
                    for (int i = 0; i < PX; i++)
#pragma unroll
                        for (int pl = 0; pl < NPL; pl++) {
                            const uint32_t* qe = (const uint32_t*)&q[i][pl];
#pragma unroll
                            for (int e4 = 0; e4 < PER; e4++) {
                                const int e = pl * 4 + e4;
                                const half2v h = __builtin_bit_cast(half2v, qe[e4]);
                                if (e == 0) acc[j][i] = __builtin_amdgcn_fdot2(h, w[i], acc[j][i], false);
                                else accx[e - 1][j][i] = __builtin_amdgcn_fdot2(h, w[i], accx[e - 1][j][i], false);
                            }
                        }
                };
                for_angles([&](int a, const float4 tb) {
                    const float base = fmaf(dyl, tb.y, fmaf(dxl, tb.x, tb.z));
                    const QT* seg = ents + a * NPL * SEG;
                    if constexpr (NS == 8 && NW == 4) {
                        // software pipeline over pixel rows: the next row's entries
                        // are in flight while this row's dot products issue (242
                        // VGPRs, two waves per SIMD; measured 3% faster than one
                        // row at a time)
                        QT q[2][PX][NPL];
                        half2v w[2][PX];
                        load_row(tb, base, seg, 0, q[0], w[0]);
#pragma unroll
                        for (int j = 0; j < PY; j++) {
                            if (j + 1 < PY) load_row(tb, base, seg, j + 1, q[(j + 1) & 1], w[(j + 1) & 1]);
                            dot_row(j, q[j & 1], w[j & 1]);
                        }
                    } else if constexpr (NS >= 4) {
                        // one row at a time (four slices: 118 VGPRs, four waves per SIMD;
                        // the whole angle at once needs 151 and allows three, measured 6%
                        // slower; sixteen slices: four planes of each pixel of the row)
#pragma unroll
                        for (int j = 0; j < PY; j++) {
                            QT q[PX][NPL];
                            half2v w[PX];
                            load_row(tb, base, seg, j, q, w);
                            dot_row(j, q, w);
                            __builtin_amdgcn_sched_barrier(0);     // keep the rows apart
                        }
                    } else {
                        // two slices: every entry of the angle in flight, then the dot products
                        QT q[PY][PX][NPL];
                        half2v w[PY][PX];
#pragma unroll
                        for (int j = 0; j < PY; j++) load_row(tb, base, seg, j, q[j], w[j]);
#pragma unroll
                        for (int j = 0; j < PY; j++) dot_row(j, q[j], w[j]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                });
            } else if (MODE == BP_LINEAR_H) {
                const __half2* pairs = (const __half2*)data;
                float cacc[PY][PX];
#pragma unroll
                for (int j = 0; j < PY; j++)
#pragma unroll
                    for (int i = 0; i < PX; i++) cacc[j][i] = 0.f;
                for (int a = 0; a < chunk; a++) {
                    const float4 tb = tab[a];
                    const __half2* seg = pairs + a * SEG;
                    const float base = fmaf(dyl, tb.y, fmaf(dxl, tb.x, tb.z));
#pragma unroll
                    for (int j = 0; j < PY; j++) {
                        const float bj = (j == 0) ? base : fmaf((float)(8 * j), tb.y, base);
#pragma unroll
                        for (int i = 0; i < PX; i++) {
                            const float u = (i == 0) ? bj : fmaf((float)(8 * i), tb.x, bj);
                            const __half2 pd = seg[(int)u];
                            const float f = __builtin_amdgcn_fractf(u);
                            cacc[j][i] += fmaf(f, __high2float(pd), __low2float(pd));
                        }
                    }
                }
                const float cs = chunk_scale;
#pragma unroll
                for (int j = 0; j < PY; j++)
#pragma unroll
                    for (int i = 0; i < PX; i++) acc[j][i] = fmaf(cacc[j][i], cs, acc[j][i]);
            } else {
                const float2* lut = (const float2*)data;
                // the index rule is fixed per launch: one loop per rule, so an
                // update converts once (no per-update select of floor and trunc)
                auto lut_loop = [&](auto rule_tag) {
                    constexpr int RULE = decltype(rule_tag)::value;
                    for (int a = 0; a < chunk; a++) {
                        const float4 tb = tab[a];
                        const int base = 16 * s0s[a];
                        const float2* la = lut + (size_t)a * SEG;
#pragma unroll
                        for (int j = 0; j < PY; j++) {
                            const float fy = (float)(Y0 + (int)dyl + 8 * j);
                            const float rt = fy * tb.y + tb.z;              // iy*fsin + foffset
#pragma unroll
                            for (int i = 0; i < PX; i++) {
                                const float fx = (float)(X0 + (int)dxl + 8 * i);
                                const float tt = fx * tb.x + rt;            // ix*fcos + (...)
                                const int idx = (RULE == 0) ? (int)floorf(tt) : (int)tt;
                                const int rel = idx - base;              // >= 0 by the choice of s0
                                const float2 e = la[rel >> 4];
                                const float lv = e.x + e.y * (float)(rel & 15);
                                const Acc v = (MODE == BP_LUT_I32) ? (Acc)(int)lv : (Acc)lv;
                                if (MODE == BP_LUT_I32)
                                    acc[j][i] = (Acc)((unsigned int)acc[j][i] + (unsigned int)v);
                                else
                                    acc[j][i] += v;
                            }
                        }
                    }
                };
                if (L.rule == 0) lut_loop(std::integral_constant<int, 0>{});
                else lut_loop(std::integral_constant<int, 1>{});
            }
            __syncthreads();
        }

        // ---- write the tile
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int iy = Y0 + (int)dyl + 8 * j;
            if (iy >= S.out_n) continue;
#pragma unroll
            for (int i = 0; i < PX; i++) {
                const int ix = X0 + (int)dxl + 8 * i;
                if (ix >= S.out_n) continue;
                const long long o = (long long)iy * S.out_n + ix;
                if (MODE == BP_LUT_I32) {
                    unsigned int* out = (unsigned int*)S.out;
                    const unsigned int v = (unsigned int)acc[j][i];
                    if (parts > 1) atomicAdd(out + o, v);       // the parts of a tile meet here
                    else out[o] = S.accumulate ? out[o] + v : v;
                } else {
                    float* out = (float*)S.out;
                    const float v = (float)acc[j][i] * gscale * S.out_scale;
                    out[o] = S.accumulate ? out[o] + v : v;
#pragma unroll
                    for (int e = 0; e + 1 < NS; e++) {
                        float* outx = (float*)S.out_x[e];
                        const float vx = (float)accx[e][j][i] * gscale * S.out_scale;
                        outx[o] = S.accumulate ? outx[o] + vx : vx;
                    }
                }
            }
        }
        if (!L.counter) break;
    }
}

// BpSlice::tgeom: the tile origin of every (tile, angle), formed exactly as
// bp_tiled_kernel forms it per chunk (same FP64 operations, so the images are
// identical with and without the table).
__global__ void __launch_bounds__(256) tile_geometry_kernel(const AngleEntry* ang, int nang, int tiles_x, int ntiles,
                                                            int tw, int th, float4* out)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nang * ntiles) return;
    const int t = (int)(i / nang);
    const int a = (int)(i - (long long)t * nang);
    const AngleEntry e = ang[a];
    const int X0 = (t % tiles_x) * tw, Y0 = (t / tiles_x) * th;
    const double ut = e.u0 + (double)X0 * e.dux + (double)Y0 * e.duy;
    const double umin = ut + fmin(0.0, (tw - 1) * e.dux) + fmin(0.0, (th - 1) * e.duy);
    const int s0 = (int)floor(umin) - 1;
    out[i] = make_float4((float)e.dux, (float)e.duy, (float)(ut - (double)s0), __int_as_float(s0));
}

void launch_tile_geometry(const AngleEntry* ang, int nang, int tiles_x, int ntiles, int tw, int th, float4* out,
                          hipStream_t stream)
{
    const long long n = (long long)nang * ntiles;
    if (n <= 0) return;
    hipLaunchKernelGGL(tile_geometry_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ang, nang,
                       tiles_x, ntiles, tw, th, out);
}

template <int STORE>
__device__ __forceinline__ float stored_val(const void* rows, long long i)
{
    if (STORE == 0) return ((const float*)rows)[i];
    if (STORE == 1) return __half2float(((const __half*)rows)[i]);
    return __bfloat162float(((const __hip_bfloat16*)rows)[i]);
}

// Item scale of the FP16-pair jobs: the largest row scale of the job's slices
// over its angle list (row scales are powers of two, so 1 / scale is exact).
__global__ void __launch_bounds__(1024) pack_scale_kernel(const PackJob* jobs)
{
    __shared__ float red[16];
    const PackJob& J = jobs[blockIdx.x];
    if (!J.half) return;
    float m = 0.f;
    const int ns = J.ns;
    for (int i = threadIdx.x; i < J.nang * ns; i += 1024) {
        const int a = i / ns, e = i - a * ns;
        m = fmaxf(m, J.scales[e][J.ang[a].row]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; w++) m = fmaxf(m, red[w]);
        *J.gscale = (m > 0.f) ? m : 1.f;
    }
}

// KIND 0: FP32 pairs of two slices, 3: FP32 pairs of four slices (two float4 per
// sample); KIND 2 / 4 / 8 / 16: FP16 pairs of that many slices.  Blocks stride
// over rows, threads over samples (no index division; each row's scales read
// once); every slice's loads are issued together.
template <int KIND, int STORE>
__global__ void __launch_bounds__(256) pack_kernel(const PackJob* jobs)
{
    constexpr bool FP32 = KIND == 0 || KIND == 3;
    constexpr int NS = KIND == 0 ? 2 : KIND == 3 ? 4 : KIND;
    const PackJob J = jobs[blockIdx.y];
    const int N = J.N;
    const float inv = FP32 ? 1.f : 1.f / *J.gscale;
    for (int row = blockIdx.x; row < J.nrows; row += gridDim.x) {
        float mul[NS];
#pragma unroll
        for (int e = 0; e < NS; e++) mul[e] = STORE ? J.scales[e][row] * inv : 1.f;
        const long long r0 = (long long)row * N;
        for (int k = threadIdx.x; k < N; k += 256) {
            const long long i = r0 + k;
            const bool last = k == N - 1;
            const long long i1 = last ? i : i + 1;      // in-row neighbour (unused at the last sample)
            float p0[NS], p1[NS];
#pragma unroll
            for (int e = 0; e < NS; e++) {
                p0[e] = stored_val<STORE>(J.rows[e], i);
                p1[e] = stored_val<STORE>(J.rows[e], i1);
            }
            if (FP32) {
                // the FP32 values and difference of the single-slice staging
                float v[2 * NS];
#pragma unroll
                for (int e = 0; e < NS; e++) {
                    const float a0 = STORE ? p0[e] * mul[e] : p0[e];
                    const float a1 = STORE ? p1[e] * mul[e] : p1[e];
                    v[2 * e] = a0;
                    v[2 * e + 1] = last ? 0.f : a1 - a0;
                }
#pragma unroll
                for (int q = 0; q < NS / 2; q++)
                    ((float4*)J.out)[i * (NS / 2) + q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
            } else {
                uint32_t w[NS];
#pragma unroll
                for (int e = 0; e < NS; e++) {
                    const float pv = __half2float(__float2half_rn(p0[e] * mul[e]));
                    const float dv = last ? 0.f : __half2float(__float2half_rn(p1[e] * mul[e])) - pv;
                    w[e] = __builtin_bit_cast(uint32_t, __halves2half2(__float2half_rn(pv), __float2half_rn(dv)));
                }
                if (NS >= 8) {
                    constexpr int NPL = NS >= 8 ? NS / 4 : 1;
#pragma unroll
                    for (int pl = 0; pl < NPL; pl++)
                        ((uint4*)J.out)[NPL * i + pl] = make_uint4(w[(4 * pl) % NS], w[(4 * pl + 1) % NS],
                                                                   w[(4 * pl + 2) % NS], w[(4 * pl + 3) % NS]);
                } else if (NS == 4) {
                    ((uint4*)J.out)[i] = make_uint4(w[0], w[1], w[2 % NS], w[3 % NS]);
                } else {
                    ((uint2*)J.out)[i] = make_uint2(w[0], w[1]);
                }
            }
        }
    }
}

// ------------------------------------------------ matrix-core back-projection
// BP_LINEAR_M32: 32 slices of one geometry per work item, FP16 samples, FP32
// sums, the multiply-adds on the MFMA units.  For one angle and one block of
// 8 x 4 pixels, linear interpolation of 32 slices is a matrix product over a
// window of 16 detector samples j0 .. j0 + 15:
//
//   D[slice][pixel] += sum_k A[slice][k] * B[k][pixel],
//   A[slice][k] = p_slice(j0 + k),  B[k][pixel] = max(0, 1 - |u_pixel - j0 - k|),
//
// i.e. 1 - f at the sample below the pixel's coordinate u, f = fract(u) at the
// one above, zero elsewhere: one v_mfma_f32_32x32x16_f16 does 1024 updates.  The
// 32 pixels of a block (8 x 4) span at most 7|cos| + 3|sin| <= 7.62 samples, so
// with j0 = floor(min u) rounded down to a multiple of 4 the window always holds
// both neighbours of every pixel.  The vector ALU only builds B: the pixel's
// coordinate, offset so that it lies on a 2^-11 grid (t in [4096, 8192)), gives
// f exactly in FP16 with (1 - f) + f = 1, and the pair (1 - f, f) is placed at its
// position with four byte permutes whose selectors come from a 22-entry table
// indexed by the integer part of t (every position a lane can meet, so no clamp).
// Per lane the MFMA operands hold eight consecutive k (lanes 0-31: k 0-7, lanes
// 32-63: k 8-15), A for slice = lane % 32 and B for pixel = lane % 32; the
// accumulator of lane l, register r is slice 8 (r / 4) + 4 (l / 32) + r % 4 of
// pixel l % 32.
//
// A 32 x 32 tile is 32 blocks; eight waves own four blocks each (a 32 x 4
// strip), 64 FP32 sums per lane, so that four waves fit per SIMD (with eight
// blocks per wave, 213 VGPRs allowed two, and the staging phases were not
// hidden: 28.9 ms for the 64-slice 2048^2 x 1800 stack).  Per chunk of angles: (1) per (angle, block)
// the window start and the coordinate offset, in FP64 per angle as the other
// kernels do; (2) the 32 slices' segments of 64 samples staged into LDS (one
// 16-byte chunk of eight samples per thread and angle, rows 136 B apart, so the
// 8-byte LDS reads of 16 lanes fall in distinct banks); (3) per angle and block
// one MFMA.  Samples outside the detector are zero (the row is zero-extended).
namespace m32 {
constexpr int kTile = 32;
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
constexpr int kBlocksPerWave = 32 / kWaves;
constexpr int kChunks = 8;                 // staged 8-sample chunks per angle and slice
constexpr int kPitch = 136;                // LDS bytes per staged row
constexpr int kMaxJ0 = kChunks * 8 - 16;   // window start bound (the analysis gives <= 44)
constexpr int kShift = 4112;               // coordinate offset: t in [4104, 4128)
constexpr int kSel = 22;                   // selector entries: lane-local positions -9 .. 12
constexpr int kItem = 16 * kSel;           // the work item's index (after the selector table)
constexpr int kHead = kItem + 16;          // per-angle tables from here
}

typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

size_t bp_m32_lds_bytes(int chunk)
{
    using namespace m32;
    return kHead + 16 * (size_t)chunk + (((size_t)4 * chunk + 15) & ~(size_t)15) + 512 * (size_t)chunk +
           (size_t)32 * kPitch * chunk;
}

__global__ void __launch_bounds__(m32::kThreads, 4) bp_m32_kernel(BpLaunch L)
{
    using namespace m32;
    const int CH = L.chunk;
    extern __shared__ float4 lds4[];
    unsigned char* lds = (unsigned char*)lds4;
    uint4* sel = (uint4*)lds;
    float4* ang = (float4*)(lds + kHead);                        // [CH] (o, cos, sin, s0)
    int* rowv = (int*)(lds + kHead + 16 * CH);                   // [CH] row index
    const int blk_off = kHead + 16 * CH + ((4 * CH + 15) & ~15);
    float4* blk = (float4*)(lds + blk_off);                      // [CH][32] (D, 2 j0, D - 8, 2 j0 + 16)
    const int data_off = blk_off + 512 * CH;                     // [CH][32 slices][kPitch]
    int* item_sh = (int*)(lds + kItem);       // in the dynamic block, so that it starts at LDS address 0

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, jl = lane & 31;
    const float xl = (float)(jl & 7), yl = (float)(jl >> 3);

    // selectors of v_perm (0x0c: zero byte): entry e places (1 - f, f) at halves
    // (p, p + 1) of the lane's eight, p = e - 9 (zero where neither falls inside)
    if (tid < kSel) {
        const unsigned int Z = 0x0c0c0c0cu, LO = 0x03020100u, SHL = 0x01000c0cu, SHR = 0x0c0c0302u;
        unsigned int d[4] = {Z, Z, Z, Z};
        const int p = tid - 9;
        if (p == -1) d[0] = SHR;
        if (p >= 0 && p <= 7) {
            if ((p & 1) == 0) d[p >> 1] = LO;
            else {
                d[p >> 1] = SHL;
                if ((p >> 1) + 1 < 4) d[(p >> 1) + 1] = SHR;
            }
        }
        sel[tid] = make_uint4(d[0], d[1], d[2], d[3]);
    }

    int iter = 0, probe = 0;
    while (true) {
        int item;
        if (L.counter) {
            if (tid == 0) *item_sh = next_work_item(L, probe);
            __syncthreads();
            item = *item_sh;
            __syncthreads();
        } else {
            item = (iter == 0) ? (int)blockIdx.x : L.total_items;
            __syncthreads();                   // the selector table is written
        }
        iter++;
        if (item >= L.total_items) break;
        if (L.item_order) item = L.item_order[item];
        int lo = 0, hi = L.nslices - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (L.slices[mid].tile_base <= item) lo = mid; else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const BpSlice& S = L.slices[lo];
        const int t = item - S.tile_base;
        const int tiles_x = S.tiles_x;
        const int X0 = (t % tiles_x) * kTile, Y0 = (t / tiles_x) * kTile;
        const int NCH = (int)S.row_stride;
        const uint4* packed = (const uint4*)S.rows;
        const int nang = S.nang;

        float16v acc[kBlocksPerWave];
#pragma unroll
        for (int b = 0; b < kBlocksPerWave; b++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[b][r] = 0.f;

        const int w0 = L.ang_end > 0 ? min(L.ang_begin, nang) : 0;
        const int w1 = L.ang_end > 0 ? min(L.ang_end, nang) : nang;
        for (int a0 = w0; a0 < w1; a0 += CH) {
            // ---- (1) per (angle, block): window start and coordinate offset
            for (int idx = tid; idx < CH * 32; idx += kThreads) {
                const int a = idx >> 5, b = idx & 31;
                const int ga = a0 + a;
                float o = 0.f, c = 0.f, s = 0.f;
                int s0 = -(1 << 28), row = -1;
                if (ga < w1) {
                    const AngleEntry e = S.ang[ga];
                    const double ut = e.u0 + (double)X0 * e.dux + (double)Y0 * e.duy;
                    const double umin = ut + fmin(0.0, (kTile - 1) * e.dux) + fmin(0.0, (kTile - 1) * e.duy);
                    s0 = ((int)floor(umin - 1.0)) & ~7;
                    o = (float)(ut - (double)s0);
                    c = (float)e.dux;
                    s = (float)e.duy;
                    row = e.row;
                }
                const float off = fmaf((float)(4 * (b >> 2)), s, (float)(8 * (b & 3)) * c);
                const float umin_b = o + off + fminf(0.f, 7.f * c) + fminf(0.f, 3.f * s);
                int j0 = ((int)floorf(umin_b - 0.015625f)) & ~3;
                j0 = min(max(j0, 0), kMaxJ0);
                const float D = (float)(kShift - j0) + off;
                blk[idx] = make_float4(D, __int_as_float(2 * j0), D - 8.f, __int_as_float(2 * j0 + 16));
                if (b == 0) {
                    ang[a] = make_float4(o, c, s, __int_as_float(s0));
                    rowv[a] = row;
                }
            }
            __syncthreads();

            // ---- (2) stage the slices' segments: thread = (slice, 8-sample chunk, angle parity)
            {
                const int sl = tid & 31, cc = (tid >> 5) & 7;
                for (int a = tid >> 8; a < CH; a += kThreads / 256) {
                    const int row = rowv[a];
                    const int gc = (__float_as_int(ang[a].w) >> 3) + cc;
                    uint4 v = make_uint4(0u, 0u, 0u, 0u);
                    if (row >= 0 && gc >= 0 && gc < NCH) v = packed[((long long)row * NCH + gc) * 32 + sl];
                    uint2* dst = (uint2*)(lds + data_off + (a * 32 + sl) * kPitch + cc * 16);
                    dst[0] = make_uint2(v.x, v.y);
                    dst[1] = make_uint2(v.z, v.w);
                }
            }
            __syncthreads();

            // ---- (3) one MFMA per angle and block
            const unsigned char* arow = lds + data_off + jl * kPitch;
            // selector entry of integer part i at selb[i] (one shift-add per lookup)
            const uint4* selb = sel - (kShift - 9);
            const float2* bq = (const float2*)blk + 2 * kBlocksPerWave * wave + h;
            // per angle: the blocks' coordinates, then every selector and sample read,
            // then the next angle's tables (LDS reads complete in order, so these do not
            // delay this angle's operands), then the permutes and products
            float4 gn = ang[0];
            float2 qn[kBlocksPerWave];
#pragma unroll
            for (int b = 0; b < kBlocksPerWave; b++) qn[b] = bq[2 * b];
            for (int a = 0; a < CH; a++) {
                const float lc = fmaf(yl, gn.z, fmaf(xl, gn.y, gn.x));
                const unsigned char* aa = arow + a * 32 * kPitch;
                float fr[kBlocksPerWave];
                uint4 sv[kBlocksPerWave];
                uint2 av0[kBlocksPerWave], av1[kBlocksPerWave];
#pragma unroll
                for (int b = 0; b < kBlocksPerWave; b++) {
                    const float tt = lc + qn[b].x;                   // kShift + u - j0 - 8 h, on the 2^-11 grid
                    fr[b] = __builtin_amdgcn_fractf(tt);
                    // entry (int)tt - (kShift - 9): tt - kShift lies in [-8, 12) (window bounds)
                    sv[b] = selb[(int)tt];
                    const uint2* ap = (const uint2*)(aa + __float_as_int(qn[b].y));
                    av0[b] = ap[0];
                    av1[b] = ap[1];
                }
                if (a + 1 < CH) {
                    gn = ang[a + 1];
#pragma unroll
                    for (int b = 0; b < kBlocksPerWave; b++) qn[b] = bq[2 * ((a + 1) * 32 + b)];
                }
#pragma unroll
                for (int b = 0; b < kBlocksPerWave; b++) {
                    const float f = fr[b];
                    const unsigned int P = __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_pkrtz(1.f - f, f));
                    const uint4 bw = make_uint4(__builtin_amdgcn_perm(P, P, sv[b].x), __builtin_amdgcn_perm(P, P, sv[b].y),
                                                __builtin_amdgcn_perm(P, P, sv[b].z), __builtin_amdgcn_perm(P, P, sv[b].w));
                    const half8v A = __builtin_bit_cast(half8v, make_uint4(av0[b].x, av0[b].y, av1[b].x, av1[b].y));
                    const half8v B = __builtin_bit_cast(half8v, bw);
                    acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, acc[b], 0, 0, 0);
                }
            }
            __syncthreads();
        }

        // ---- write: lane l, register r holds slice 8 (r / 4) + 4 h + r % 4 of pixel jl
        const float scale = *S.gscale * S.out_scale;
        const int out_n = S.out_n;
#pragma unroll
        for (int b = 0; b < kBlocksPerWave; b++) {
            const int bb = kBlocksPerWave * wave + b;
            const int X = X0 + 8 * (bb & 3) + (jl & 7);
            const int Y = Y0 + 4 * (bb >> 2) + (jl >> 3);
            if (X >= out_n || Y >= out_n) continue;
            const long long o = (long long)Y * out_n + X;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int sl0 = 8 * (r >> 2) + (r & 3);
                float* o0 = (float*)(sl0 == 0 ? S.out : S.out_x[sl0 - 1]);
                float* o1 = (float*)S.out_x[sl0 + 3];
                float* out = h ? o1 : o0;
                const float v = acc[b][r] * scale;
                out[o] = S.accumulate ? out[o] + v : v;
            }
        }
        if (!L.counter) break;
    }
}

// BP_LINEAR_M32 rows: samples p_k of 32 slices times row scale / item scale in
// FP16, chunks of eight samples with the slices side by side; lanes = slices,
// so a wave writes 1 KB contiguously.
template <int STORE>
__global__ void __launch_bounds__(256) pack_m32_kernel(const PackJob* jobs)
{
    const PackJob* J = jobs + blockIdx.y;
    const int N = J->N;
    const int NCH = (int)bp_m32_chunks(N);
    const int nrows = J->nrows;
    const float inv = 1.f / *J->gscale;
    const int s = threadIdx.x & 31;
    const void* rows = J->rows[s];
    const float* scales = J->scales[s];
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const float mul = scales[row] * inv;
        const long long r0 = (long long)row * N;
        uint4* out = (uint4*)J->out + (long long)row * NCH * 32;
        for (int c = threadIdx.x >> 5; c < NCH; c += 8) {
            unsigned int w[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = 8 * c + 2 * q;
                const float v0 = k < N ? stored_val<STORE>(rows, r0 + k) * mul : 0.f;
                const float v1 = k + 1 < N ? stored_val<STORE>(rows, r0 + k + 1) * mul : 0.f;
                w[q] = __builtin_bit_cast(unsigned int, __halves2half2(__float2half_rn(v0), __float2half_rn(v1)));
            }
            out[c * 32 + s] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

template <int KIND>
static void launch_pack_kind(const PackJob* d_jobs, int njobs, int store, int max_rows, hipStream_t stream)
{
    int bx = 16384 / njobs;
    bx = bx < 1 ? 1 : bx > max_rows ? max_rows : bx;
    const dim3 g((unsigned)bx, (unsigned)njobs);
    if (store == 0) hipLaunchKernelGGL((pack_kernel<KIND, 0>), g, dim3(256), 0, stream, d_jobs);
    else if (store == 1) hipLaunchKernelGGL((pack_kernel<KIND, 1>), g, dim3(256), 0, stream, d_jobs);
    else hipLaunchKernelGGL((pack_kernel<KIND, 2>), g, dim3(256), 0, stream, d_jobs);
}

void launch_pack_scales(const PackJob* d_jobs, int njobs, hipStream_t stream)
{
    if (njobs > 0) hipLaunchKernelGGL(pack_scale_kernel, dim3((unsigned)njobs), dim3(1024), 0, stream, d_jobs);
}

void launch_pack(const PackJob* d_jobs, int njobs, int kind, int store, int max_rows, hipStream_t stream)
{
    if (njobs <= 0 || max_rows <= 0) return;
    if (kind == 32) {
        int bx = 16384 / njobs;
        bx = bx < 1 ? 1 : bx > max_rows ? max_rows : bx;
        const dim3 g((unsigned)bx, (unsigned)njobs);
        if (store == 1) hipLaunchKernelGGL(pack_m32_kernel<1>, g, dim3(256), 0, stream, d_jobs);
        else hipLaunchKernelGGL(pack_m32_kernel<2>, g, dim3(256), 0, stream, d_jobs);
        return;
    }
    if (kind == 0) launch_pack_kind<0>(d_jobs, njobs, store, max_rows, stream);
    else if (kind == 3) launch_pack_kind<3>(d_jobs, njobs, store, max_rows, stream);
    else if (kind == 2) launch_pack_kind<2>(d_jobs, njobs, store, max_rows, stream);
    else if (kind == 4) launch_pack_kind<4>(d_jobs, njobs, store, max_rows, stream);
    else if (kind == 8) launch_pack_kind<8>(d_jobs, njobs, store, max_rows, stream);
    else launch_pack_kind<16>(d_jobs, njobs, store, max_rows, stream);
}

// Tile configurations: id -> (PX, PY).  0: 64x64, 1: 64x128 (W x H), 2: 32x32,
// 3: 128x64, 4: 32x64.
int bp_tile_side(int tile)
{
    switch (tile) {
    case 1: return 64;
    case 2: return 32;
    case 3: return 128;
    case 4: return 32;
    default: return 64;
    }
}

int bp_tile_h(int tile)
{
    switch (tile) {
    case 1: return 128;
    case 2: return 32;
    case 3: return 64;
    case 4: return 64;
    default: return 64;
    }
}

static int tile_seg(int tile)
{
    switch (tile) {
    case 1: return TileCfg<4, 8>::SEG;
    case 2: return TileCfg<2, 2>::SEG;
    case 3: return TileCfg<8, 4>::SEG;
    case 4: return TileCfg<2, 4>::SEG;
    default: return TileCfg<4, 4>::SEG;
    }
}
static_assert(TileCfg<2, 4>::SEG == TileCfg<2, 2, 8>::SEG, "32x64 tiles stage the same segment");
static_assert(TileCfg<4, 4>::SEG == TileCfg<4, 2, 8>::SEG, "64x64 tiles stage the same segment");

int bp_default_chunk(int mode, int tile)
{
    if (mode == BP_LINEAR) return (tile == 2) ? 64 : 32;
    if (mode == BP_LINEAR2) return (tile == 2) ? 32 : 16;
    if (mode == BP_LINEAR4) return (tile == 2) ? 16 : 8;
    if (mode == BP_LINEAR_H) return (tile == 2) ? 64 : 32;
    if (mode == BP_LINEAR_H2) return (tile == 2) ? 64 : 32;
    if (mode == BP_LINEAR_H4) return (tile == 2) ? 32 : 16;
    if (mode == BP_LINEAR_H8) return (tile == 2) ? 32 : 24;
    if (mode == BP_LINEAR_H16) return 16;
    if (mode == BP_LINEAR_M32) return 16;
    return (tile == 2) ? 64 : 32;
}

int bp_max_chunk(int mode) { return mode == BP_LINEAR_H ? 64 : mode == BP_LINEAR_M32 ? 16 : 256; }

size_t bp_lds_bytes(int mode, int tile, int chunk)
{
    if (mode == BP_LINEAR_M32) return bp_m32_lds_bytes(chunk);
    const int seg = tile_seg(tile);
    const size_t head = (size_t)chunk * 16 + (size_t)((2 * chunk + 3) / 4) * 16;
    return head + (size_t)chunk * entry_bytes(mode, seg);
}

// Eight slices per item exist for tiles 0, 2 and 4 (the larger tiles would need
// 256 accumulators per thread), sixteen for tiles 2 and 4; the plan forms no
// such items otherwise.  On the 64x64 tile eight slices run eight waves per
// workgroup (8 pixels x 8 slices per lane), sixteen slices on the 32x64 tile as
// well (4 pixels x 16 slices), so that four waves fit per SIMD.
int bp_threads(int mode, int tile)
{
    return ((mode == BP_LINEAR_H8 && tile == 0) || (mode == BP_LINEAR_H16 && tile == 4) || mode == BP_LINEAR_M32)
               ? 512 : 256;
}

// Modes that read the tile geometry table: the FP32 pairs.  The kernel also
// builds the table form for single FP32 slices and FP16 pairs, but the sixteen-
// slice FP16 kernel measured slower with it (64-slice C2 stack: 26.4 vs 24.0 ms;
// it spills and the compiler's schedule changes), and singles are not measured.
bool bp_has_tgeom(int mode)
{
    return mode == BP_LINEAR2 || mode == BP_LINEAR4;
}

// Persistent launches that deal their work list to the XCDs in runs of
// neighbouring tiles.  On the 64-slice 2048^2 x 1800 stack the dealing raises
// the L2 hit rate of the FP32 pair kernel from 36% to 61% and of the
// sixteen-slice FP16 kernel from 55% to 75% (fetched bytes -40% and -48%), but
// only the wide FP16 items gain time from it (sixteen per item 23.5 -> 23.15 ms,
// 32 on the matrix cores 21.9 -> 21.7 ms).  The FP32 pairs lost 1.3-4% with runs
// of 32 to 1024 tiles, but those builds also carried the dealing's three VGPRs,
// which cost the pair kernel a wave per SIMD; the four-slice kernel, at four
// waves either way, measured neutral (64-slice stack +0.4%, uneven batch -0.3%);
// the mixed eight/four/two-slice FP16 batch lost 0.7%.  Those keep one queue (a
// single run of the whole list).
bool bp_deals_runs(int mode)
{
    return mode == BP_LINEAR_H16 || mode == BP_LINEAR_M32;      // bp_tiled_kernel and bp_m32_kernel agree
}

template <int MODE, bool TG = false>
static const void* kernel_ptr(int tile)
{
    if constexpr (MODE == BP_LINEAR_H8) {
        if (tile == 2) return (const void*)&bp_tiled_kernel<MODE, 2, 2, 4, TG>;
        if (tile == 4) return (const void*)&bp_tiled_kernel<MODE, 2, 4, 4, TG>;
        return tile == 0 ? (const void*)&bp_tiled_kernel<MODE, 4, 2, 8, TG> : nullptr;
    } else if constexpr (MODE == BP_LINEAR_H16) {
        if (tile == 2) return (const void*)&bp_tiled_kernel<MODE, 2, 2, 4, TG>;
        return tile == 4 ? (const void*)&bp_tiled_kernel<MODE, 2, 2, 8, TG> : nullptr;
    } else {
        switch (tile) {
        case 1: return (const void*)&bp_tiled_kernel<MODE, 4, 8, 4, TG>;
        case 2: return (const void*)&bp_tiled_kernel<MODE, 2, 2, 4, TG>;
        case 3: return (const void*)&bp_tiled_kernel<MODE, 8, 4, 4, TG>;
        case 4: return (const void*)&bp_tiled_kernel<MODE, 2, 4, 4, TG>;
        default: return (const void*)&bp_tiled_kernel<MODE, 4, 4, 4, TG>;
        }
    }
}

static const void* kernel_for(int mode, int tile, bool tg = false)
{
    if (mode == BP_LINEAR_M32) return tile == 2 ? (const void*)&bp_m32_kernel : nullptr;
    if (tg && mode == BP_LINEAR2) return kernel_ptr<BP_LINEAR2, true>(tile);
    if (tg && mode == BP_LINEAR4) return kernel_ptr<BP_LINEAR4, true>(tile);
    if (mode == BP_LINEAR) return kernel_ptr<BP_LINEAR>(tile);
    if (mode == BP_LINEAR2) return kernel_ptr<BP_LINEAR2>(tile);
    if (mode == BP_LINEAR4) return kernel_ptr<BP_LINEAR4>(tile);
    if (mode == BP_LINEAR_H) return kernel_ptr<BP_LINEAR_H>(tile);
    if (mode == BP_LINEAR_H2) return kernel_ptr<BP_LINEAR_H2>(tile);
    if (mode == BP_LINEAR_H4) return kernel_ptr<BP_LINEAR_H4>(tile);
    if (mode == BP_LINEAR_H8) return kernel_ptr<BP_LINEAR_H8>(tile);
    if (mode == BP_LINEAR_H16) return kernel_ptr<BP_LINEAR_H16>(tile);
    if (mode == BP_LUT_F32) return kernel_ptr<BP_LUT_F32>(tile);
    return kernel_ptr<BP_LUT_I32>(tile);
}

// Kernels that use more than 64 KiB of LDS must say so (once per kernel and
// size, so that a later launch inside a stream capture makes no such call).
static void allow_lds(const void* f, size_t lds)
{
    static const void* seen_f[64];
    static size_t seen_lds[64];
    static int nseen = 0;
    if (lds <= 65536) return;
    for (int i = 0; i < nseen; i++)
        if (seen_f[i] == f && seen_lds[i] >= lds) return;
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (nseen < 64) { seen_f[nseen] = f; seen_lds[nseen] = lds; nseen++; }
}

int bp_blocks_per_cu(int mode, int tile, int chunk)
{
    const void* f = kernel_for(mode, tile);
    const size_t lds = bp_lds_bytes(mode, tile, chunk);
    allow_lds(f, lds);
    int nb = 0;
    const int threads = bp_threads(mode, tile);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads, lds) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

int bp_blocks_per_cu_tg(int mode, int tile, int chunk, bool tgeom)
{
    if (!(tgeom && bp_has_tgeom(mode))) return bp_blocks_per_cu(mode, tile, chunk);
    const void* f = kernel_for(mode, tile, true);
    const size_t lds = bp_lds_bytes(mode, tile, chunk);
    allow_lds(f, lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, bp_threads(mode, tile), lds) != hipSuccess || nb < 1)
        nb = 1;
    return nb;
}

template <int MODE, bool TG>
static void launch_mode_tg(const BpLaunch& L, int grid, size_t lds, hipStream_t stream)
{
    allow_lds(kernel_ptr<MODE, TG>(L.tile), lds);
    if constexpr (MODE == BP_LINEAR_H8) {
        if (L.tile == 2) hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 2, 4, TG>), dim3(grid), dim3(256), lds, stream, L);
        else if (L.tile == 4) hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 4, 4, TG>), dim3(grid), dim3(256), lds, stream, L);
        else if (L.tile == 0)
            hipLaunchKernelGGL((bp_tiled_kernel<MODE, 4, 2, 8, TG>), dim3(grid), dim3(512), lds, stream, L);
    } else if constexpr (MODE == BP_LINEAR_H16) {
        if (L.tile == 2) hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 2, 4, TG>), dim3(grid), dim3(256), lds, stream, L);
        else if (L.tile == 4)
            hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 2, 8, TG>), dim3(grid), dim3(512), lds, stream, L);
    } else {
        switch (L.tile) {
        case 1: hipLaunchKernelGGL((bp_tiled_kernel<MODE, 4, 8, 4, TG>), dim3(grid), dim3(256), lds, stream, L); break;
        case 2: hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 2, 4, TG>), dim3(grid), dim3(256), lds, stream, L); break;
        case 3: hipLaunchKernelGGL((bp_tiled_kernel<MODE, 8, 4, 4, TG>), dim3(grid), dim3(256), lds, stream, L); break;
        case 4: hipLaunchKernelGGL((bp_tiled_kernel<MODE, 2, 4, 4, TG>), dim3(grid), dim3(256), lds, stream, L); break;
        default: hipLaunchKernelGGL((bp_tiled_kernel<MODE, 4, 4, 4, TG>), dim3(grid), dim3(256), lds, stream, L); break;
        }
    }
}

template <int MODE>
static void launch_mode(const BpLaunch& L, int grid, hipStream_t stream)
{
    const size_t lds = bp_lds_bytes(MODE, L.tile, L.chunk);
    if constexpr (MODE == BP_LINEAR2 || MODE == BP_LINEAR4) {
        if (L.tgeom) {
            launch_mode_tg<MODE, true>(L, grid, lds, stream);
            return;
        }
    }
    launch_mode_tg<MODE, false>(L, grid, lds, stream);
}

__global__ void __launch_bounds__(64) ptr_patch_kernel(PtrPatch P)
{
    const int i = threadIdx.x;
    if (i < P.n) *P.dst[i] = P.val[i];
}

void launch_ptr_patch(const PtrPatch& P, hipStream_t stream)
{
    if (P.n > 0) hipLaunchKernelGGL(ptr_patch_kernel, dim3(1), dim3(64), 0, stream, P);
}

void launch_bp(const BpLaunch& L, int grid, hipStream_t stream)
{
    if (L.mode == BP_LINEAR_M32) {
        if (L.tile != 2 || L.chunk < 1 || L.chunk > 16) return;     // the plan forms no other
        const size_t lds = bp_m32_lds_bytes(L.chunk);
        allow_lds((const void*)&bp_m32_kernel, lds);
        hipLaunchKernelGGL(bp_m32_kernel, dim3(grid), dim3(m32::kThreads), lds, stream, L);
        return;
    }
    if (L.mode == BP_LINEAR) launch_mode<BP_LINEAR>(L, grid, stream);
    else if (L.mode == BP_LINEAR2) launch_mode<BP_LINEAR2>(L, grid, stream);
    else if (L.mode == BP_LINEAR4) launch_mode<BP_LINEAR4>(L, grid, stream);
    else if (L.mode == BP_LINEAR_H) launch_mode<BP_LINEAR_H>(L, grid, stream);
    else if (L.mode == BP_LINEAR_H2) launch_mode<BP_LINEAR_H2>(L, grid, stream);
    else if (L.mode == BP_LINEAR_H4) launch_mode<BP_LINEAR_H4>(L, grid, stream);
    else if (L.mode == BP_LINEAR_H8) launch_mode<BP_LINEAR_H8>(L, grid, stream);
    else if (L.mode == BP_LINEAR_H16) launch_mode<BP_LINEAR_H16>(L, grid, stream);
    else if (L.mode == BP_LUT_F32) launch_mode<BP_LUT_F32>(L, grid, stream);
    else launch_mode<BP_LUT_I32>(L, grid, stream);
}

}  // namespace rv
