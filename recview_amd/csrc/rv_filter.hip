// rv_filter.hip -- sinogram ramp-filter stage as one LDS-resident FFT per row.
//
// Replaces the per-projection CPU filter of cudaReconstHost.cpp:731-756 (row
// prep, CFft::FFT1Rev, multiply by fFilter, CFft::FFT1, crop) and the CUDAFFT
// pipeline of cudaReconst.cu:123-192.  One workgroup owns one row: the padded
// complex row (n <= 16384 points, 8n bytes) lives in LDS for the whole
// transform, butterflies are register-staged so a single LDS buffer serves
// every stage, and the filter multiply and the crop are fused into the last
// stage of each transform.
//
// The butterfly network, twiddles and rounding order are those of CFft's
// radix-2 Stockham passes (cfft.cpp:193-369): stage l reads x[b] and
// x[b + n/2] for butterfly b = j*al + m and writes x[2*j*al + m] and
// x[2*j*al + m + al].  Keeping the same network (this file is compiled with
// -ffp-contract=off) makes the filtered rows bit-identical to the CPU
// restatement, which is what lets the LUT16 integer path match the reference
// bit for bit.
#include "rv_internal.h"
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <map>
#include <mutex>

namespace rv {

constexpr int kMaxBfly = 8;  // butterflies per thread and stage (n/2 <= 8 * blockDim)

template <int IN_KIND>
__device__ __forceinline__ float load_in(const void* base, long long off)
{
    if constexpr (IN_KIND == 0) return (float)((const int16_t*)base)[off];
    else return ((const float*)base)[off];
}

// Value of padded position m (cudaReconstHost.cpp:734-751 for pad_mode 0,
// cudaReconst.cu:124-134 after the memset of cu:174 for pad_mode 1).
template <int IN_KIND>
__device__ float prep_value(const FilterJob& J, const void* row, int m)
{
    const int z = J.z, W = J.W, ic = J.ic;
    const int rel = m - J.place;
    if (z == 1) {
        // no zoom: sample ic + rel, edge-replicated (pad_mode 0) or zero outside
        const int k = ic + rel;
        if (k >= 0 && k < W) return load_in<IN_KIND>(row, k);
        if (J.pad_mode != 0) return 0.f;
        return load_in<IN_KIND>(row, k < 0 ? 0 : W - 1);
    }
    const int kk = (rel >= 0) ? rel / z : -((-rel + z - 1) / z);
    const int j = rel - kk * z;
    const int k = ic + kk;
    if (J.pad_mode == 0) {
        const int idx0 = (0 - ic) * z + J.place;
        const int idx1 = (W - 1 - ic) * z + J.place + 1;
        if (m < idx0) return load_in<IN_KIND>(row, 0);
        if (m >= idx1) return load_in<IN_KIND>(row, W - 1);
        if (j == 0) return (k >= 0 && k <= W - 1) ? load_in<IN_KIND>(row, k) : 0.f;
        if (k >= 0 && k < W - 1 && m - j >= 0) {
            if constexpr (IN_KIND == 0) {
                const int s0 = ((const int16_t*)row)[k], s1 = ((const int16_t*)row)[k + 1];
                return (float)(s0 * (z - j) / z + s1 * j / z);
            } else {
                const float s0 = ((const float*)row)[k], s1 = ((const float*)row)[k + 1];
                return (s0 * (float)(z - j) + s1 * (float)j) / (float)z;
            }
        }
        return 0.f;
    } else {
        if (j == 0) return (k >= 0 && k <= W - 1) ? load_in<IN_KIND>(row, k) : 0.f;
        if (k >= 0 && k < W - 1) {
            if constexpr (IN_KIND == 0) {
                const int s0 = ((const int16_t*)row)[k], s1 = ((const int16_t*)row)[k + 1];
                return (float)(s0 * (z - j) + s1 * j) / z;
            } else {
                const float s0 = ((const float*)row)[k], s1 = ((const float*)row)[k + 1];
                return (s0 * (float)(z - j) + s1 * (float)j) / (float)z;
            }
        }
        return 0.f;
    }
}

__device__ __forceinline__ float2 cmul_tw(float2 b, float2 w)
{
    // scrr = br*wr - bi*wi ; scri = br*wi + bi*wr  (cfft.cpp:239-240)
    return make_float2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
}

// Middle Stockham stage l (1 <= l <= np-2) in place, register staged.
__device__ __forceinline__ void fft_mid_stage(float2* buf, const float2* tw, int half, int l, int T,
                                              int tid)
{
    const int al = 1 << l, bt = half >> l;
    float2 A[kMaxBfly], B[kMaxBfly];
#pragma unroll
    for (int i = 0; i < kMaxBfly; i++) {
        const int b = tid + i * T;
        if (b < half) { A[i] = buf[b]; B[i] = buf[b + half]; }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMaxBfly; i++) {
        const int b = tid + i * T;
        if (b < half) {
            const int j = b >> l, m = b & (al - 1);
            const float2 s = cmul_tw(B[i], tw[m * bt]);
            const int o = (j << (l + 1)) + m;
            buf[o] = make_float2(A[i].x + s.x, A[i].y + s.y);
            buf[o + al] = make_float2(A[i].x - s.x, A[i].y - s.y);
        }
    }
    __syncthreads();
}

// First stage: x1[2j] = x[i] + x[i+n/2], x1[2j+1] = x[i] - x[i+n/2], with
// i = n/2-1-j for FFTrev (cfft.cpp:199-211) and i = j for FFT (:284-294).
template <bool REV>
__device__ __forceinline__ void fft_first_stage(float2* buf, int half, int T, int tid)
{
    float2 A[kMaxBfly], B[kMaxBfly];
#pragma unroll
    for (int i = 0; i < kMaxBfly; i++) {
        const int j = tid + i * T;
        if (j < half) {
            const int src = REV ? (half - 1 - j) : j;
            A[i] = buf[src];
            B[i] = buf[src + half];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMaxBfly; i++) {
        const int j = tid + i * T;
        if (j < half) {
            buf[2 * j] = make_float2(A[i].x + B[i].x, A[i].y + B[i].y);
            buf[2 * j + 1] = make_float2(A[i].x - B[i].x, A[i].y - B[i].y);
        }
    }
    __syncthreads();
}

// Per-row power-of-two scale so that |v| / scale <= 2^14, then the 16-bit store
// of the N cropped values held in outf (LDS).
__device__ void store_lp_row(const FilterJob& J, void* out_lp, float* out_scale, const float* outf, int rout,
                             float* red)
{
    const int T = blockDim.x, tid = threadIdx.x;
    __syncthreads();
    float mx = 0.f;
    for (int c = tid; c < J.N; c += T) mx = fmaxf(mx, fabsf(outf[c]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        float m2 = 0.f;
        for (int w = 0; w < (T + 63) / 64; w++) m2 = fmaxf(m2, red[w]);
        int e = 0;
        if (m2 > 0.f) { frexpf(m2 / 16384.f, &e); }
        red[16] = ldexpf(1.f, e);
    }
    __syncthreads();
    const float sc = red[16];
    const float inv = 1.f / sc;            // power of two: exact
    if (tid == 0) out_scale[rout] = sc;
    if (J.store == 1) {
        __half* o = (__half*)out_lp + (long long)rout * J.out_stride;
        for (int c = tid; c < J.N; c += T) o[c] = __float2half_rn(outf[c] * inv);
    } else {
        __hip_bfloat16* o = (__hip_bfloat16*)out_lp + (long long)rout * J.out_stride;
        for (int c = tid; c < J.N; c += T) o[c] = __float2bfloat16(outf[c] * inv);
    }
    __syncthreads();
}

constexpr int fft_threads(int np) { return (1 << np) / 16 < 64 ? 64 : (1 << np) / 16; }

// NP: log2 of the transform length (compile time, so every stage's strides
// and twiddle steps are constants); launched with fft_threads(NP) threads.
template <int IN_KIND, int NP>
__global__ void __launch_bounds__(1024) filter_rows_kernel(FilterJob J, FilterBatch FB)
{
    extern __shared__ float2 buf[];
    const int r = blockIdx.x, bi = blockIdx.y;
    const void* sino = FB.sino[bi];
    float* out = FB.out[bi];
    const int rin = J.row_in ? J.row_in[r] : r;
    const int rout = J.row_out ? J.row_out[r] : r;
    constexpr int n = 1 << NP, half = n >> 1, np = NP, T = fft_threads(NP);
    const int tid = threadIdx.x;
    const float2* tw = (const float2*)J.tw;
    const void* row = (IN_KIND == 0)
        ? (const void*)((const int16_t*)sino + (long long)rin * J.in_stride)
        : (const void*)((const float*)sino + (long long)rin * J.in_stride);

    // Row preparation straight into LDS.
    for (int m = tid; m < n; m += T) buf[m] = make_float2(prep_value<IN_KIND>(J, row, m), 0.f);
    __syncthreads();

    // ---- FFT1Rev (cfft.cpp:431-434 -> FFTrev with nmin < 0)
    fft_first_stage<true>(buf, half, T, tid);
#pragma unroll
    for (int l = 1; l < np - 1; l++) fft_mid_stage(buf, tw, half, l, T, tid);
    {
        // last stage (cfft.cpp:249-268) fused with p[k] *= fFilter[k]
        // (cudaReconstHost.cpp:753)
        float2 A[kMaxBfly], B[kMaxBfly];
#pragma unroll
        for (int i = 0; i < kMaxBfly; i++) {
            const int k = tid + i * T;
            if (k < half) { A[i] = buf[k]; B[i] = buf[half + k]; }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kMaxBfly; i++) {
            const int k = tid + i * T;
            if (k < half) {
                const float2 s = (k == 0) ? B[i] : cmul_tw(B[i], tw[k]);
                const int ip = (k == 0) ? 0 : (n - k);
                const int im = half - k;
                const float hp = J.filt[ip], hm = J.filt[im];
                const float pr = A[i].x + s.x, pi = A[i].y + s.y;
                const float mr = A[i].x - s.x, mi = A[i].y - s.y;
                buf[ip] = make_float2(pr * hp, pi * hp);
                buf[im] = make_float2(mr * hm, mi * hm);
            }
        }
        __syncthreads();
    }

    // ---- FFT1 (cfft.cpp:412-429 -> FFT with nmin < 0, then reverse and 1/n)
    fft_first_stage<false>(buf, half, T, tid);
#pragma unroll
    for (int l = 1; l < np - 1; l++) fft_mid_stage(buf, tw, half, l, T, tid);
    {
        float2 A[kMaxBfly], B[kMaxBfly];
#pragma unroll
        for (int i = 0; i < kMaxBfly; i++) {
            const int k = tid + i * T;
            if (k < half) { A[i] = buf[k]; B[i] = buf[half + k]; }
        }
        __syncthreads();
        const float a = 1.f / (float)n;
        float* outf = (float*)buf;                 // reuse LDS as the cropped real row
        const bool lp = (J.store != 0);
#pragma unroll
        for (int i = 0; i < kMaxBfly; i++) {
            const int k = tid + i * T;
            if (k < half) {
                const float2 s = cmul_tw(B[i], tw[k]);
                // f[k] = x1[k] - scr, f[k + n/2] = x1[k] + scr (cfft.cpp:364-365);
                // FFT1 then stores f[n-1-q] * (1/n) at q (cfft.cpp:421-425) and
                // the caller keeps p[c + crop].re (cudaReconstHost.cpp:756).
                const float v0 = (A[i].x - s.x) * a;   // f[k]
                const float v1 = (A[i].x + s.x) * a;   // f[k + half]
                const int c0 = n - 1 - k - J.crop;
                const int c1 = n - 1 - (k + half) - J.crop;
                if (!lp) {
                    float* o = out + (long long)rout * J.out_stride;
                    if (c0 >= 0 && c0 < J.N) o[c0] = v0;
                    if (c1 >= 0 && c1 < J.N) o[c1] = v1;
                } else {
                    if (c0 >= 0 && c0 < J.N) outf[c0] = v0;
                    if (c1 >= 0 && c1 < J.N) outf[c1] = v1;
                }
            }
        }
        if (lp) {
            __shared__ float red[32];
            store_lp_row(J, FB.out_lp[bi], FB.out_scale[bi], outf, rout, red);
        }
    }
}

// ------------------------------------------------ transforms longer than LDS
// n = 2^L above 16384 points: the complex row (8n bytes) no longer fits in LDS.
// CFft's radix-2 Stockham network is then run in two passes per transform
// through a work buffer, with every butterfly -- inputs, twiddle entry, operation
// order -- that of the one-pass network, so the filtered rows stay bit-identical
// to the reference filter.  Stage l maps the index (c | j | m) -- c the top bit,
// m the low l bits -- to (j | c' | m).  Hence:
//  * pass A, stages 0..K-1, only combines elements that share the low L-K bits
//    j0 of their stage-0 index: 2^(L-K) independent sub-networks of 2^K elements
//    at stride 2^(L-K), each ending in the contiguous block j0 * 2^K + (0..2^K);
//  * pass B, stages K..L-1, never moves the low K bits r of an index: 2^K
//    sub-networks of the elements r + q * 2^K, where local stage l' is global
//    stage l' + K with twiddle index ((m << K) + r) * (n/2 >> (l' + K)).
// The last global stage keeps its fused operations: the filter multiply and the
// reversed output order of FFT1Rev (cfft.cpp:249-268), the crop of FFT1.
// A workgroup runs 2048 / 2^LS sub-networks (adjacent j0 or r, so global reads
// and writes coalesce); LDS holds them element-major with an odd pitch.
constexpr int kLgBlock = 2048;
constexpr int kLgThreads = 256;

struct LgPass {
    FilterJob J;
    const float2* src;     // work rows of the batch (passes 1-3)
    float2* dst;           // work rows (passes 0-2); pass 3: FP32 rows of 16-bit stores
    int row0;              // first row of the batch
    int L, K;              // log2 n; stages of pass A
};

// PASS 0: FFT1Rev stages 0..K-1 with row prep and the reversed first-stage read
// (FFTrev reads x[n/2-1-j], cfft.cpp:199-211); 1: FFT1Rev stages K..L-1 and the
// filter; 2: FFT1 stages 0..K-1; 3: FFT1 stages K..L-1, 1/n and the crop.
template <int PASS, int LS, int IN_KIND>
__global__ void __launch_bounds__(kLgThreads) fft_large_kernel(LgPass P)
{
    constexpr int NS = 1 << LS, G = kLgBlock >> LS, PITCH = G + 1;
    constexpr bool passA = (PASS == 0 || PASS == 2);
    constexpr int BF = kLgBlock / 2 / kLgThreads;           // butterflies per thread and stage
    __shared__ float2 buf[NS * PITCH];
    const FilterJob& J = P.J;
    const int L = P.L, K = P.K;
    const int n = 1 << L, half = n >> 1;
    const int r = P.row0 + (int)blockIdx.y;
    const size_t wrow = (size_t)blockIdx.y * n;
    const int tid = threadIdx.x;
    const int gbase = (int)blockIdx.x * G;                  // first j0 (pass A) or residue (pass B)
    const int estride = passA ? (1 << (L - K)) : (1 << K);
    const float2* tw = (const float2*)J.tw;

    if constexpr (PASS == 0) {
        const int rin = J.row_in ? J.row_in[r] : r;
        const void* row = (IN_KIND == 0)
            ? (const void*)((const int16_t*)J.sino + (long long)rin * J.in_stride)
            : (const void*)((const float*)J.sino + (long long)rin * J.in_stride);
        for (int idx = tid; idx < kLgBlock; idx += kLgThreads) {
            const int g = idx % G, e = idx / G;
            const int i = e * estride + gbase + g;
            const int m = (i < half) ? half - 1 - i : n + half - 1 - i;
            buf[e * PITCH + g] = make_float2(prep_value<IN_KIND>(J, row, m), 0.f);
        }
    } else {
        for (int idx = tid; idx < kLgBlock; idx += kLgThreads) {
            const int g = idx % G, e = idx / G;
            buf[e * PITCH + g] = P.src[wrow + (size_t)e * estride + gbase + g];
        }
    }
    __syncthreads();

    constexpr int NSTAGE = passA ? LS : LS - 1;            // pass B ends with the fused last stage
#pragma unroll 1
    for (int l = 0; l < NSTAGE; l++) {
        const int al = 1 << l;
        const int lg = passA ? l : l + K;                  // global stage
        const int bt = half >> lg;
        float2 A[BF], B[BF];
#pragma unroll
        for (int i = 0; i < BF; i++) {
            const int idx = tid + i * kLgThreads;
            const int g = idx % G, bl = idx / G;
            A[i] = buf[bl * PITCH + g];
            B[i] = buf[(bl + NS / 2) * PITCH + g];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < BF; i++) {
            const int idx = tid + i * kLgThreads;
            const int g = idx % G, bl = idx / G;
            const int j = bl >> l, m = bl & (al - 1);
            float2 sv = B[i];
            if (lg > 0) {
                const int mg = passA ? m : (m << K) + gbase + g;
                sv = cmul_tw(B[i], tw[mg * bt]);
            }
            const int o = (j << (l + 1)) + m;
            buf[o * PITCH + g] = make_float2(A[i].x + sv.x, A[i].y + sv.y);
            buf[(o + al) * PITCH + g] = make_float2(A[i].x - sv.x, A[i].y - sv.y);
        }
        __syncthreads();
    }

    if constexpr (passA) {
        // after K stages sub-network j0 holds the contiguous block j0 * 2^K + e
        for (int idx = tid; idx < kLgBlock; idx += kLgThreads) {
            const int e = idx % NS, g = idx / NS;
            P.dst[wrow + (size_t)(gbase + g) * NS + e] = buf[e * PITCH + g];
        }
    } else {
        // global last stage: butterfly k = (bl << K) + r reads x[k] and x[k + n/2]
#pragma unroll
        for (int i = 0; i < BF; i++) {
            const int idx = tid + i * kLgThreads;
            const int g = idx % G, bl = idx / G;
            const int k = (bl << K) + gbase + g;
            const float2 a = buf[bl * PITCH + g];
            const float2 b = buf[(bl + NS / 2) * PITCH + g];
            if constexpr (PASS == 1) {
                const float2 sv = (k == 0) ? b : cmul_tw(b, tw[k]);
                const int ip = (k == 0) ? 0 : (n - k);
                const int im = half - k;
                const float hp = J.filt[ip], hm = J.filt[im];
                const float pr = a.x + sv.x, pi = a.y + sv.y;
                const float mr = a.x - sv.x, mi = a.y - sv.y;
                P.dst[wrow + ip] = make_float2(pr * hp, pi * hp);
                P.dst[wrow + im] = make_float2(mr * hm, mi * hm);
            } else {
                const float sc = 1.f / (float)n;
                const float2 sv = cmul_tw(b, tw[k]);
                const float v0 = (a.x - sv.x) * sc;        // f[k]
                const float v1 = (a.x + sv.x) * sc;        // f[k + n/2]
                const int c0 = n - 1 - k - J.crop;
                const int c1 = n - 1 - (k + half) - J.crop;
                float* o;
                if (J.store == 0) {
                    const int rout = J.row_out ? J.row_out[r] : r;
                    o = J.out + (long long)rout * J.out_stride;
                } else {
                    o = (float*)(P.dst + wrow);             // FP32 row, scaled and stored below
                }
                if (c0 >= 0 && c0 < J.N) o[c0] = v0;
                if (c1 >= 0 && c1 < J.N) o[c1] = v1;
            }
        }
    }
}

// 16-bit stores of the two-pass filter: per-row scale and conversion of the FP32 rows
__global__ void __launch_bounds__(kLgThreads) lp_rows_kernel(FilterJob J, const float2* stage, int row0, int n)
{
    const int r = row0 + (int)blockIdx.x;
    const int rout = J.row_out ? J.row_out[r] : r;
    __shared__ float red[32];
    store_lp_row(J, J.out_lp, J.out_scale, (const float*)(stage + (size_t)blockIdx.x * n), rout, red);
}

template <int PASS, int IN_KIND>
static void launch_lg(const LgPass& P, int ls, int nb, hipStream_t st)
{
    const dim3 grid((unsigned)((1 << P.L) / kLgBlock), (unsigned)nb), block(kLgThreads);
    switch (ls) {
    case 7: hipLaunchKernelGGL((fft_large_kernel<PASS, 7, IN_KIND>), grid, block, 0, st, P); break;
    case 8: hipLaunchKernelGGL((fft_large_kernel<PASS, 8, IN_KIND>), grid, block, 0, st, P); break;
    case 9: hipLaunchKernelGGL((fft_large_kernel<PASS, 9, IN_KIND>), grid, block, 0, st, P); break;
    default: hipLaunchKernelGGL((fft_large_kernel<PASS, 10, IN_KIND>), grid, block, 0, st, P); break;
    }
}

size_t filter_large_rows(int n, size_t budget_bytes, int nrows)
{
    size_t rows = budget_bytes / (2 * sizeof(float2) * (size_t)n);
    if (rows < 1) rows = 1;
    if (nrows > 0 && rows > (size_t)nrows) rows = (size_t)nrows;
    return rows;
}

hipError_t launch_filter_large(const FilterJob& J, void* work, size_t work_rows, hipStream_t st)
{
    if (J.nrows <= 0) return hipSuccess;
    if (J.pair || J.np <= kMaxLdsFftLog2 || J.np > kMaxFftLog2 || J.n != (1 << J.np) || !work || work_rows < 1)
        return hipErrorInvalidValue;
    const int L = J.np, K = L / 2, n = J.n;
    float2* w0 = (float2*)work;
    float2* w1 = w0 + work_rows * (size_t)n;
    for (int row0 = 0; row0 < J.nrows; row0 += (int)work_rows) {
        const int nb = (int)std::min<size_t>(work_rows, (size_t)(J.nrows - row0));
        LgPass P{};
        P.J = J;
        P.row0 = row0;
        P.L = L;
        P.K = K;
        P.src = nullptr; P.dst = w0;
        if (J.in_kind == 0) launch_lg<0, 0>(P, K, nb, st);
        else launch_lg<0, 1>(P, K, nb, st);
        P.src = w0; P.dst = w1;
        launch_lg<1, 1>(P, L - K, nb, st);
        P.src = w1; P.dst = w0;
        launch_lg<2, 1>(P, K, nb, st);
        P.src = w0; P.dst = w1;
        launch_lg<3, 1>(P, L - K, nb, st);
        if (J.store != 0)
            hipLaunchKernelGGL(lp_rows_kernel, dim3((unsigned)nb), dim3(kLgThreads), 0, st, J, (const float2*)w1,
                               row0, n);
    }
    return hipGetLastError();
}

// ------------------------------------------------ paired rows, radix-16 transform
// Production filter for the LINEAR modes: rows a and b become the real and
// imaginary parts of one complex row; with the Hermitian-symmetrised filter
// Hs[k] = (H[k] + H[n-k]) / 2 (real and even) the result is
//   y = IFFT(Hs * FFT(a + i b)) = a (*) h + i (b (*) h),  h = IFFT(Hs) = h_re,
// the same circular convolution the CFft pipeline computes (one-sided filter,
// real part), in the standard DFT convention.  The transforms are Stockham
// autosort passes after Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors" (SC08): a pass of radix R and sub-length
// Ns reads x[j + r*n/R], twiddles by exp(-2 pi i r (j mod Ns) / (Ns R)), does an
// R-point DFT and writes x'[(j / Ns) * Ns * R + j mod Ns + r * Ns].  Passes are
// radix 16 (a thread's 16 points in registers, 4 x 4 Cooley-Tukey inside), the
// last one radix 8, 4 or 2 when log2 n is not a multiple of 4: a 4096-point
// transform is three LDS round trips.  The inverse is conj(FFT(conj(.))) / n,
// the conjugation folded into the filter multiply.
__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// LDS position of element i of a transform: inside each 16-element block the
// element index is XORed with the low four bits of the block index, and the
// blocks of each pair swap when bit 4 of the block index is set.  A wave's reads
// at stride n/R and writes at stride Ns (every pass of every length up to 2^14),
// and its contiguous loads and stores, then meet each LDS bank the minimum number
// of times for 64 8-byte accesses, with 32 or 64 banks.
__device__ __forceinline__ int lds_pos(int i)
{
    const int blk = i >> 4;
    return ((blk ^ ((blk >> 4) & 1)) << 4) | ((i & 15) ^ (blk & 15));
}

// 4-point forward DFT in place (W4 = -i)
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3)
{
    const float2 t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = csub(x1, x3);
    x0 = cadd(t0, t2);
    x1 = make_float2(t1.x + t3.y, t1.y - t3.x);          // t1 - i t3
    x2 = csub(t0, t2);
    x3 = make_float2(t1.x - t3.y, t1.y + t3.x);          // t1 + i t3
}

// R-point forward DFT of v in place; output k sits in v[dft_slot<R>(k)].
// R = 16: input p = q + 4 s; 4-point DFTs over s, twiddles W16^(q k1), 4-point
// DFTs over q, X[k1 + 4 k2] in v[4 k1 + k2].  R = 8: p = q + 2 s, twiddles
// W8^(q k1), X[k1 + 4 k2] in v[2 k1 + k2].
template <int R>
__host__ __device__ constexpr int dft_slot(int k)
{
    return R == 16 ? 4 * (k & 3) + (k >> 2) : R == 8 ? 2 * (k & 3) + (k >> 2) : k;
}

template <int R>
__device__ __forceinline__ void dft_r(float2 (&v)[R])
{
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f, RH = 0.70710678118654752f;
    if constexpr (R == 16) {
#pragma unroll
        for (int q = 0; q < 4; q++) dft4(v[q], v[q + 4], v[q + 8], v[q + 12]);
        // W16^m = (cos 2 pi m / 16, -sin 2 pi m / 16) for m = q k1
        v[5] = cmulf(v[5], make_float2(C1, -S1));        // q 1, k1 1: m 1
        v[9] = cmulf(v[9], make_float2(RH, -RH));        // q 1, k1 2: m 2
        v[13] = cmulf(v[13], make_float2(S1, -C1));      // q 1, k1 3: m 3
        v[6] = cmulf(v[6], make_float2(RH, -RH));        // q 2, k1 1: m 2
        v[10] = make_float2(v[10].y, -v[10].x);          // q 2, k1 2: m 4 (-i)
        v[14] = cmulf(v[14], make_float2(-RH, -RH));     // q 2, k1 3: m 6
        v[7] = cmulf(v[7], make_float2(S1, -C1));        // q 3, k1 1: m 3
        v[11] = cmulf(v[11], make_float2(-RH, -RH));     // q 3, k1 2: m 6
        v[15] = cmulf(v[15], make_float2(-C1, S1));      // q 3, k1 3: m 9
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    } else if constexpr (R == 8) {
        dft4(v[0], v[2], v[4], v[6]);
        dft4(v[1], v[3], v[5], v[7]);
        v[3] = cmulf(v[3], make_float2(RH, -RH));        // k1 1: W8^1
        v[5] = make_float2(v[5].y, -v[5].x);             // k1 2: W8^2 = -i
        v[7] = cmulf(v[7], make_float2(-RH, -RH));       // k1 3: W8^3
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            const float2 a = v[2 * k1], b = v[2 * k1 + 1];
            v[2 * k1] = cadd(a, b);
            v[2 * k1 + 1] = csub(a, b);
        }
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
}

// Element access of a pass: the transform's points in LDS (positions lds_pos).
struct LdsLoad {
    const float2* buf;
    __device__ __forceinline__ float2 operator()(int m) const { return buf[lds_pos(m)]; }
};
struct LdsStore {
    float2* buf;
    __device__ __forceinline__ void operator()(int, int, int k, float2 v) const { buf[lds_pos(k)] = v; }
};

// One Stockham pass of radix R and sub-length NS over n = 2^NP points, by
// fft_threads(NP) threads: from 1024 points on they cover the pass's n / R
// butterflies exactly (16 / R each, no lane guards).  Point m comes from ld(m),
// output k of butterfly i, slot r goes to st(i, r, k, value) (i, r compile-time
// after unrolling, so a store may keep it in registers); IN_LDS / OUT_LDS say whether those touch the
// LDS buffer, which decides the barriers: LDS reads finish before LDS writes
// (the passes work in place), and LDS writes before the next pass reads.
template <int R, int NP, int NS, bool IN_LDS, bool OUT_LDS, class Ld, class St>
__device__ __forceinline__ void stockham_pass(const float2* twf, int tid, const Ld& ld, const St& st)
{
    constexpr int n = 1 << NP, T = fft_threads(NP), nb = n / R, NB = (nb + T - 1) / T;
    constexpr int tstep = n / (NS * R);
    float2 v[NB][R];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int j = tid + i * T;
        if (nb % T == 0 || j < nb) {
            const int jm = j & (NS - 1);
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = ld(j + r * nb);
            if constexpr (NS > 1) {
                // w^r for w = exp(-2 pi i jm tstep / n): w, w^2, w^4, w^8 from the table
                // (rounded from double), the others as products of at most four of them
                float2 w[R];
                const int e = jm * tstep;                       // (R - 1) e < n
#pragma unroll
                for (int b = 1; b < R; b <<= 1) w[b] = twf[b * e];
#pragma unroll
                for (int r = 3; r < R; r++)
                    if (r & (r - 1)) w[r] = cmulf(w[r & (r - 1)], w[r & -r]);
#pragma unroll
                for (int r = 1; r < R; r++) v[i][r] = cmulf(v[i][r], w[r]);
            }
        }
    }
    if constexpr (IN_LDS && OUT_LDS) __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int j = tid + i * T;
        if (nb % T == 0 || j < nb) {
            const int jm = j & (NS - 1);
            const int d = (j - jm) * R + jm;
            dft_r<R>(v[i]);
#pragma unroll
            for (int r = 0; r < R; r++) st(i, r, d + r * NS, v[i][dft_slot<R>(r)]);
        }
    }
    if constexpr (OUT_LDS) __syncthreads();
}

// Forward DFT of n = 2^NP points: radix-16 passes, then one of radix 8, 4 or 2.
// The first pass reads its points through ld (IN_LDS: from buf), the last pass
// writes through st (OUT_LDS: into buf); the passes between use buf.
template <int NP, int NS, int REM, bool IN_LDS, bool OUT_LDS, class Ld, class St>
__device__ __forceinline__ void fft_r16(float2* buf, const float2* twf, int tid, const Ld& ld, const St& st)
{
    constexpr int LR = REM >= 4 ? 4 : REM, R = 1 << LR;
    constexpr bool first = NS == 1, last = REM == LR;
    const LdsLoad lds_ld{buf};
    const LdsStore lds_st{buf};
    if constexpr (first && last) stockham_pass<R, NP, NS, IN_LDS, OUT_LDS>(twf, tid, ld, st);
    else if constexpr (first) stockham_pass<R, NP, NS, IN_LDS, true>(twf, tid, ld, lds_st);
    else if constexpr (last) stockham_pass<R, NP, NS, true, OUT_LDS>(twf, tid, lds_ld, st);
    else stockham_pass<R, NP, NS, true, true>(twf, tid, lds_ld, lds_st);
    if constexpr (!last) fft_r16<NP, (NS << LR), REM - LR, IN_LDS, OUT_LDS>(buf, twf, tid, ld, st);
}

// NP: log2 of the transform length (compile time); launched with
// fft_threads(NP) threads.  The forward transform reads the prepared rows from
// global memory in its first pass and applies conj(Hs * X) in its last; the
// inverse transform's last pass writes the cropped columns (LP false: FP32 rows,
// J.store 0) or keeps them in registers for the row scales (LP true: 16-bit
// rows, J.store 1 or 2).
template <int IN_KIND, int NP, bool LP>
__global__ void __launch_bounds__(fft_threads(NP), 4) filter_pair_kernel(FilterJob J, FilterBatch FB)
{
    constexpr int n = 1 << NP, T = fft_threads(NP);
    extern __shared__ float2 buf[];
    const int ra = 2 * (int)blockIdx.x, bi = blockIdx.y;
    const void* sino = FB.sino[bi];
    const bool hasb = ra + 1 < J.nrows;
    const int rin = J.row_in ? J.row_in[ra] : ra;
    const int rout = J.row_out ? J.row_out[ra] : ra;
    const int rin_b = hasb ? (J.row_in ? J.row_in[ra + 1] : ra + 1) : 0;
    const int rout_b = hasb ? (J.row_out ? J.row_out[ra + 1] : ra + 1) : 0;
    const int tid = threadIdx.x;
    const float2* twf = (const float2*)J.tw;
    const void* row = (IN_KIND == 0)
        ? (const void*)((const int16_t*)sino + (long long)rin * J.in_stride)
        : (const void*)((const float*)sino + (long long)rin * J.in_stride);
    const void* row_b = (IN_KIND == 0)
        ? (const void*)((const int16_t*)sino + (long long)rin_b * J.in_stride)
        : (const void*)((const float*)sino + (long long)rin_b * J.in_stride);
    const float* filt = J.filt;
    auto prep = [&](int m) {
        return make_float2(prep_value<IN_KIND>(J, row, m), hasb ? prep_value<IN_KIND>(J, row_b, m) : 0.f);
    };
    auto filter = [&](int, int, int k, float2 x) {          // conj(Hs * X)
        const float h = filt[k];
        buf[lds_pos(k)] = make_float2(x.x * h, -(x.y * h));
    };
    fft_r16<NP, 1, NP, false, true>(buf, twf, tid, prep, filter);
    // y = conj(buf) / n; rows: real part -> a, imaginary part -> b
    const float a = 1.f / (float)n;
    const LdsLoad lds_ld{buf};
    if constexpr (!LP) {
        float* o = FB.out[bi] + (long long)rout * J.out_stride;
        float* ob = FB.out[bi] + (long long)rout_b * J.out_stride;
        const int crop = J.crop, N = J.N;
        auto out = [&](int, int, int k, float2 y) {
            const int c = k - crop;
            if (c >= 0 && c < N) {
                o[c] = y.x * a;
                if (hasb) ob[c] = -y.y * a;
            }
        };
        fft_r16<NP, 1, NP, true, false>(buf, twf, tid, lds_ld, out);
    } else {
        // 16-bit rows: the inverse's last pass keeps its outputs in registers; per row
        // a power-of-two scale so that |value| / scale <= 2^14 (the largest cropped
        // value over the workgroup), then the scaled values stored from registers
        constexpr int LRL = NP % 4 == 0 ? 4 : NP % 4, RL = 1 << LRL;     // the last pass
        constexpr int NSL = n / RL, NBL = (NSL + T - 1) / T;
        float2 keep[NBL][RL];
        auto hold = [&](int i, int r, int, float2 y) { keep[i][r] = y; };
        fft_r16<NP, 1, NP, true, false>(buf, twf, tid, lds_ld, hold);
        // output k of butterfly i, slot r of the last pass: j = tid + i T < NSL, k = j + r NSL
        const int crop = J.crop, N = J.N;
        float ma = 0.f, mb = 0.f;
#pragma unroll
        for (int i = 0; i < NBL; i++) {
            const int j = tid + i * T;
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int c = j + r * NSL - crop;
                if ((NSL % T == 0 || j < NSL) && c >= 0 && c < N) {
                    ma = fmaxf(ma, fabsf(keep[i][r].x * a));
                    mb = fmaxf(mb, fabsf(-keep[i][r].y * a));
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            ma = fmaxf(ma, __shfl_xor(ma, off));
            mb = fmaxf(mb, __shfl_xor(mb, off));
        }
        __shared__ float red[2 * 16 + 2];
        if ((tid & 63) == 0) { red[tid >> 6] = ma; red[16 + (tid >> 6)] = mb; }
        __syncthreads();
        if (tid == 0) {
            float m2a = 0.f, m2b = 0.f;
            for (int w = 0; w < T / 64; w++) { m2a = fmaxf(m2a, red[w]); m2b = fmaxf(m2b, red[16 + w]); }
            int ea = 0, eb = 0;
            if (m2a > 0.f) frexpf(m2a / 16384.f, &ea);
            if (m2b > 0.f) frexpf(m2b / 16384.f, &eb);
            red[32] = ldexpf(1.f, ea);
            red[33] = ldexpf(1.f, eb);
            FB.out_scale[bi][rout] = red[32];
            if (hasb) FB.out_scale[bi][rout_b] = red[33];
        }
        __syncthreads();
        const float inva = 1.f / red[32], invb = 1.f / red[33];      // powers of two: exact
#pragma unroll
        for (int i = 0; i < NBL; i++) {
            const int j = tid + i * T;
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int c = j + r * NSL - crop;
                if ((NSL % T == 0 || j < NSL) && c >= 0 && c < N) {
                    const float va = keep[i][r].x * a, vb = -keep[i][r].y * a;
                    if (J.store == 1) {
                        __half* o = (__half*)FB.out_lp[bi];
                        o[(long long)rout * J.out_stride + c] = __float2half_rn(va * inva);
                        if (hasb) o[(long long)rout_b * J.out_stride + c] = __float2half_rn(vb * invb);
                    } else {
                        __hip_bfloat16* o = (__hip_bfloat16*)FB.out_lp[bi];
                        o[(long long)rout * J.out_stride + c] = __float2bfloat16(va * inva);
                        if (hasb) o[(long long)rout_b * J.out_stride + c] = __float2bfloat16(vb * invb);
                    }
                }
            }
        }
    }
}

// Transforms longer than 8192 points need more than the default 64 KiB of LDS;
// raised once, outside any stream capture.
void prepare_filter(int n)
{
    // per device: the attribute applies to the current device's copy of the kernels
    static std::mutex mu;
    static std::map<int, int> done;
    const size_t lds = sizeof(float2) * (size_t)n;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    if (lds <= 65536 || n <= done[dev]) return;
    (void)hipFuncSetAttribute((const void*)&filter_rows_kernel<0, 14>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)&filter_rows_kernel<1, 14>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)&filter_pair_kernel<0, 14, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)&filter_pair_kernel<1, 14, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)&filter_pair_kernel<0, 14, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)&filter_pair_kernel<1, 14, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    done[dev] = n;
}

template <int NP>
static void launch_np(const FilterJob& job, const FilterBatch& fb, int count, size_t lds, hipStream_t stream)
{
    const dim3 T(fft_threads(NP));
    if (job.pair) {
        const dim3 g((job.nrows + 1) / 2, count);
        const bool lp = job.store != 0;
        if (job.in_kind == 0 && !lp)
            hipLaunchKernelGGL((filter_pair_kernel<0, NP, false>), g, T, lds, stream, job, fb);
        else if (job.in_kind == 0)
            hipLaunchKernelGGL((filter_pair_kernel<0, NP, true>), g, T, lds, stream, job, fb);
        else if (!lp)
            hipLaunchKernelGGL((filter_pair_kernel<1, NP, false>), g, T, lds, stream, job, fb);
        else
            hipLaunchKernelGGL((filter_pair_kernel<1, NP, true>), g, T, lds, stream, job, fb);
    } else {
        const dim3 g(job.nrows, count);
        if (job.in_kind == 0)
            hipLaunchKernelGGL((filter_rows_kernel<0, NP>), g, T, lds, stream, job, fb);
        else
            hipLaunchKernelGGL((filter_rows_kernel<1, NP>), g, T, lds, stream, job, fb);
    }
}

hipError_t launch_filter_batch(const FilterJob& job, const FilterBatch& fb, int count, hipStream_t stream)
{
    if (job.nrows <= 0) return hipSuccess;
    if (count <= 0 || count > kFilterBatch || job.np < 3 || job.np > 14 || job.n != (1 << job.np))
        return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)job.n;
    prepare_filter(job.n);
    switch (job.np) {
    case 3: launch_np<3>(job, fb, count, lds, stream); break;
    case 4: launch_np<4>(job, fb, count, lds, stream); break;
    case 5: launch_np<5>(job, fb, count, lds, stream); break;
    case 6: launch_np<6>(job, fb, count, lds, stream); break;
    case 7: launch_np<7>(job, fb, count, lds, stream); break;
    case 8: launch_np<8>(job, fb, count, lds, stream); break;
    case 9: launch_np<9>(job, fb, count, lds, stream); break;
    case 10: launch_np<10>(job, fb, count, lds, stream); break;
    case 11: launch_np<11>(job, fb, count, lds, stream); break;
    case 12: launch_np<12>(job, fb, count, lds, stream); break;
    case 13: launch_np<13>(job, fb, count, lds, stream); break;
    case 14: launch_np<14>(job, fb, count, lds, stream); break;
    default: break;
    }
    return hipGetLastError();
}

hipError_t launch_filter(const FilterJob& job, hipStream_t stream)
{
    FilterBatch fb{};
    fb.sino[0] = job.sino;
    fb.out[0] = job.out;
    fb.out_lp[0] = job.out_lp;
    fb.out_scale[0] = job.out_scale;
    return launch_filter_batch(job, fb, 1, stream);
}

// ------------------------------------------------ compatibility kernels

// cudaReconst.cu:230-240 (formula 1) / general.cpp:2131-2136 (formula 0);
// `scale` is BACKPROJ_SCALE (100) for the reference LUT.
__global__ void px2igp_kernel(const float* px, int* igp, int N, int formula, float scale)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= N) return;
    const int iy = blockIdx.y;
    float p0, p1p0;
    if (formula == 0) {
        p0 = px[ix] * scale;
        p1p0 = (ix == N - 1) ? 0.0f : (px[ix + 1] - px[ix]) / 16.f * scale;
    } else {
        p0 = px[ix] * scale;
        p1p0 = (ix == N - 1) ? 0.0f : (px[ix + 1] * scale - p0) / 16.f;
    }
    igp[ix * 16 + iy] = (int)(p0 + p1p0 * (float)iy);
}

void launch_px2igp(const float* d_px, int* d_igp, int N, int formula, float scale, int block,
                   hipStream_t stream)
{
    if (block <= 0 || block > 1024) block = 256;
    hipLaunchKernelGGL(px2igp_kernel, dim3((N + block - 1) / block, 16), dim3(block), 0, stream,
                       d_px, d_igp, N, formula, scale);
}

// One projection's LUT added into the int32 image (cudaReconst.cu:31-39).
// Every pixel belongs to one thread and launches on a stream are ordered, so
// the reference's atomicAdd is a plain add here.  t uses the reference CPU
// operation order ix*fcos + (iy*fsin + foffset) without contraction.
__global__ void bp_lut_global_kernel(int* ifp, const int* igp, int N, float fcos, float fsin,
                                     float foffset, int rule)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= N) return;
    const int iy = blockIdx.y;
    const float fyoff = (float)iy * fsin + foffset;
    const float t = (float)ix * fcos + fyoff;
    const int ng = N * 16;
    int idx;
    if (rule == 0) {
        if (!(t >= 0.f && t < (float)ng)) return;
        idx = (int)t;
    } else {
        if (!(t > -2147483648.f && t < 2147483648.f)) return;
        idx = (int)t;
        if (idx < 0 || idx >= ng) return;
    }
    unsigned int* p = (unsigned int*)ifp + (long long)iy * N + ix;
    *p = *p + (unsigned int)igp[idx];
}

void launch_bp_lut_global(int* d_ifp, const int* d_igp, int N, float fcos, float fsin,
                          float foffset, int rule, int block, hipStream_t stream)
{
    if (block <= 0 || block > 1024) block = 256;
    hipLaunchKernelGGL(bp_lut_global_kernel, dim3((N + block - 1) / block, N), dim3(block), 0,
                       stream, d_ifp, d_igp, N, fcos, fsin, foffset, rule);
}

__global__ void zero_u32_kernel(unsigned int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

void launch_zero_u32(unsigned int* p, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
}

}  // namespace rv
