// rv_unet.hip -- CDNA4 kernels of the learned restoration stage.
//
// 3x3 convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward / data gradient: M = 16 pixels of an image row, N = 16 output
//   channels, K = 32 input channels of one tap.  A block owns a 16 x 16 pixel
//   patch and 16*NF output channels; per 32-channel chunk it stages the 18 x 18
//   halo (96-byte pixel pitch, conflict-free 16-byte reads) and the chunk's
//   weights (already in fragment order) in LDS, then each wave runs
//   9 taps x 4 rows x NF fragments of MFMA.
//   weight gradient: M = 16 output channels, N = 16 input channels, K = 32
//   pixels.  Pixels are the contraction, so dY and X are staged transposed
//   (pixel-contiguous) and X in three column-shifted copies, which keeps every
//   fragment read an aligned 16-byte LDS load for all nine taps.
// Operand layouts (checked on gfx950): A[i][k] in lane i + 16*(k/8), element
// k%8; B[k][j] in lane j + 16*(k/8), element k%8; D[i][j] in lane j + 16*(i/4),
// element i%4.
#include "rv_unet.h"

#include <algorithm>
#include <cmath>

namespace rv {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ uint16_t f2bf(float f) {
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
static __device__ __forceinline__ float bf2f(uint16_t u) {
    return __uint_as_float((uint32_t)u << 16);
}
static __device__ __forceinline__ uint4 ld16(const uint16_t* p) { return *(const uint4*)p; }
static __device__ __forceinline__ void st16(uint16_t* p, uint4 v) { *(uint4*)p = v; }

// ------------------------------------------------------------ conv forward
// Row pitches of LDS tiles read as 16-byte fragments (16 rows x 4 groups per
// wave): measured conflict-free on gfx950 when the pitch is 32 mod 64 bytes
// (80, 112, 144, 272, 336 bytes conflict 2x, 128 bytes 4x).
constexpr int HALO_PITCH = 48;          // bf16 per staged pixel: 32 channels + 16 pad
constexpr int CONV_ROWS = 16;           // patch rows (4 per wave)
constexpr int HALO_ROWS = CONV_ROWS + 2;
constexpr int HALO_PIX = HALO_ROWS * 18;
constexpr int HALO_ITEMS = HALO_PIX * 4;          // 16-byte loads per chunk
constexpr int HALO_PER = (HALO_ITEMS + 255) / 256;

// One block: a 16-wide, 4*MR-row pixel patch of one image and 16*NF output
// channels.  Wave w owns patch rows MR*w .. MR*w+MR-1 (MR M fragments) and all
// NF N fragments, so per tap it reads MR A and NF B fragments for MR*NF MFMAs.
// The next 32-channel chunk is fetched into registers while the current one is
// multiplied; the epilogue goes through LDS so every store is 16 bytes.
template <int NF, int MR>
__global__ __launch_bounds__(256, 2) void conv3x3_kernel(ConvArgs a) {
    constexpr int ROWS = 4 * MR;
    constexpr int HALO_PIX_K = (ROWS + 2) * 18;
    constexpr int HALO_ITEMS_K = HALO_PIX_K * 4;
    constexpr int HALO_PER_K = (HALO_ITEMS_K + 255) / 256;
    constexpr int NPIX = 16 * ROWS;
    constexpr int HSZ = HALO_PIX_K * HALO_PITCH;
    constexpr int WSZ = NF * 576 * 8;
    constexpr int TC = 16 * NF;
    constexpr int OPITCH = TC + 8;
    constexpr int OSZ = NPIX * OPITCH;
    constexpr int SMEM = HSZ + WSZ > OSZ ? HSZ + WSZ : OSZ;
    constexpr int W_ITEMS = NF * 576;
    constexpr int W_PER = (W_ITEMS + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint16_t smem[SMEM];
    uint16_t* in_s = smem;
    uint16_t* w_s = smem + HSZ;
    const int tiles_x = (a.W + 15) >> 4;
    const int x0 = (blockIdx.x % tiles_x) * 16, y0 = (blockIdx.x / tiles_x) * ROWS;
    const int bb = blockIdx.z, f0 = blockIdx.y * NF;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int nch = a.Cin >> 5;
    const uint16_t* in = a.in + (long long)bb * a.H * a.W * a.in_stride + a.in_off;
    const uint16_t* up = a.up ? a.up + (long long)bb * (a.H >> 1) * (a.W >> 1) * a.Cup : nullptr;
    floatx4 acc[MR][NF];
#pragma unroll
    for (int m = 0; m < MR; m++)
#pragma unroll
        for (int f = 0; f < NF; f++) acc[m][f] = floatx4{0.f, 0.f, 0.f, 0.f};

    uint4 hreg[HALO_PER_K], wreg[W_PER];
    auto fetch = [&](int cc) {
#pragma unroll
        for (int k = 0; k < HALO_PER_K; k++) {
            const int i = t + 256 * k;
            const int hp = i >> 2, g = i & 3;
            const int hy = hp / 18, hx = hp - hy * 18;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            hreg[k] = make_uint4(0, 0, 0, 0);
            if (i < HALO_ITEMS_K && y >= 0 && y < a.H && x >= 0 && x < a.W) {
                if (cc * 32 < a.Cup)
                    hreg[k] = ld16(up + ((long long)(y >> 1) * (a.W >> 1) + (x >> 1)) * a.Cup + cc * 32 +
                                   g * 8);
                else
                    hreg[k] = ld16(in + ((long long)y * a.W + x) * a.in_stride + cc * 32 + g * 8);
            }
        }
#pragma unroll
        for (int k = 0; k < W_PER; k++) {
            const int i = t + 256 * k;
            if (i < W_ITEMS) {
                const int f = i / 576, r = i - f * 576;
                wreg[k] = ld16(a.wfrag + ((long long)((f0 + f) * nch + cc) * 576 + r) * 8);
            }
        }
    };
    fetch(0);
    for (int cc = 0; cc < nch; cc++) {
#pragma unroll
        for (int k = 0; k < HALO_PER_K; k++) {
            const int i = t + 256 * k;
            if (i < HALO_ITEMS_K) st16(in_s + (i >> 2) * HALO_PITCH + (i & 3) * 8, hreg[k]);
        }
#pragma unroll
        for (int k = 0; k < W_PER; k++) {
            const int i = t + 256 * k;
            if (i < W_ITEMS) st16(w_s + i * 8, wreg[k]);
        }
        __syncthreads();
        if (cc + 1 < nch) fetch(cc + 1);
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int ky = tap / 3, kx = tap % 3;
            bf16x8 av[MR], bv[NF];
#pragma unroll
            for (int m = 0; m < MR; m++)
                av[m] = *(const bf16x8*)(in_s + ((MR * w + m + ky) * 18 + (l & 15) + kx) * HALO_PITCH +
                                         8 * (l >> 4));
#pragma unroll
            for (int f = 0; f < NF; f++) bv[f] = *(const bf16x8*)(w_s + ((f * 9 + tap) * 64 + l) * 8);
            // all fragment reads of the tap are in flight before its first MFMA
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int m = 0; m < MR; m++)
                    acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[f], acc[m][f], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: bias, ReLU, bf16 into an LDS tile [NPIX pixels][TC], then 16-byte stores
    uint16_t* o_s = smem;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const int co = f * 16 + (l & 15);
        const float bias = a.bias ? a.bias[(f0 * 16) + co] : 0.f;
#pragma unroll
        for (int m = 0; m < MR; m++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float v = acc[m][f][i] + bias;
                if (a.relu) v = fmaxf(v, 0.f);
                o_s[((MR * w + m) * 16 + 4 * (l >> 4) + i) * OPITCH + co] = f2bf(v);
            }
    }
    __syncthreads();
    if (a.pool_out) {
        // 8 x ROWS/2 pooled pixels of this patch, from the rounded outputs
        const int Ho = a.H >> 1, Wo = a.W >> 1;
        for (int it = t; it < 4 * ROWS * (TC / 8); it += 256) {
            const int q = it / (TC / 8), g = it - q * (TC / 8);
            const int py = q >> 3, pxx = q & 7;
            const int yo = (y0 >> 1) + py, xo = (x0 >> 1) + pxx;
            if (yo >= Ho || xo >= Wo) continue;
            float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int px = (2 * py + (k >> 1)) * 16 + 2 * pxx + (k & 1);
                const uint4 v = ld16(o_s + px * OPITCH + 8 * g);
                const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
                for (int e = 0; e < 8; e++) sum[e] += bf2f(v8[e]);
            }
            uint4 r;
            uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
            for (int e = 0; e < 8; e++) r8[e] = f2bf(0.25f * sum[e]);
            st16(a.pool_out + (((long long)bb * Ho + yo) * Wo + xo) * a.Cout + f0 * 16 + 8 * g, r);
        }
    }
    for (int it = t; it < NPIX * (TC / 8); it += 256) {
        const int px = it / (TC / 8), g = it - px * (TC / 8);
        const int y = y0 + (px >> 4), x = x0 + (px & 15);
        if (y >= a.H || x >= a.W) continue;
        const long long pix = ((long long)bb * a.H + y) * a.W + x;
        const int c = f0 * 16 + 8 * g;
        uint4 v = ld16(o_s + px * OPITCH + 8 * g);
        if (a.mask) {
            const uint4 mv = ld16(a.mask + pix * a.mask_stride + a.mask_off + c);
            uint16_t* v8 = (uint16_t*)&v;
            const uint16_t* m8 = (const uint16_t*)&mv;
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (!(bf2f(m8[e]) > 0.f)) v8[e] = 0;
        }
        st16(a.out + pix * a.out_stride + a.out_off + c, v);
    }
}

// 32-channel inputs (one chunk): a block walks PPB consecutive 16 x 16
// patches with the weights staged once.  The next patch's halo is fetched into
// registers while the current one is multiplied; the output tile has its own
// LDS so the weights survive the epilogue.
template <int NF, int PPB>
__global__ __launch_bounds__(256, 2) void conv3x3_c32_kernel(ConvArgs a) {
    constexpr int HSZ = HALO_PIX * HALO_PITCH;
    constexpr int WSZ = NF * 576 * 8;
    constexpr int TC = 16 * NF;
    constexpr int OPITCH = TC + 8;
    constexpr int W_ITEMS = NF * 576;
    constexpr int W_PER = (W_ITEMS + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint16_t smem[HSZ + WSZ + 256 * OPITCH];
    uint16_t* in_s = smem;
    uint16_t* w_s = smem + HSZ;
    uint16_t* o_s = w_s + WSZ;
    const int tiles_x = (a.W + 15) >> 4;
    const int ntiles = tiles_x * ((a.H + CONV_ROWS - 1) / CONV_ROWS);
    const int tile0 = blockIdx.x * PPB;
    const int npatch = ntiles - tile0 < PPB ? ntiles - tile0 : PPB;
    const int bb = blockIdx.z, f0 = blockIdx.y * NF;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const uint16_t* in = a.in + (long long)bb * a.H * a.W * a.in_stride + a.in_off;
    const uint16_t* up = a.up ? a.up + (long long)bb * (a.H >> 1) * (a.W >> 1) * a.Cup : nullptr;

#pragma unroll
    for (int k = 0; k < W_PER; k++) {
        const int i = t + 256 * k;
        if (i < W_ITEMS) {
            const int f = i / 576, r = i - f * 576;
            st16(w_s + i * 8, ld16(a.wfrag + ((long long)(f0 + f) * 576 + r) * 8));
        }
    }
    uint4 hreg[HALO_PER];
    auto fetch = [&](int p) {
        const int tile = tile0 + p;
        const int x0 = (tile % tiles_x) * 16, y0 = (tile / tiles_x) * CONV_ROWS;
#pragma unroll
        for (int k = 0; k < HALO_PER; k++) {
            const int i = t + 256 * k;
            const int hp = i >> 2, g = i & 3;
            const int hy = hp / 18, hx = hp - hy * 18;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            hreg[k] = make_uint4(0, 0, 0, 0);
            if (i < HALO_ITEMS && y >= 0 && y < a.H && x >= 0 && x < a.W) {
                if (a.Cup > 0)
                    hreg[k] = ld16(up + ((long long)(y >> 1) * (a.W >> 1) + (x >> 1)) * a.Cup + g * 8);
                else
                    hreg[k] = ld16(in + ((long long)y * a.W + x) * a.in_stride + g * 8);
            }
        }
    };
    fetch(0);
    for (int p = 0; p < npatch; p++) {
        const int tile = tile0 + p;
        const int x0 = (tile % tiles_x) * 16, y0 = (tile / tiles_x) * CONV_ROWS;
#pragma unroll
        for (int k = 0; k < HALO_PER; k++) {
            const int i = t + 256 * k;
            if (i < HALO_ITEMS) st16(in_s + (i >> 2) * HALO_PITCH + (i & 3) * 8, hreg[k]);
        }
        __syncthreads();
        if (p + 1 < npatch) fetch(p + 1);
        floatx4 acc[4][NF];
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int f = 0; f < NF; f++) acc[m][f] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int ky = tap / 3, kx = tap % 3;
            bf16x8 av[4], bv[NF];
#pragma unroll
            for (int m = 0; m < 4; m++)
                av[m] = *(const bf16x8*)(in_s + ((4 * w + m + ky) * 18 + (l & 15) + kx) * HALO_PITCH +
                                         8 * (l >> 4));
#pragma unroll
            for (int f = 0; f < NF; f++) bv[f] = *(const bf16x8*)(w_s + ((f * 9 + tap) * 64 + l) * 8);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int m = 0; m < 4; m++)
                    acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[f], acc[m][f], 0, 0, 0);
        }
        // bias, ReLU, bf16 into the output tile [256 pixels][TC]
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const int co = f * 16 + (l & 15);
            const float bias = a.bias ? a.bias[(f0 * 16) + co] : 0.f;
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float v = acc[m][f][i] + bias;
                    if (a.relu) v = fmaxf(v, 0.f);
                    o_s[((4 * w + m) * 16 + 4 * (l >> 4) + i) * OPITCH + co] = f2bf(v);
                }
        }
        __syncthreads();
        if (a.pool_out) {
            const int Ho = a.H >> 1, Wo = a.W >> 1;
            for (int it = t; it < 64 * (TC / 8); it += 256) {
                const int q = it / (TC / 8), g = it - q * (TC / 8);
                const int py = q >> 3, pxx = q & 7;
                const int yo = (y0 >> 1) + py, xo = (x0 >> 1) + pxx;
                if (yo >= Ho || xo >= Wo) continue;
                float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int px = (2 * py + (k >> 1)) * 16 + 2 * pxx + (k & 1);
                    const uint4 v = ld16(o_s + px * OPITCH + 8 * g);
                    const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
                    for (int e = 0; e < 8; e++) sum[e] += bf2f(v8[e]);
                }
                uint4 r;
                uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
                for (int e = 0; e < 8; e++) r8[e] = f2bf(0.25f * sum[e]);
                st16(a.pool_out + (((long long)bb * Ho + yo) * Wo + xo) * a.Cout + f0 * 16 + 8 * g, r);
            }
        }
        for (int it = t; it < 256 * (TC / 8); it += 256) {
            const int px = it / (TC / 8), g = it - px * (TC / 8);
            const int y = y0 + (px >> 4), x = x0 + (px & 15);
            if (y >= a.H || x >= a.W) continue;
            const long long pix = ((long long)bb * a.H + y) * a.W + x;
            const int c = f0 * 16 + 8 * g;
            uint4 v = ld16(o_s + px * OPITCH + 8 * g);
            if (a.mask) {
                const uint4 mv = ld16(a.mask + pix * a.mask_stride + a.mask_off + c);
                uint16_t* v8 = (uint16_t*)&v;
                const uint16_t* m8 = (const uint16_t*)&mv;
#pragma unroll
                for (int e = 0; e < 8; e++)
                    if (!(bf2f(m8[e]) > 0.f)) v8[e] = 0;
            }
            st16(a.out + pix * a.out_stride + a.out_off + c, v);
        }
        // the next iteration's halo store and output-tile writes are each
        // separated from this iteration's reads by one barrier
    }
}

// Encoder level 0 in one pass for 32 base channels: the edge convolution
// (1 -> 32, ReLU) of each patch's 18 x 18 halo is computed from a 20 x 20 tile
// of the input image in LDS, in edge_in_kernel's operation order (bias, the
// nine taps by fmaf, ReLU, bf16), straight into the halo tile of the 32 -> 32
// convolution; the MFMA loop and the epilogue (bias, ReLU, 2x2 pool, stores)
// are conv3x3_c32_kernel's.  The 32-channel edge output never goes to HBM.
// Halo item i = t + 256 k covers pixel i / 4 and channel group i % 4 = t % 4,
// so a thread keeps one channel group's 72 edge weights in registers.
template <int PPB>
__global__ __launch_bounds__(256, 2) void edge_c32_kernel(ConvArgs a, const float* x, const float* ew,
                                                          const float* eb) {
    constexpr int NF = 2;
    constexpr int HSZ = HALO_PIX * HALO_PITCH;
    constexpr int WSZ = NF * 576 * 8;
    constexpr int TC = 16 * NF;
    constexpr int OPITCH = TC + 8;
    constexpr int W_ITEMS = NF * 576;
    constexpr int W_PER = (W_ITEMS + 255) / 256;
    constexpr int XW = 20, XPIX = (CONV_ROWS + 4) * XW;     // input tile: halo 2 around the patch
    constexpr int X_PER = (XPIX + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint16_t smem[HSZ + WSZ + 256 * OPITCH];
    __shared__ float xs[XPIX];
    uint16_t* in_s = smem;
    uint16_t* w_s = smem + HSZ;
    uint16_t* o_s = w_s + WSZ;
    const int tiles_x = (a.W + 15) >> 4;
    const int ntiles = tiles_x * ((a.H + CONV_ROWS - 1) / CONV_ROWS);
    const int tile0 = blockIdx.x * PPB;
    const int npatch = ntiles - tile0 < PPB ? ntiles - tile0 : PPB;
    const int bb = blockIdx.z;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const float* xb = x + (long long)bb * a.H * a.W;

#pragma unroll
    for (int k = 0; k < W_PER; k++) {
        const int i = t + 256 * k;
        if (i < W_ITEMS) st16(w_s + i * 8, ld16(a.wfrag + (long long)i * 8));
    }
    const int g = t & 3;
    float ewr[8][9], ebr[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        ebr[e] = eb[8 * g + e];
#pragma unroll
        for (int tap = 0; tap < 9; tap++) ewr[e][tap] = ew[(8 * g + e) * 9 + tap];
    }
    float xreg[X_PER];
    auto fetch = [&](int p) {
        const int tile = tile0 + p;
        const int x0 = (tile % tiles_x) * 16, y0 = (tile / tiles_x) * CONV_ROWS;
#pragma unroll
        for (int k = 0; k < X_PER; k++) {
            const int i = t + 256 * k;
            const int ty = i / XW, tx = i - ty * XW;
            const int y = y0 - 2 + ty, xx = x0 - 2 + tx;
            xreg[k] = (i < XPIX && y >= 0 && y < a.H && xx >= 0 && xx < a.W) ? xb[(long long)y * a.W + xx] : 0.f;
        }
    };
    fetch(0);
    for (int p = 0; p < npatch; p++) {
        const int tile = tile0 + p;
        const int x0 = (tile % tiles_x) * 16, y0 = (tile / tiles_x) * CONV_ROWS;
#pragma unroll
        for (int k = 0; k < X_PER; k++) {
            const int i = t + 256 * k;
            if (i < XPIX) xs[i] = xreg[k];
        }
        __syncthreads();
        if (p + 1 < npatch) fetch(p + 1);
        // edge convolution of the halo: zero outside the image (the 32 -> 32
        // convolution's padding), the tile's zeros are the edge layer's padding
#pragma unroll
        for (int k = 0; k < HALO_PER; k++) {
            const int i = t + 256 * k;
            if (i < HALO_ITEMS) {
                const int hp = i >> 2;
                const int hy = hp / 18, hx = hp - hy * 18;
                const int y = y0 - 1 + hy, xx = x0 - 1 + hx;
                uint4 r = make_uint4(0, 0, 0, 0);
                if (y >= 0 && y < a.H && xx >= 0 && xx < a.W) {
                    float v[9];
#pragma unroll
                    for (int tap = 0; tap < 9; tap++) v[tap] = xs[(hy + tap / 3) * XW + hx + tap % 3];
                    uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        float s = ebr[e];
#pragma unroll
                        for (int tap = 0; tap < 9; tap++) s = fmaf(ewr[e][tap], v[tap], s);
                        r8[e] = f2bf(fmaxf(s, 0.f));
                    }
                }
                st16(in_s + hp * HALO_PITCH + g * 8, r);
            }
        }
        __syncthreads();
        floatx4 acc[4][NF];
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int f = 0; f < NF; f++) acc[m][f] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int ky = tap / 3, kx = tap % 3;
            bf16x8 av[4], bv[NF];
#pragma unroll
            for (int m = 0; m < 4; m++)
                av[m] = *(const bf16x8*)(in_s + ((4 * w + m + ky) * 18 + (l & 15) + kx) * HALO_PITCH +
                                         8 * (l >> 4));
#pragma unroll
            for (int f = 0; f < NF; f++) bv[f] = *(const bf16x8*)(w_s + ((f * 9 + tap) * 64 + l) * 8);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int m = 0; m < 4; m++)
                    acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[f], acc[m][f], 0, 0, 0);
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const int co = f * 16 + (l & 15);
            const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float v = acc[m][f][i] + bias;
                    if (a.relu) v = fmaxf(v, 0.f);
                    o_s[((4 * w + m) * 16 + 4 * (l >> 4) + i) * OPITCH + co] = f2bf(v);
                }
        }
        __syncthreads();
        if (a.pool_out) {
            const int Ho = a.H >> 1, Wo = a.W >> 1;
            for (int it = t; it < 64 * (TC / 8); it += 256) {
                const int q = it / (TC / 8), gg = it - q * (TC / 8);
                const int py = q >> 3, pxx = q & 7;
                const int yo = (y0 >> 1) + py, xo = (x0 >> 1) + pxx;
                if (yo >= Ho || xo >= Wo) continue;
                float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int px = (2 * py + (k >> 1)) * 16 + 2 * pxx + (k & 1);
                    const uint4 v = ld16(o_s + px * OPITCH + 8 * gg);
                    const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
                    for (int e = 0; e < 8; e++) sum[e] += bf2f(v8[e]);
                }
                uint4 r;
                uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
                for (int e = 0; e < 8; e++) r8[e] = f2bf(0.25f * sum[e]);
                st16(a.pool_out + (((long long)bb * Ho + yo) * Wo + xo) * a.Cout + 8 * gg, r);
            }
        }
        for (int it = t; it < 256 * (TC / 8); it += 256) {
            const int px = it / (TC / 8), gg = it - px * (TC / 8);
            const int y = y0 + (px >> 4), xx = x0 + (px & 15);
            if (y >= a.H || xx >= a.W) continue;
            const long long pix = ((long long)bb * a.H + y) * a.W + xx;
            st16(a.out + pix * a.out_stride + a.out_off + 8 * gg, ld16(o_s + px * OPITCH + 8 * gg));
        }
        // the next patch's tile and halo stores and output-tile writes are each
        // separated from this patch's reads by one barrier
    }
}

hipError_t launch_edge_conv3x3_c32(const ConvArgs& a, const float* x, const float* ew, const float* eb,
                                   hipStream_t s) {
    if (a.Cin != 32 || a.Cout != 32 || a.out_stride % 8 || a.out_off % 8 || a.B <= 0 || a.H <= 0 ||
        a.W <= 0 || a.out_off + a.Cout > a.out_stride || a.up || a.Cup || a.mask ||
        (a.pool_out && (a.H % 2 || a.W % 2)))
        return hipErrorInvalidValue;
    const int tiles = ((a.W + 15) / 16) * ((a.H + CONV_ROWS - 1) / CONV_ROWS);
    hipLaunchKernelGGL((edge_c32_kernel<4>), dim3((tiles + 3) / 4, 1, a.B), dim3(256), 0, s, a, x, ew, eb);
    return hipGetLastError();
}

hipError_t launch_conv3x3(const ConvArgs& a, hipStream_t s) {
    if (a.Cin % 32 || a.Cout % 32 || a.in_stride % 8 || a.in_off % 8 || a.out_stride % 8 ||
        a.out_off % 8 || a.B <= 0 || a.H <= 0 || a.W <= 0 || a.in_off + a.Cin > a.in_stride ||
        a.out_off + a.Cout > a.out_stride ||
        (a.mask && (a.mask_stride % 8 || a.mask_off % 8 || a.mask_off + a.Cout > a.mask_stride)) ||
        (a.up && (a.Cup % 32 || a.Cup > a.Cin || a.H % 2 || a.W % 2)) || (!a.up && a.Cup) ||
        (a.pool_out && (a.H % 2 || a.W % 2)))
        return hipErrorInvalidValue;
    const int tiles = ((a.W + 15) / 16) * ((a.H + CONV_ROWS - 1) / CONV_ROWS);
    const int tiles24 = ((a.W + 15) / 16) * ((a.H + 23) / 24);
    // 32 -> 32 layers (one chunk, 64 output channels of work per 16 KB of halo)
    // walk four patches per block with the weights staged once; other 32-output
    // layers use 24-row patches (6 M fragments per wave against 2 N fragments).
    if (a.Cout % 64 == 0)
        hipLaunchKernelGGL((conv3x3_kernel<4, 4>), dim3(tiles, a.Cout / 64, a.B), dim3(256), 0, s, a);
    else if (a.Cin == 32)
        hipLaunchKernelGGL((conv3x3_c32_kernel<2, 4>), dim3((tiles + 3) / 4, a.Cout / 32, a.B), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((conv3x3_kernel<2, 6>), dim3(tiles24, a.Cout / 32, a.B), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ weight gradient
constexpr int DY_PITCH = 144;   // bf16 per staged output channel: 128 pixels + 16 pad
constexpr int XT_PITCH = 176;   // bf16 per staged (shift, channel): 10 x 16 pixels + 16 pad

template <int NFW>
__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(WgradArgs a) {
    constexpr int TCO = 16 * NFW;
    constexpr int CF = NFW == 4 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) uint16_t dy_s[TCO * DY_PITCH];
    __shared__ __attribute__((aligned(16))) uint16_t x_s[3 * 32 * XT_PITCH];
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int co0 = blockIdx.x * TCO, cc = blockIdx.y, split = blockIdx.z;
    const int cof = w % NFW, cib = (w / NFW) * CF;
    const int tiles_x = (a.W + 15) >> 4, tiles_y = (a.H + 7) >> 3;
    const int per_img = tiles_x * tiles_y, npatch = a.B * per_img;
    floatx4 acc[9][CF];
#pragma unroll
    for (int k = 0; k < 9; k++)
#pragma unroll
        for (int c = 0; c < CF; c++) acc[k][c] = floatx4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int patch = split; patch < npatch; patch += a.nsplit) {
        const int bb = patch / per_img, r = patch - bb * per_img;
        const int y0 = (r / tiles_x) * 8, x0 = (r % tiles_x) * 16;
        for (int i = t; i < 128 * (TCO / 8); i += 256) {
            const int px = i / (TCO / 8), g = i - px * (TCO / 8);
            const int y = y0 + (px >> 4), x = x0 + (px & 15);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (y < a.H && x < a.W)
                v = ld16(a.dy + (((long long)bb * a.H + y) * a.W + x) * a.dy_stride + a.dy_off + co0 +
                         8 * g);
            const uint16_t* e8 = (const uint16_t*)&v;
#pragma unroll
            for (int e = 0; e < 8; e++) dy_s[(8 * g + e) * DY_PITCH + px] = e8[e];
        }
        for (int i = t; i < 720; i += 256) {
            const int hp = i >> 2, g = i & 3;
            const int hy = hp / 18, hx = hp - hy * 18;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (y >= 0 && y < a.H && x >= 0 && x < a.W)
                v = ld16(a.x + (((long long)bb * a.H + y) * a.W + x) * a.x_stride + a.x_off + cc * 32 +
                         8 * g);
            const uint16_t* e8 = (const uint16_t*)&v;
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int c = hx - kx;
                if (c < 0 || c >= 16) continue;
#pragma unroll
                for (int e = 0; e < 8; e++) x_s[(kx * 32 + 8 * g + e) * XT_PITCH + hy * 16 + c] = e8[e];
            }
        }
        __syncthreads();
        if (cc == 0 && t < TCO) {
            for (int px = 0; px < 128; px += 8) {
                const uint4 v = ld16(dy_s + t * DY_PITCH + px);
                const uint16_t* e8 = (const uint16_t*)&v;
#pragma unroll
                for (int e = 0; e < 8; e++) bsum += bf2f(e8[e]);
            }
        }
        const int q = l >> 4;
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            const bf16x8 av = *(const bf16x8*)(dy_s + (cof * 16 + (l & 15)) * DY_PITCH + kb * 32 + 8 * q);
            const int rr = 2 * kb + (q >> 1), c0 = 8 * (q & 1);
#pragma unroll
            for (int tap = 0; tap < 9; tap++) {
                const int ky = tap / 3, kx = tap % 3;
#pragma unroll
                for (int cf = 0; cf < CF; cf++) {
                    const int ci = (cib + cf) * 16 + (l & 15);
                    const bf16x8 bv =
                        *(const bf16x8*)(x_s + (kx * 32 + ci) * XT_PITCH + (rr + ky) * 16 + c0);
                    acc[tap][cf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[tap][cf], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float* P = a.part + (long long)split * a.Cout * 9 * a.Cin;
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int cf = 0; cf < CF; cf++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int co = co0 + cof * 16 + 4 * (l >> 4) + e;
                const int ci = cc * 32 + (cib + cf) * 16 + (l & 15);
                P[((long long)co * 9 + tap) * a.Cin + ci] = acc[tap][cf][e];
            }
    if (cc == 0 && t < TCO) a.partb[(long long)split * a.Cout + co0 + t] = bsum;
}

int wgrad_splits(int Cin, int Cout, int B, int H, int W) {
    const int nfw = Cout % 64 == 0 ? 4 : 2;
    const int tiles = (Cout / (16 * nfw)) * (Cin / 32);
    const long long npatch = (long long)B * ((H + 7) / 8) * ((W + 15) / 16);
    long long ns = (1024 + tiles - 1) / tiles;
    return (int)std::max(1LL, std::min(ns, npatch));
}

hipError_t launch_conv3x3_wgrad(const WgradArgs& a, hipStream_t s) {
    if (a.Cin % 32 || a.Cout % 32 || a.dy_stride % 8 || a.dy_off % 8 || a.x_stride % 8 ||
        a.x_off % 8 || a.nsplit <= 0 || a.dy_off + a.Cout > a.dy_stride || a.x_off + a.Cin > a.x_stride ||
        a.B <= 0 || a.H <= 0 || a.W <= 0)
        return hipErrorInvalidValue;
    if (a.Cout % 64 == 0)
        hipLaunchKernelGGL(conv3x3_wgrad_kernel<4>, dim3(a.Cout / 64, a.Cin / 32, a.nsplit), dim3(256),
                           0, s, a);
    else
        hipLaunchKernelGGL(conv3x3_wgrad_kernel<2>, dim3(a.Cout / 32, a.Cin / 32, a.nsplit), dim3(256),
                           0, s, a);
    return hipGetLastError();
}

__global__ void sum_splits_kernel(const float* part, int nsplit, int n, float* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += part[(long long)k * n + i];
    out[i] = s;
}

hipError_t launch_sum_splits(const float* part, int nsplit, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_splits_kernel, dim3((n + 255) / 256), dim3(256), 0, s, part, nsplit, n, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ weight packing
__global__ void pack_kernel(const float* w, int Cin, int Cout, int transpose, uint16_t* out,
                            long long total) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int CinP = transpose ? Cout : Cin;   // contraction channels of the launch
    const int nch = CinP / 32;
    const int e = idx & 7, l = (idx >> 3) & 63;
    const long long r = idx >> 9;              // (f*nch + cc)*9 + tap
    const int tap = (int)(r % 9);
    const long long fc = r / 9;
    const int cc = (int)(fc % nch), f = (int)(fc / nch);
    const int o = f * 16 + (l & 15);           // output channel of the launch
    const int c = cc * 32 + 8 * (l >> 4) + e;  // contraction channel of the launch
    const float v = transpose ? w[((long long)c * 9 + (8 - tap)) * Cin + o]
                              : w[((long long)o * 9 + tap) * Cin + c];
    out[idx] = f2bf(v);
}

hipError_t launch_pack_conv_weights(const float* w, int Cin, int Cout, int transpose,
                                    uint16_t* wfrag, hipStream_t s) {
    if (Cin % 32 || Cout % 32) return hipErrorInvalidValue;
    const long long total = (long long)Cin * Cout * 9;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, w, Cin,
                       Cout, transpose, wfrag, total);
    return hipGetLastError();
}

// ------------------------------------------------------------ edge layers
static __device__ __forceinline__ void pix_coords(long long p, int H, int W, int& b, int& y, int& x) {
    x = (int)(p % W);
    const long long r = p / W;
    y = (int)(r % H);
    b = (int)(r / H);
}

// One block = 256 consecutive pixels; outputs are staged in LDS so the stores
// are contiguous 16-byte chunks.
template <int C>
__global__ __launch_bounds__(256) void edge_in_kernel(const float* x, const float* w, const float* b,
                                                      uint16_t* out, int os, int oo, int B, int H, int W) {
    constexpr int PITCH = C + 8;
    __shared__ __attribute__((aligned(16))) uint16_t o_s[256 * PITCH];
    const long long n = (long long)B * H * W;
    const long long p0 = (long long)blockIdx.x * 256;
    const long long p = p0 + threadIdx.x;
    if (p < n) {
        int bb, y, xx;
        pix_coords(p, H, W, bb, y, xx);
        float v[9];
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int yy = y + tap / 3 - 1, x2 = xx + tap % 3 - 1;
            v[tap] = (yy >= 0 && yy < H && x2 >= 0 && x2 < W) ? x[((long long)bb * H + yy) * W + x2] : 0.f;
        }
#pragma unroll
        for (int g = 0; g < C / 8; g++) {
            uint4 r;
            uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int c = 8 * g + e;
                float s = b[c];
#pragma unroll
                for (int tap = 0; tap < 9; tap++) s = fmaf(w[c * 9 + tap], v[tap], s);
                r8[e] = f2bf(fmaxf(s, 0.f));
            }
            st16(o_s + threadIdx.x * PITCH + 8 * g, r);
        }
    }
    __syncthreads();
    for (int it = threadIdx.x; it < 256 * (C / 8); it += 256) {
        const int px = it / (C / 8), g = it - px * (C / 8);
        if (p0 + px < n) st16(out + (p0 + px) * os + oo + 8 * g, ld16(o_s + px * PITCH + 8 * g));
    }
}

hipError_t launch_edge_in(const float* x, const float* w, const float* b, int C, uint16_t* out,
                          int out_stride, int out_off, int B, int H, int W, hipStream_t s) {
    const long long n = (long long)B * H * W;
    const dim3 g((unsigned)((n + 255) / 256));
    if (out_stride % 8 || out_off % 8) return hipErrorInvalidValue;
    if (C == 32)
        hipLaunchKernelGGL(edge_in_kernel<32>, g, dim3(256), 0, s, x, w, b, out, out_stride, out_off, B, H, W);
    else if (C == 64)
        hipLaunchKernelGGL(edge_in_kernel<64>, g, dim3(256), 0, s, x, w, b, out, out_stride, out_off, B, H, W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int edge_blocks(int B, int H, int W) {
    const long long n = (long long)B * H * W;
    return (int)std::max(1LL, std::min(1024LL, (n + 4095) / 4096));
}

// Pixels of block k: [k*ppb, min(n, (k+1)*ppb)).  Thread t handles channel
// group t % (C/8) for every (256 / (C/8))-th pixel.  part[blk] = weights
// [c][tap] (9*C) then biases [c] (C), the layout of the parameters.
template <int C>
__global__ void edge_in_wgrad_kernel(const uint16_t* dy, const float* x, float* part, int B, int H,
                                     int W) {
    constexpr int G = C / 8, LANES = 256 / G;
    __shared__ float red[C * 10];
    const long long n = (long long)B * H * W;
    const long long ppb = (n + gridDim.x - 1) / gridDim.x;
    const long long p0 = blockIdx.x * ppb, p1 = p0 + ppb < n ? p0 + ppb : n;
    const int t = threadIdx.x, g = t % G;
    float acc[8][10];
#pragma unroll
    for (int e = 0; e < 8; e++)
#pragma unroll
        for (int k = 0; k < 10; k++) acc[e][k] = 0.f;
    for (int i = t; i < C * 10; i += 256) red[i] = 0.f;
    for (long long p = p0 + t / G; p < p1; p += LANES) {
        int bb, y, xx;
        pix_coords(p, H, W, bb, y, xx);
        const uint4 d = ld16(dy + p * C + 8 * g);
        const uint16_t* d8 = (const uint16_t*)&d;
        float v[9];
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int yy = y + tap / 3 - 1, x2 = xx + tap % 3 - 1;
            v[tap] = (yy >= 0 && yy < H && x2 >= 0 && x2 < W) ? x[((long long)bb * H + yy) * W + x2] : 0.f;
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float de = bf2f(d8[e]);
#pragma unroll
            for (int tap = 0; tap < 9; tap++) acc[e][tap] = fmaf(de, v[tap], acc[e][tap]);
            acc[e][9] += de;
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; e++)
#pragma unroll
        for (int k = 0; k < 10; k++)
            atomicAdd(&red[k < 9 ? (8 * g + e) * 9 + k : 9 * C + 8 * g + e], acc[e][k]);
    __syncthreads();
    for (int i = t; i < C * 10; i += 256) part[(long long)blockIdx.x * C * 10 + i] = red[i];
}

hipError_t launch_edge_in_wgrad(const uint16_t* dy, int C, const float* x, float* part, int nblk,
                                int B, int H, int W, hipStream_t s) {
    if (C == 32)
        hipLaunchKernelGGL(edge_in_wgrad_kernel<32>, dim3(nblk), dim3(256), 0, s, dy, x, part, B, H, W);
    else if (C == 64)
        hipLaunchKernelGGL(edge_in_wgrad_kernel<64>, dim3(nblk), dim3(256), 0, s, dy, x, part, B, H, W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// One block = a 16 x 16 pixel patch; the 18 x 18 halo of X is staged in LDS
// with coalesced 16-byte loads, then each thread reduces 9 taps x C channels.
template <int C>
__global__ __launch_bounds__(256) void edge_out_kernel(const uint16_t* x, const float* w, const float* b,
                                                       const float* xres, float* y, int B, int H, int W) {
    constexpr int PITCH = C + 8;
    constexpr int G = C / 8;
    __shared__ __attribute__((aligned(16))) uint16_t h_s[324 * PITCH];
    const int tiles_x = (W + 15) >> 4;
    const int x0 = (blockIdx.x % tiles_x) * 16, y0 = (blockIdx.x / tiles_x) * 16;
    const int bb = blockIdx.y;
    for (int i = threadIdx.x; i < 324 * G; i += 256) {
        const int hp = i / G, g = i - hp * G;
        const int hy = hp / 18, hx = hp - hy * 18;
        const int yy = y0 - 1 + hy, xx = x0 - 1 + hx;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = ld16(x + (((long long)bb * H + yy) * W + xx) * C + 8 * g);
        st16(h_s + hp * PITCH + 8 * g, v);
    }
    __syncthreads();
    const int py = threadIdx.x >> 4, pxl = threadIdx.x & 15;
    const int yy = y0 + py, xx = x0 + pxl;
    if (yy >= H || xx >= W) return;
    float s = b[0];
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const uint16_t* q = h_s + ((py + tap / 3) * 18 + pxl + tap % 3) * PITCH;
#pragma unroll
        for (int g = 0; g < G; g++) {
            const uint4 v = ld16(q + 8 * g);
            const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
            for (int e = 0; e < 8; e++) s = fmaf(w[tap * C + 8 * g + e], bf2f(v8[e]), s);
        }
    }
    const long long p = ((long long)bb * H + yy) * W + xx;
    y[p] = xres[p] + s;
}

hipError_t launch_edge_out(const uint16_t* x, int C, const float* w, const float* b,
                           const float* xres, float* y, int B, int H, int W, hipStream_t s) {
    const dim3 g((unsigned)(((W + 15) / 16) * ((H + 15) / 16)), (unsigned)B);
    if (C == 32)
        hipLaunchKernelGGL(edge_out_kernel<32>, g, dim3(256), 0, s, x, w, b, xres, y, B, H, W);
    else if (C == 64)
        hipLaunchKernelGGL(edge_out_kernel<64>, g, dim3(256), 0, s, x, w, b, xres, y, B, H, W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int C>
__global__ void edge_out_bwd_kernel(const float* dy, const float* w, const uint16_t* x, uint16_t* dx,
                                    int B, int H, int W) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    int bb, yy, xx;
    pix_coords(p, H, W, bb, yy, xx);
    float d[9];
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int y2 = yy - (tap / 3 - 1), x2 = xx - (tap % 3 - 1);
        d[tap] = (y2 >= 0 && y2 < H && x2 >= 0 && x2 < W) ? dy[((long long)bb * H + y2) * W + x2] : 0.f;
    }
#pragma unroll
    for (int g = 0; g < C / 8; g++) {
        const uint4 xv = ld16(x + p * C + 8 * g);
        const uint16_t* x8 = (const uint16_t*)&xv;
        uint4 r;
        uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int c = 8 * g + e;
            float s = 0.f;
#pragma unroll
            for (int tap = 0; tap < 9; tap++) s = fmaf(w[tap * C + c], d[tap], s);
            r8[e] = f2bf(bf2f(x8[e]) > 0.f ? s : 0.f);
        }
        st16(dx + p * C + 8 * g, r);
    }
}

hipError_t launch_edge_out_bwd(const float* dy, const float* w, const uint16_t* x, int C,
                               uint16_t* dx, int B, int H, int W, hipStream_t s) {
    const long long n = (long long)B * H * W;
    const dim3 g((unsigned)((n + 255) / 256));
    if (C == 32)
        hipLaunchKernelGGL(edge_out_bwd_kernel<32>, g, dim3(256), 0, s, dy, w, x, dx, B, H, W);
    else if (C == 64)
        hipLaunchKernelGGL(edge_out_bwd_kernel<64>, g, dim3(256), 0, s, dy, w, x, dx, B, H, W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int C>
__global__ void edge_out_wgrad_kernel(const float* dy, const uint16_t* x, float* part, int B, int H,
                                      int W) {
    constexpr int G = C / 8, LANES = 256 / G;
    __shared__ float red[9 * C + 1];
    const long long n = (long long)B * H * W;
    const long long ppb = (n + gridDim.x - 1) / gridDim.x;
    const long long p0 = blockIdx.x * ppb, p1 = p0 + ppb < n ? p0 + ppb : n;
    const int t = threadIdx.x, g = t % G;
    float acc[9][8];
    float bs = 0.f;
#pragma unroll
    for (int k = 0; k < 9; k++)
#pragma unroll
        for (int e = 0; e < 8; e++) acc[k][e] = 0.f;
    for (int i = t; i < 9 * C + 1; i += 256) red[i] = 0.f;
    for (long long p = p0 + t / G; p < p1; p += LANES) {
        int bb, yy, xx;
        pix_coords(p, H, W, bb, yy, xx);
        const float d = dy[p];
        if (g == 0) bs += d;
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int y2 = yy + tap / 3 - 1, x2 = xx + tap % 3 - 1;
            if (y2 < 0 || y2 >= H || x2 < 0 || x2 >= W) continue;
            const uint4 v = ld16(x + (((long long)bb * H + y2) * W + x2) * C + 8 * g);
            const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
            for (int e = 0; e < 8; e++) acc[tap][e] = fmaf(d, bf2f(v8[e]), acc[tap][e]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 9; k++)
#pragma unroll
        for (int e = 0; e < 8; e++) atomicAdd(&red[k * C + 8 * g + e], acc[k][e]);
    if (g == 0) atomicAdd(&red[9 * C], bs);
    __syncthreads();
    for (int i = t; i < 9 * C + 1; i += 256) part[(long long)blockIdx.x * (9 * C + 1) + i] = red[i];
}

hipError_t launch_edge_out_wgrad(const float* dy, const uint16_t* x, int C, float* part, int nblk,
                                 int B, int H, int W, hipStream_t s) {
    if (C == 32)
        hipLaunchKernelGGL(edge_out_wgrad_kernel<32>, dim3(nblk), dim3(256), 0, s, dy, x, part, B, H, W);
    else if (C == 64)
        hipLaunchKernelGGL(edge_out_wgrad_kernel<64>, dim3(nblk), dim3(256), 0, s, dy, x, part, B, H, W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ------------------------------------------------------------ pooling / upsampling
__global__ void avgpool2_kernel(const uint16_t* in, int is, int io, int C, uint16_t* out, int B,
                                int H, int W) {
    const int G = C / 8, Ho = H / 2, Wo = W / 2;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)B * Ho * Wo * G) return;
    const int g = (int)(i % G);
    const long long q = i / G;
    int bb, y, x;
    pix_coords(q, Ho, Wo, bb, y, x);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long p = ((long long)bb * H + 2 * y + (k >> 1)) * W + 2 * x + (k & 1);
        const uint4 v = ld16(in + p * is + io + 8 * g);
        const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += bf2f(v8[e]);
    }
    uint4 r;
    uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
    for (int e = 0; e < 8; e++) r8[e] = f2bf(0.25f * s[e]);
    st16(out + q * C + 8 * g, r);
}

hipError_t launch_avgpool2(const uint16_t* in, int in_stride, int in_off, int C, uint16_t* out,
                           int B, int H, int W, hipStream_t s) {
    if (C % 8 || in_stride % 8 || in_off % 8 || H % 2 || W % 2) return hipErrorInvalidValue;
    const long long n = (long long)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(avgpool2_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in,
                       in_stride, in_off, C, out, B, H, W);
    return hipGetLastError();
}

__global__ void upsample2_kernel(const uint16_t* in, int C, uint16_t* out, int os, int oo, int B,
                                 int H, int W) {
    const int G = C / 8;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)B * H * W * G) return;
    const int g = (int)(i % G);
    const long long p = i / G;
    int bb, y, x;
    pix_coords(p, H, W, bb, y, x);
    const long long q = ((long long)bb * (H / 2) + y / 2) * (W / 2) + x / 2;
    st16(out + p * os + oo + 8 * g, ld16(in + q * C + 8 * g));
}

hipError_t launch_upsample2(const uint16_t* in, int C, uint16_t* out, int out_stride, int out_off,
                            int B, int H, int W, hipStream_t s) {
    if (C % 8 || out_stride % 8 || out_off % 8 || H % 2 || W % 2) return hipErrorInvalidValue;
    const long long n = (long long)B * H * W * (C / 8);
    hipLaunchKernelGGL(upsample2_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, C,
                       out, out_stride, out_off, B, H, W);
    return hipGetLastError();
}

// H, W: full-resolution size of g; out is H/2 x W/2 with C channels.
__global__ void upsample2_bwd_kernel(const uint16_t* g, int gs, int go, int C, uint16_t* out, int B,
                                     int H, int W) {
    const int G = C / 8, Ho = H / 2, Wo = W / 2;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)B * Ho * Wo * G) return;
    const int gg = (int)(i % G);
    const long long q = i / G;
    int bb, y, x;
    pix_coords(q, Ho, Wo, bb, y, x);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long p = ((long long)bb * H + 2 * y + (k >> 1)) * W + 2 * x + (k & 1);
        const uint4 v = ld16(g + p * gs + go + 8 * gg);
        const uint16_t* v8 = (const uint16_t*)&v;
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += bf2f(v8[e]);
    }
    uint4 r;
    uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
    for (int e = 0; e < 8; e++) r8[e] = f2bf(s[e]);
    st16(out + q * C + 8 * gg, r);
}

hipError_t launch_upsample2_bwd(const uint16_t* g, int g_stride, int g_off, int C, uint16_t* out,
                                int B, int H, int W, hipStream_t s) {
    if (C % 8 || g_stride % 8 || g_off % 8 || H % 2 || W % 2) return hipErrorInvalidValue;
    const long long n = (long long)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(upsample2_bwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g,
                       g_stride, g_off, C, out, B, H, W);
    return hipGetLastError();
}

__global__ void pool_bwd_add_kernel(const uint16_t* skip, int ss, int so, const uint16_t* dp,
                                    const uint16_t* e, int es, int eo, int C, uint16_t* out, int B,
                                    int H, int W) {
    const int G = C / 8;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)B * H * W * G) return;
    const int g = (int)(i % G);
    const long long p = i / G;
    int bb, y, x;
    pix_coords(p, H, W, bb, y, x);
    const long long q = ((long long)bb * (H / 2) + y / 2) * (W / 2) + x / 2;
    const uint4 sv = ld16(skip + p * ss + so + 8 * g), dv = ld16(dp + q * C + 8 * g),
                ev = ld16(e + p * es + eo + 8 * g);
    const uint16_t *s8 = (const uint16_t*)&sv, *d8 = (const uint16_t*)&dv, *e8 = (const uint16_t*)&ev;
    uint4 r;
    uint16_t* r8 = (uint16_t*)&r;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float v = fmaf(0.25f, bf2f(d8[k]), bf2f(s8[k]));
        r8[k] = f2bf(bf2f(e8[k]) > 0.f ? v : 0.f);
    }
    st16(out + p * C + 8 * g, r);
}

hipError_t launch_pool_bwd_add(const uint16_t* skip, int skip_stride, int skip_off,
                               const uint16_t* dp, const uint16_t* e, int e_stride, int e_off,
                               int C, uint16_t* out, int B, int H, int W, hipStream_t s) {
    if (C % 8 || skip_stride % 8 || skip_off % 8 || e_stride % 8 || e_off % 8 || H % 2 || W % 2)
        return hipErrorInvalidValue;
    const long long n = (long long)B * H * W * (C / 8);
    hipLaunchKernelGGL(pool_bwd_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, skip,
                       skip_stride, skip_off, dp, e, e_stride, e_off, C, out, B, H, W);
    return hipGetLastError();
}

// ------------------------------------------------------------ loss / optimiser
__global__ void scale_kernel(const float* x, float scale, float* y, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i] * scale;
}

hipError_t launch_scale(const float* x, float scale, float* y, long long n, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, scale, y, n);
    return hipGetLastError();
}

int loss_blocks(long long n) { return (int)std::max(1LL, std::min(1024LL, (n + 2047) / 2048)); }

__global__ void mse_kernel(const float* y, const float* t, float* dy, float* part, long long n) {
    __shared__ float red[4];
    float s = 0.f;
    const float k = 2.f / (float)n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float d = y[i] - t[i];
        dy[i] = k * d;
        s = fmaf(d, d, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

hipError_t launch_mse(const float* y, const float* t, float* dy, float* part, long long n,
                      hipStream_t s) {
    hipLaunchKernelGGL(mse_kernel, dim3(loss_blocks(n)), dim3(256), 0, s, y, t, dy, part, n);
    return hipGetLastError();
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = fmaf(b1, m[i], (1.f - b1) * gi);
    const float vi = fmaf(b2, v[i], (1.f - b2) * gi * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s) {
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, step)));
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, g, m, v, n,
                       lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace rv
