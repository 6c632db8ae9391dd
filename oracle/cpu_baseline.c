/*
 * cpu_baseline.c -- the reference CPU reconstruction, restated for timing.
 *
 * TEST AND CPU-BASELINE INFRASTRUCTURE ONLY (loaded by tests/ and by the cpu_baseline
 * leg of bench.py).  It is timed beside the GPU path as a baseline, never used
 * as a parity reference.
 *
 * What it restates:
 *  - DeconvBackProjThread's INTEL body (general.cpp:2019-2163): every thread
 *    filters every projection itself (row prep, CFft, filter, CFft, int LUT16
 *    with the CPU formula), then back-projects its band of image rows
 *    [iy0, iy1) (general.cpp:2063-2066);
 *  - projx64's AVX-512 loop (projx64.asm:249-280): t = (ix+lane)*fcos +
 *    (iy*fsin + foffset), mask 0 <= t < 16N, truncating convert, masked
 *    gather, masked int32 add; the AVX2 loop (projx64.asm:190-221) when the
 *    host lacks AVX-512, the scalar loop (general.cpp:2148-2158) otherwise;
 *  - the host sum of the per-thread partial images (gazoDoc.cpp:3974-3984).
 * MASM cannot be assembled in this environment, hence intrinsics.
 */
#include "oracle.h"
#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define RV_GINTP 16

__attribute__((target("avx512f,avx512dq")))
static void bp_rows_avx512(int32_t* ifp, const int32_t* igp, int N, float fcos, float fsin,
                           float foffset, int iy0, int iy1)
{
    const __m512 vcos = _mm512_set1_ps(fcos);
    const __m512 vng = _mm512_set1_ps((float)(N * RV_GINTP));
    const __m512 vn = _mm512_set1_ps((float)N);
    const __m512 v16 = _mm512_set1_ps(16.0f);
    const __m512 zero = _mm512_setzero_ps();
    const __m512 lane = _mm512_setr_ps(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    int iy;
    for (iy = iy0; iy < iy1; iy++) {
        const __m512 fyoff = _mm512_add_ps(_mm512_mul_ps(_mm512_set1_ps(fsin), _mm512_set1_ps((float)iy)),
                                           _mm512_set1_ps(foffset));
        int32_t* row = ifp + (size_t)iy * N;
        __m512 vx = lane;
        int ix;
        for (ix = 0; ix < N; ix += 16) {
            const __m512 t = _mm512_add_ps(_mm512_mul_ps(vcos, vx), fyoff);
            const __mmask16 k1 = _mm512_cmp_ps_mask(t, vng, _CMP_LT_OS) & _mm512_cmp_ps_mask(t, zero, _CMP_GE_OS);
            const __m512i it = _mm512_cvttps_epi32(t);
            const __m512i gv = _mm512_mask_i32gather_epi32(_mm512_setzero_si512(), k1, it, igp, 4);
            const __mmask16 k2 = _mm512_cmp_ps_mask(vx, vn, _CMP_LT_OS);
            const __m512i cur = _mm512_maskz_loadu_epi32(k2, row + ix);
            _mm512_mask_storeu_epi32(row + ix, k2, _mm512_add_epi32(gv, cur));
            vx = _mm512_add_ps(vx, v16);
        }
    }
}

__attribute__((target("avx2")))
static void bp_rows_avx2(int32_t* ifp, const int32_t* igp, int N, float fcos, float fsin,
                         float foffset, int iy0, int iy1)
{
    const __m256 vcos = _mm256_set1_ps(fcos);
    const __m256 vng = _mm256_set1_ps((float)(N * RV_GINTP));
    const __m256 vn = _mm256_set1_ps((float)N);
    const __m256 v8 = _mm256_set1_ps(8.0f);
    const __m256 zero = _mm256_setzero_ps();
    const __m256 lane = _mm256_setr_ps(0, 1, 2, 3, 4, 5, 6, 7);
    int iy;
    for (iy = iy0; iy < iy1; iy++) {
        const __m256 fyoff = _mm256_add_ps(_mm256_mul_ps(_mm256_set1_ps(fsin), _mm256_set1_ps((float)iy)),
                                           _mm256_set1_ps(foffset));
        int32_t* row = ifp + (size_t)iy * N;
        __m256 vx = lane;
        int ix;
        for (ix = 0; ix < N; ix += 8) {
            const __m256 t = _mm256_add_ps(_mm256_mul_ps(vcos, vx), fyoff);
            const __m256 m1 = _mm256_and_ps(_mm256_cmp_ps(t, vng, _CMP_LT_OS), _mm256_cmp_ps(t, zero, _CMP_GE_OS));
            const __m256i it = _mm256_cvttps_epi32(t);
            const __m256i gv = _mm256_mask_i32gather_epi32(_mm256_setzero_si256(), igp, it,
                                                           _mm256_castps_si256(m1), 4);
            const __m256i m2 = _mm256_castps_si256(_mm256_cmp_ps(vx, vn, _CMP_LT_OS));
            const __m256i cur = _mm256_maskload_epi32(row + ix, m2);
            _mm256_maskstore_epi32(row + ix, m2, _mm256_add_epi32(gv, cur));
            vx = _mm256_add_ps(vx, v8);
        }
    }
}

static void bp_rows_scalar(int32_t* ifp, const int32_t* igp, int N, float fcos, float fsin,
                           float foffset, int iy0, int iy1)
{
    const int ng = N * RV_GINTP;
    int iy;
    for (iy = iy0; iy < iy1; iy++) {
        const float fyoff = iy * fsin + foffset;
        uint32_t* row = (uint32_t*)ifp + (size_t)iy * N;
        int ix;
        for (ix = 0; ix < N; ix++) {
            const int i0 = (int)(ix * fcos + fyoff);
            if (i0 < 0 || i0 >= ng) continue;
            row[ix] += (uint32_t)igp[i0];
        }
    }
}

typedef struct {
    const int16_t* sino; int P; const float* fdeg; float tilt;
    const orc_geom* g; const float* filt; const float* wr; const float* wi;
    int32_t* img; int iy0, iy1; int isa;
} cb_job;

static void* cb_thread(void* arg)
{
    cb_job* j = (cb_job*)arg;
    const orc_geom* g = j->g;
    const int n = g->ndim, N = g->N;
    float* re = (float*)malloc(sizeof(float) * n * 6);
    float* im = re + n;
    float* scratch = re + 2 * n;
    float* px = (float*)malloc(sizeof(float) * N);
    int32_t* igp = (int32_t*)malloc(sizeof(int32_t) * (size_t)N * RV_GINTP);
    int i, k;
    for (i = 0; i < j->P; i++) {
        float c, s, o;
        orc_row_prep_i16(j->sino + (size_t)i * g->ixdim, g, re);
        memset(im, 0, sizeof(float) * n);
        orc_fft1rev(g->ndimp, j->wr, j->wi, re, im, scratch);
        for (k = 0; k < n; k++) { re[k] *= j->filt[k]; im[k] *= j->filt[k]; }
        orc_fft1(g->ndimp, j->wr, j->wi, re, im, scratch);
        for (k = 0; k < N; k++) px[k] = re[k + g->ihoffset];
        orc_lut16_i32(px, N, 0, igp);
        orc_bp_params(j->fdeg[i], j->tilt, g, 0, &c, &s, &o);
        if (j->isa == 2) bp_rows_avx512(j->img, igp, N, c, s, o, j->iy0, j->iy1);
        else if (j->isa == 1) bp_rows_avx2(j->img, igp, N, c, s, o, j->iy0, j->iy1);
        else bp_rows_scalar(j->img, igp, N, c, s, o, j->iy0, j->iy1);
    }
    free(igp); free(px); free(re);
    return NULL;
}

int orc_cpu_isa(void)
{
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq")) return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
    return 0;
}

/* Reconstructs one int16 slice the way the reference CPU path does, on
 * nthreads threads.  isa < 0 picks the best the host supports.  Returns the
 * wall time in seconds of filter + BP + partial-image sum (the reference timer
 * window, gazoDoc.cpp:3873-3992), or a negative value on allocation failure. */
double orc_cpu_baseline(const int16_t* sino, int P, const float* fdeg, float tilt,
                        const orc_geom* g, const float* filt, int nthreads, int isa,
                        int32_t* img)
{
    const int N = g->N, n = g->ndim;
    const size_t n2 = (size_t)N * N;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    cb_job* jobs = (cb_job*)malloc(sizeof(cb_job) * nthreads);
    int32_t** part = (int32_t**)calloc(nthreads, sizeof(int32_t*));
    float* wr = (float*)malloc(sizeof(float) * n / 2);
    float* wi = (float*)malloc(sizeof(float) * n / 2);
    struct timespec t0, t1;
    int t;
    size_t q;
    if (!th || !jobs || !part || !wr || !wi) return -1.0;
    if (isa < 0) isa = orc_cpu_isa();
    orc_cfft_twiddles(g->ndimp, wr, wi);
    part[0] = img;
    for (t = 1; t < nthreads; t++) {
        part[t] = (int32_t*)aligned_alloc(64, ((sizeof(int32_t) * n2 + 63) / 64) * 64);
        if (!part[t]) return -1.0;
    }
    for (t = 0; t < nthreads; t++) memset(part[t], 0, sizeof(int32_t) * n2);
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (t = 0; t < nthreads; t++) {
        cb_job* j = &jobs[t];
        j->sino = sino; j->P = P; j->fdeg = fdeg; j->tilt = tilt; j->g = g; j->filt = filt;
        j->wr = wr; j->wi = wi; j->img = part[t]; j->isa = isa;
        j->iy0 = (N * t + nthreads - 1) / nthreads;          /* general.cpp:2065-2066 */
        j->iy1 = (N * (t + 1) + nthreads - 1) / nthreads;
        pthread_create(&th[t], NULL, cb_thread, j);
    }
    for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    for (t = 1; t < nthreads; t++) {                          /* gazoDoc.cpp:3974-3984 */
        uint32_t* dst = (uint32_t*)img;
        const uint32_t* src = (const uint32_t*)part[t];
        for (q = 0; q < n2; q++) dst[q] += src[q];
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    for (t = 1; t < nthreads; t++) free(part[t]);
    free(part); free(jobs); free(th); free(wr); free(wi);
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
