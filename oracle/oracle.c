/*
 * CPU oracle for the gemmMfma path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -march, no fast-math,
 * so every fp32 operation rounds exactly as the reference's scalar loop does).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ff_gpu_src/ff_gpu.cpp:110-139 (InitRealData): fills dim0*dim1*dim2 floats by
 * a pattern keyed on the init_type value; idx counts every element in row-major
 * order of (dim2, dim1, dim0), padding columns included. */
void orc_init_real_data(float *x, float t, uint32_t dim0, uint32_t dim1, uint32_t dim2)
{
    if (t == 0.0f) return; /* reference leaves the malloc'd buffer untouched */
    for (uint32_t bt = 0; bt < dim2; bt++)
        for (uint32_t ch = 0; ch < dim1; ch++)
            for (uint32_t i = 0; i < dim0; i++) {
                uint32_t idx = dim0 * (dim1 * bt + ch) + i;
                float v;
                if (t == 0.5f) v = 0.0f;
                else if (t == 1.0f) v = 1.0f;
                else if (t == 2.0f) v = 1.0f * (float)idx;
                else if (t == 2.1f) v = 0.01f * (float)(idx % 1137);
                else if (t == 2.2f) v = (float)(i + 1) * 0.01f + (float)ch * 1.0f + (float)bt * 100.0f;
                else if (t == 2.3f) v = 0.1f * (float)(idx % 17);
                else if (t == 2.4f) v = 1.0f * (float)(idx % 7);
                else if (t == -1.0f) v = -1.0f;
                else if (t == -2.0f) v = -1.0f * (float)idx;
                else if (t == -2.1f) v = -0.01f * (float)(idx % 1137);
                else if (t == -2.2f) v = (float)(i + 1) * -0.01f + (float)ch * -1.0f + (float)bt * -100.0f;
                else if (t == -2.3f) v = -0.1f * (float)(idx % 17);
                else if (t == -2.4f) v = -1.0f * (float)(idx % 11);
                else v = t;
                x[idx] = v;
            }
}

/* inc/ff_utils.h:90-111 (cvtFP32toFP16): round-to-nearest-even, overflow to
 * +-inf, every NaN becomes sign|0x7e00.  Written here as a significand shift
 * with an explicit round bit and sticky remainder. */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t x = f2u(f);
    uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    uint32_t e = (x >> 23) & 0xffu, man = x & 0x7fffffu;
    if (e == 0xffu) return sign | (man ? 0x7e00u : 0x7c00u);
    int32_t E = (int32_t)e - 127 + 15;
    if (E >= 31) return sign | 0x7c00u;
    if (E >= 1) {
        uint32_t h = ((uint32_t)E << 10) | (man >> 13);
        uint32_t rest = man & 0x1fffu;
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    if (E < -10) return sign;
    uint32_t full = man | 0x800000u;
    int shift = 14 - E; /* 14..24 */
    uint32_t q = full >> shift, rem = full & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

/* IEEE binary16 -> binary32 (exact). */
float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        float v = ldexpf((float)m, -24);
        return sign ? -v : v;
    }
    if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* inc/ff_utils.h:112-118 (cvtFP16toFP32) as it actually behaves: its
 * "* pow(1.0, 112)" scale is 1.0, so the exponent field is re-biased for every
 * code including subnormals and inf/NaN.  Result: +-0 for +-0, otherwise
 * +-2^(e-15) * (1 + f/1024) with e the 5-bit field and f the 10-bit field
 * (subnormal 0x00a8 -> 3.55e-5, inf 0x7c00 -> 65536). */
float orc_f16_to_f32_quirk(uint16_t h)
{
    if ((h & 0x7fffu) == 0) return (h & 0x8000u) ? -0.0f : 0.0f;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v = ldexpf(1.0f + (float)m / 1024.0f, (int)e - 15);
    return (h & 0x8000u) ? -v : v;
}

/* inc/ff_utils.h:119-122 (cvtFP32toBF16): truncation to the high 16 bits. */
uint16_t orc_f32_to_bf16_trunc(float f) { return (uint16_t)(f2u(f) >> 16); }

/* RNE fp32 -> bf16 (the rounding the device conversions use). NaN stays NaN. */
uint16_t orc_f32_to_bf16_rne(float f)
{
    uint32_t x = f2u(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    uint32_t lsb = (x >> 16) & 1u;
    return (uint16_t)((x + 0x7fffu + lsb) >> 16);
}

/* inc/ff_utils.h:123-127 (cvtBF16toFP32). */
float orc_bf16_to_f32(uint16_t b) { return u2f((uint32_t)b << 16); }

/* OCP FP8 e4m3 ("e4m3fn"): bias 7, no inf, NaN = S.1111.111, max 448.
 * Encoding clamps to +-448 first (saturating), then rounds to nearest even. */
uint8_t orc_f32_to_e4m3(float f)
{
    uint32_t x = f2u(f);
    uint8_t sign = (uint8_t)((x >> 24) & 0x80u);
    if ((x & 0x7fffffffu) > 0x7f800000u) return sign | 0x7fu;
    float a = fabsf(f);
    if (a > 448.0f) a = 448.0f;
    x = f2u(a);
    int32_t E = (int32_t)((x >> 23) & 0xffu) - 127;
    uint32_t man = x & 0x7fffffu;
    if (E >= -6) {
        uint32_t h = ((uint32_t)(E + 7) << 3) | (man >> 20);
        uint32_t rest = man & 0xfffffu;
        if (rest > 0x80000u || (rest == 0x80000u && (h & 1u))) h++;
        if (h > 0x7eu) h = 0x7eu;
        return (uint8_t)(sign | h);
    }
    if (x == 0 || E < -10) return sign;
    uint32_t full = man | 0x800000u;
    int shift = 14 - E; /* 21..24: value in units of 2^-9 */
    uint32_t q = full >> shift, rem = full & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint8_t)(sign | q);
}

float orc_e4m3_to_f32(uint8_t q)
{
    uint32_t e = (q >> 3) & 0xfu, m = q & 7u;
    float v;
    if (e == 15 && m == 7) return u2f(0x7fc00000u | ((uint32_t)(q & 0x80u) << 24));
    if (e == 0) v = ldexpf((float)m, -9);
    else v = ldexpf(1.0f + (float)m / 8.0f, (int)e - 7);
    return (q & 0x80u) ? -v : v;
}

void orc_cvt_f32_to_f16_n(const float *s, uint16_t *d, int64_t n)
{ for (int64_t i = 0; i < n; i++) d[i] = orc_f32_to_f16(s[i]); }
void orc_cvt_f32_to_bf16_n(const float *s, uint16_t *d, int64_t n, int truncate)
{ for (int64_t i = 0; i < n; i++) d[i] = truncate ? orc_f32_to_bf16_trunc(s[i]) : orc_f32_to_bf16_rne(s[i]); }
void orc_cvt_f16_to_f32_n(const uint16_t *s, float *d, int64_t n)
{ for (int64_t i = 0; i < n; i++) d[i] = orc_f16_to_f32(s[i]); }
void orc_cvt_bf16_to_f32_n(const uint16_t *s, float *d, int64_t n)
{ for (int64_t i = 0; i < n; i++) d[i] = orc_bf16_to_f32(s[i]); }
void orc_cvt_f32_to_e4m3_n(const float *s, uint8_t *d, int64_t n, float scale)
{ for (int64_t i = 0; i < n; i++) d[i] = orc_f32_to_e4m3(s[i] * scale); }
void orc_cvt_e4m3_to_f32_n(const uint8_t *s, float *d, int64_t n)
{ for (int64_t i = 0; i < n; i++) d[i] = orc_e4m3_to_f32(s[i]); }

/* bf16x3 split (Henry, Tang & Heinecke 2019): hi = rne(a), lo = rne(a - hi). */
void orc_split_bf16x3_n(const float *s, uint16_t *hi, uint16_t *lo, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        uint16_t h = orc_f32_to_bf16_rne(s[i]);
        hi[i] = h;
        lo[i] = orc_f32_to_bf16_rne(s[i] - orc_bf16_to_f32(h));
    }
}

/* Element addressing of op(A)(m,k) and op(B)(k,n), column-major storage.
 * TN is the reference's layout: A[lda*m + k], B[ldb*n + k] (gemmMfma.cpp:334-335). */
static inline int64_t a_idx(int layout, int64_t m, int64_t k, int64_t lda)
{ return (layout >= ORC_TN) ? (k + m * lda) : (m + k * lda); }
static inline int64_t b_idx(int layout, int64_t k, int64_t n, int64_t ldb)
{ return (layout & 1) ? (n + k * ldb) : (k + n * ldb); }

/* src/gemmMfma.cpp:325-346 (GemmMfmaProblem::cpuCompute): for b, n, m: fp32
 * sum over k (k innermost), then D = alpha*sum + beta*C with each operation
 * rounded to fp32 (no fused multiply-add). Columns [n0, n1). */
static void gemm_f32_cols(int layout, int64_t M, int64_t n0, int64_t n1, int64_t K, int64_t batch,
                          float alpha, const float *A, int64_t lda, int64_t bsA, const float *B,
                          int64_t ldb, int64_t bsB, float beta, const float *C, int64_t ldc,
                          int64_t bsC, float *D, int64_t ldd, int64_t bsD)
{
    for (int64_t b = 0; b < batch; b++)
        for (int64_t n = n0; n < n1; n++)
            for (int64_t m = 0; m < M; m++) {
                float t = 0.0f;
                for (int64_t k = 0; k < K; k++) {
                    float p = A[bsA * b + a_idx(layout, m, k, lda)] * B[bsB * b + b_idx(layout, k, n, ldb)];
                    t = t + p;
                }
                float r = alpha * t;
                if (beta != 0.0f) {
                    float bc = beta * C[bsC * b + ldc * n + m];
                    r = r + bc;
                }
                D[bsD * b + ldd * n + m] = r;
            }
}

void orc_gemm_f32(int layout, int64_t M, int64_t N, int64_t K, int64_t batch, float alpha,
                  const float *A, int64_t lda, int64_t bsA, const float *B, int64_t ldb, int64_t bsB,
                  float beta, const float *C, int64_t ldc, int64_t bsC, float *D, int64_t ldd,
                  int64_t bsD)
{
    gemm_f32_cols(layout, M, 0, N, K, batch, alpha, A, lda, bsA, B, ldb, bsB, beta, C, ldc, bsC, D,
                  ldd, bsD);
}

typedef struct {
    int layout; int64_t M, n0, n1, K, batch; float alpha, beta;
    const float *A, *B, *C; float *D; int64_t lda, bsA, ldb, bsB, ldc, bsC, ldd, bsD;
} gemm_job;

static void *gemm_job_run(void *p)
{
    gemm_job *j = (gemm_job *)p;
    gemm_f32_cols(j->layout, j->M, j->n0, j->n1, j->K, j->batch, j->alpha, j->A, j->lda, j->bsA,
                  j->B, j->ldb, j->bsB, j->beta, j->C, j->ldc, j->bsC, j->D, j->ldd, j->bsD);
    return NULL;
}

void orc_gemm_f32_mt(int nthreads, int layout, int64_t M, int64_t N, int64_t K, int64_t batch,
                     float alpha, const float *A, int64_t lda, int64_t bsA, const float *B,
                     int64_t ldb, int64_t bsB, float beta, const float *C, int64_t ldc,
                     int64_t bsC, float *D, int64_t ldd, int64_t bsD)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > N) nthreads = (int)(N > 0 ? N : 1);
    pthread_t th[256];
    gemm_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        gemm_job j = {layout, M, N * t / nthreads, N * (t + 1) / nthreads, K, batch, alpha, beta,
                      A, B, C, D, lda, bsA, ldb, bsB, ldc, bsC, ldd, bsD};
        jobs[t] = j;
        pthread_create(&th[t], NULL, gemm_job_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* FP64 oracle: the same loop nest in double precision. */
void orc_gemm_f64(int layout, int64_t M, int64_t N, int64_t K, int64_t batch, double alpha,
                  const double *A, int64_t lda, int64_t bsA, const double *B, int64_t ldb,
                  int64_t bsB, double beta, const double *C, int64_t ldc, int64_t bsC, double *D,
                  int64_t ldd, int64_t bsD)
{
    for (int64_t b = 0; b < batch; b++)
        for (int64_t n = 0; n < N; n++)
            for (int64_t m = 0; m < M; m++) {
                double t = 0.0;
                for (int64_t k = 0; k < K; k++)
                    t += A[bsA * b + a_idx(layout, m, k, lda)] * B[bsB * b + b_idx(layout, k, n, ldb)];
                double r = alpha * t;
                if (beta != 0.0) r += beta * C[bsC * b + ldc * n + m];
                D[bsD * b + ldd * n + m] = r;
            }
}
