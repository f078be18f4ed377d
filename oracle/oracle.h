/*
 * CPU oracle for the gemmMfma path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only to check results (or as the timed CPU baseline).
 * The product (tensilelite_amd, libffgemm.so) never links or calls it.
 *
 * Every function restates the reference's algorithm for this path and cites
 * the file:line it follows (paths relative to the TensileLite tree).
 *
 * Parity pins (see tests/test_oracle.py):
 *   orc_init_real_data     closed-form patterns, hand-computed values
 *   orc_f32_to_f16         bit-exact vs the reference header compiled into
 *                          oracle/_ref (when built) and vs IEEE RNE (numpy)
 *   orc_f16_to_f32_quirk   bit-exact vs oracle/_ref; closed form of the quirk
 *   orc_f16_to_f32         exhaustive 65536 codes vs numpy float16
 *   orc_f32_to_bf16_trunc  bit-exact vs oracle/_ref; closed form (bits >> 16)
 *   orc_f32_to_bf16_rne    vs torch bfloat16 conversion (RNE)
 *   orc_gemm_f32           integer KAT (exact sums), FP64 torch.matmul on CPU
 *                          for all four layouts, brute-force Python loops
 *   orc_gemm_f64           torch.matmul float64; exact integer KAT
 *   orc_f32_to_e4m3 / orc_e4m3_to_f32
 *                          all 256 codes vs the OCP e4m3fn closed form and
 *                          torch.float8_e4m3fn; round-trip; midpoint RNE
 */
#ifndef FFGEMM_ORACLE_H
#define FFGEMM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* layouts: op(A), op(B) in BLAS column-major naming */
enum { ORC_NN = 0, ORC_NT = 1, ORC_TN = 2, ORC_TT = 3 };

void orc_init_real_data(float *x, float init_type, uint32_t dim0, uint32_t dim1, uint32_t dim2);

uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
float orc_f16_to_f32_quirk(uint16_t h);
uint16_t orc_f32_to_bf16_trunc(float f);
uint16_t orc_f32_to_bf16_rne(float f);
float orc_bf16_to_f32(uint16_t b);
uint8_t orc_f32_to_e4m3(float f);
float orc_e4m3_to_f32(uint8_t q);

void orc_cvt_f32_to_f16_n(const float *src, uint16_t *dst, int64_t n);
void orc_cvt_f32_to_bf16_n(const float *src, uint16_t *dst, int64_t n, int truncate);
void orc_cvt_f16_to_f32_n(const uint16_t *src, float *dst, int64_t n);
void orc_cvt_bf16_to_f32_n(const uint16_t *src, float *dst, int64_t n);
void orc_cvt_f32_to_e4m3_n(const float *src, uint8_t *dst, int64_t n, float scale);
void orc_cvt_e4m3_to_f32_n(const uint8_t *src, float *dst, int64_t n);
void orc_split_bf16x3_n(const float *src, uint16_t *hi, uint16_t *lo, int64_t n);

/* fp32 reference GEMM (gemmMfma.cpp:325-346), column-major, layout-generalised. */
void orc_gemm_f32(int layout, int64_t M, int64_t N, int64_t K, int64_t batch, float alpha,
                  const float *A, int64_t lda, int64_t bsA, const float *B, int64_t ldb, int64_t bsB,
                  float beta, const float *C, int64_t ldc, int64_t bsC, float *D, int64_t ldd,
                  int64_t bsD);
/* same loop, columns split over nthreads pthreads (bit-identical result). */
void orc_gemm_f32_mt(int nthreads, int layout, int64_t M, int64_t N, int64_t K, int64_t batch,
                     float alpha, const float *A, int64_t lda, int64_t bsA, const float *B,
                     int64_t ldb, int64_t bsB, float beta, const float *C, int64_t ldc,
                     int64_t bsC, float *D, int64_t ldd, int64_t bsD);
/* FP64 oracle: same indexing, double accumulation, double output. */
void orc_gemm_f64(int layout, int64_t M, int64_t N, int64_t K, int64_t batch, double alpha,
                  const double *A, int64_t lda, int64_t bsA, const double *B, int64_t ldb,
                  int64_t bsB, double beta, const double *C, int64_t ldc, int64_t bsC, double *D,
                  int64_t ldd, int64_t bsD);

#ifdef __cplusplus
}
#endif
#endif
