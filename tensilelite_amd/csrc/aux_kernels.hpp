// Declarations of the auxiliary kernels (aux_kernels.hip).
#pragma once
#include "gemm_common.hpp"

namespace ffgemm {

struct F64Args {
    const double* A;
    const double* B;
    const double* C;
    double* D;
    int64_t M, N, K, lda, ldb, ldc, ldd, bsA, bsB, bsC, bsD;
    double alpha, beta;
    int transA, transB;
};

__global__ void zero_u32_kernel(uint32_t* p, int n);
__global__ void amax_kernel(const float* __restrict__ x, int64_t rows, int64_t k, int64_t ldx,
                            int trans_x, uint32_t* __restrict__ amax_bits);
__global__ void fp8_scale_kernel(const uint32_t* __restrict__ amax_bits, float* __restrict__ scale);
__global__ void fp8_quant_kernel(const float* __restrict__ x, int64_t rows, int64_t k, int64_t ldx,
                                 int trans_x, uint8_t* __restrict__ q, int64_t ldq,
                                 const float* __restrict__ scale);
__global__ void __launch_bounds__(256) x3_split_kernel(const float* __restrict__ x, int64_t rows, int64_t k,
                                                       int64_t ldx, int64_t bsx, int trans_x,
                                                       uint16_t* __restrict__ out, int64_t kp, int order);
__global__ void convert_kernel(const float* __restrict__ src, void* __restrict__ dst, int64_t n, int to);

// K-major repack of a 16-bit operand stored M/N-contiguous: element (r, k),
// at x[b * bsx + k * ldx + r], goes to y[b * bsy + r * ldy + k].  Grid:
// x = cdiv(rows, 64) * cdiv(K, 256) workgroups of 256 threads, y = batch.
// Needs 16-byte aligned x and y, ldx, bsx, ldy and bsy multiples of 8
// elements, and K a multiple of 8.
__global__ void __launch_bounds__(256) kpack16_kernel(const uint16_t* __restrict__ x, int64_t rows, int64_t K,
                                                      int64_t ldx, int64_t bsx, uint16_t* __restrict__ y,
                                                      int64_t ldy, int64_t bsy);

// Split-K reduction: D = alpha * sum_s W[s] + beta * C, converted to D's type.
// W holds the fp32 partial products of S K-slices, slice s at W + s * slice
// (M x N, leading dimension ldw, a multiple of 4).
struct SplitReduceArgs {
    const float* W;
    int64_t slice;
    int32_t S, M, N, ldw;
    const void* C;  // null when beta == 0
    void* D;
    int32_t ldc, ldd;
    float alpha, beta;
    int32_t out;    // OutKind
    int32_t trunc;  // bf16 output by truncation
};
__global__ void __launch_bounds__(256) splitk_reduce_kernel(SplitReduceArgs a);
__global__ void __launch_bounds__(256) gemm_f64_kernel(F64Args a);

}  // namespace ffgemm
