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

// K-major repack of 16-bit operands stored M/N-contiguous: element (r, k)
// of batch b, at x[b * bsx + k * ldx + r], goes to y[b * bsy + r * K + k].
// One launch repacks up to two operands (o1.blocks 0: one), in 64 x 64
// blocks: blocks = cdiv(rows, 64) * cdiv(K, 64) per batch.  Needs 16-byte
// aligned x and y, ldx, bsx and bsy multiples of 8 elements, and K a multiple
// of 8.  Any grid of 256-thread workgroups covers every block.
struct KPackOperand {
    const uint16_t* x;
    uint16_t* y;
    int64_t rows, ldx, bsx, bsy;
    uint32_t tiles_r;  // cdiv(rows, 64)
    uint32_t blocks;   // 64 x 64 blocks per batch
};
__global__ void __launch_bounds__(256) kpack16_kernel(KPackOperand o0, KPackOperand o1, int64_t K, uint32_t batch);

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
