// Auxiliary device kernels: FP8 per-tensor quantisation, fp32 -> bf16/fp16
// conversion, and the FP64 MFMA GEMM used as the error oracle.
#include "aux_kernels.hpp"

namespace ffgemm {

// Workspace layout for FP8 quantisation: [0] amax bits (uint32).
__global__ void zero_u32_kernel(uint32_t* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// max |x| over a rows x k matrix with leading dimension ldx; element (r, kk)
// at kk + r*ldx when trans_x (K-contiguous) else r + kk*ldx.  The order of
// traversal does not matter for a maximum.
__global__ void amax_kernel(const float* __restrict__ x, int64_t rows, int64_t k, int64_t ldx,
                            int trans_x, uint32_t* __restrict__ amax_bits) {
    const int64_t inner = trans_x ? k : rows;
    const int64_t outer = trans_x ? rows : k;
    const int64_t total = inner * outer;
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = i / inner, in = i - o * inner;
        float v = fabsf(x[o * ldx + in]);
        m = v > m ? v : m;  // NaN inputs are ignored
    }
    // wave reduction then one atomic per wave (non-negative floats order like their bits)
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(amax_bits, __builtin_bit_cast(uint32_t, m));
}

__global__ void fp8_scale_kernel(const uint32_t* __restrict__ amax_bits, float* __restrict__ scale) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float a = __builtin_bit_cast(float, *amax_bits);
        scale[0] = (a > 0.0f) ? 448.0f / a : 1.0f;
    }
}

__device__ inline uint8_t f32_to_e4m3(float v) {
    // saturate to the e4m3 range, then the hardware RNE conversion
    v = fminf(fmaxf(v, -448.0f), 448.0f);
    int p = __builtin_amdgcn_cvt_pk_fp8_f32(v, 0.0f, 0, false);
    return (uint8_t)(p & 0xff);
}

// q[kk + r*ldq] = e4m3(x(r, kk) * scale).  32 x 32 tiles through LDS so both
// the read and the write are coalesced when the layouts differ.
__global__ void fp8_quant_kernel(const float* __restrict__ x, int64_t rows, int64_t k, int64_t ldx,
                                 int trans_x, uint8_t* __restrict__ q, int64_t ldq,
                                 const float* __restrict__ scale) {
    __shared__ float tile[32][33];
    const float s = *scale;
    const int64_t r0 = (int64_t)blockIdx.y * 32, k0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int yy = ty; yy < 32; yy += 8) {
        int64_t r, kk;
        if (trans_x) { r = r0 + yy; kk = k0 + tx; } else { r = r0 + tx; kk = k0 + yy; }
        float v = 0.0f;
        if (r < rows && kk < k) v = trans_x ? x[r * ldx + kk] : x[kk * ldx + r];
        if (trans_x) tile[yy][tx] = v; else tile[tx][yy] = v;
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        int64_t r = r0 + yy, kk = k0 + tx;
        if (r < rows && kk < k) q[r * ldq + kk] = f32_to_e4m3(tile[yy][tx] * s);
    }
}

// bf16x3 operand planes (Henry, Tang & Heinecke 2019): for element (r, kk) of
// a rows x k fp32 matrix (trans_x: at kk + r*ldx, else r + kk*ldx),
// hi = rne_bf16(x) and lo = rne_bf16(x - hi).  Row r of the output holds three
// K-segments of kp elements (kp a multiple of 64), out[r*3kp + s*kp + kk] =
// (order >> s) & 1 ? lo : hi, zero for kk in [k, kp), so hi.hi + hi.lo + lo.hi
// of two operands is one bf16 GEMM over K = 3kp.  A block covers 64 rows x
// 64 kk; each thread converts 8 consecutive kk of one row and stores them as
// one 16-byte vector per segment.  M-contiguous input goes through an LDS
// transpose so the fp32 reads stay coalesced.  blockIdx.z is the batch.
__device__ inline void x3_store8(const float* v, uint16_t* o, int64_t kp, int order) {
    uint32_t hw[4], lw[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const __bf16 h0 = (__bf16)v[2 * e], h1 = (__bf16)v[2 * e + 1];
        const __bf16 l0 = (__bf16)(v[2 * e] - (float)h0), l1 = (__bf16)(v[2 * e + 1] - (float)h1);
        hw[e] = (uint32_t)__builtin_bit_cast(uint16_t, h0) | ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
        lw[e] = (uint32_t)__builtin_bit_cast(uint16_t, l0) | ((uint32_t)__builtin_bit_cast(uint16_t, l1) << 16);
    }
    const i32x4_t hv = {(int)hw[0], (int)hw[1], (int)hw[2], (int)hw[3]};
    const i32x4_t lv = {(int)lw[0], (int)lw[1], (int)lw[2], (int)lw[3]};
    *(i32x4_t*)o = (order & 1) ? lv : hv;
    *(i32x4_t*)(o + kp) = (order & 2) ? lv : hv;
    *(i32x4_t*)(o + 2 * kp) = (order & 4) ? lv : hv;
}

__global__ void __launch_bounds__(256) x3_split_kernel(const float* __restrict__ x, int64_t rows, int64_t k,
                                                       int64_t ldx, int64_t bsx, int trans_x,
                                                       uint16_t* __restrict__ out, int64_t kp, int order) {
    __shared__ float tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 64;
    const float* xb = x + (int64_t)blockIdx.z * bsx;
    uint16_t* ob = out + (int64_t)blockIdx.z * rows * 3 * kp;
    const int t = threadIdx.x;
    const int kc = t & 7;  // this thread's 8-element K chunk
    // 16-byte loads where the operand's base and leading dimension keep them aligned
    const bool vec = (((uintptr_t)xb | (uintptr_t)(ldx * 4)) & 15) == 0;
    if (!trans_x) {
        // M-contiguous: 64 kk columns of 64 rows, each thread 4 consecutive rows of one kk
        const bool rows_in = vec && r0 + 64 <= rows;
        for (int i = t; i < 64 * 16; i += 256) {
            const int kk = i >> 4, rq = (i & 15) * 4;
            const int64_t kg = k0 + kk;
            float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (kg < k) {
                const float* src = xb + kg * ldx + r0 + rq;
                if (rows_in) {
                    const float4 q = *(const float4*)src;
                    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = (r0 + rq + e < rows) ? src[e] : 0.0f;
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) tile[rq + e][kk] = v[e];
        }
        __syncthreads();
    }
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        const int rl = (t >> 3) + 32 * pass;
        const int64_t r = r0 + rl, kk = k0 + kc * 8;
        if (r >= rows) continue;
        float v[8];
        if (trans_x) {
            const float* src = xb + r * ldx + kk;
            if (vec && kk + 8 <= k) {
                const float4 q0 = *(const float4*)src, q1 = *(const float4*)(src + 4);
                v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w;
                v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (kk + e < k) ? src[e] : 0.0f;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = tile[rl][kc * 8 + e];
        }
        x3_store8(v, ob + r * 3 * kp + kk, kp, order);
    }
}

// Each wave moves 64 x 64 blocks of (r, k), one after another (block t,
// t + 4 * gridDim.x, ...), loading its next block before storing the current
// one, so loads and stores of the launch overlap.  Within a block, lane
// (a, c) = (lane & 7, lane >> 3) moves the 8 x 8 piece at rows 8a.., k 8c..:
// eight 16-byte loads along r (one per k), a transpose in registers, eight
// 16-byte stores along k (one per r), so every load instruction reads eight
// whole 128-byte runs of x and every store writes eight of y.  Blocks run
// along r first (neighbouring waves read neighbouring runs of the same rows
// of x).  Loads go through the caches normally: non-temporal loads measured
// 25-35 % slower here on MI355X.
__global__ void __launch_bounds__(256) kpack16_kernel(KPackOperand o0, KPackOperand o1, int64_t K, uint32_t batch) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const uint32_t n0 = o0.blocks * batch, total = n0 + o1.blocks * batch;
    const uint32_t stride = gridDim.x * 4u;
    uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (t >= total) return;  // no barriers below
    struct Piece {
        const uint16_t* src;
        uint16_t* dst;
        int64_t ldx;
        int nr;  // rows of this lane's piece inside the operand (0: none)
    };
    auto locate = [&](uint32_t u) {
        const bool second = u >= n0;
        if (second) u -= n0;
        const uint32_t blocks = second ? o1.blocks : o0.blocks, tiles_r = second ? o1.tiles_r : o0.tiles_r;
        const uint32_t bi = u / blocks, rem = u - bi * blocks;
        const uint32_t tk = rem / tiles_r, tr = rem - tk * tiles_r;
        const int64_t rows = second ? o1.rows : o0.rows;
        const int64_t r0 = (int64_t)tr * 64 + 8 * (lane & 7), k0 = (int64_t)tk * 64 + 8 * (lane >> 3);
        Piece q;
        q.ldx = second ? o1.ldx : o0.ldx;
        q.nr = (r0 < rows && k0 < K) ? (int)(rows - r0 < 8 ? rows - r0 : 8) : 0;  // K % 8 == 0: k0 + 7 < K
        q.src = (second ? o1.x : o0.x) + (int64_t)bi * (second ? o1.bsx : o0.bsx) + k0 * q.ldx + r0;
        q.dst = (second ? o1.y : o0.y) + (int64_t)bi * (second ? o1.bsy : o0.bsy) + r0 * K + k0;
        return q;
    };
    auto load = [&](const Piece& q, u32x4 (&v)[8]) {
        if (q.nr == 8) {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = *(const u32x4*)(q.src + i * q.ldx);
        } else if (q.nr > 0) {
            // a short last block: element loads, nothing past the operand's last row
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                uint16_t e[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) e[j] = j < q.nr ? q.src[i * q.ldx + j] : (uint16_t)0;
#pragma unroll
                for (int h = 0; h < 4; ++h) v[i][h] = (uint32_t)e[2 * h] | ((uint32_t)e[2 * h + 1] << 16);
            }
        }
    };
    // row j of the piece: element k is element j of v[k]
    auto store = [&](const Piece& q, const u32x4 (&v)[8]) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j >= q.nr) break;
            u32x4 o;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint32_t lo = v[2 * h][j >> 1], hi = v[2 * h + 1][j >> 1];
                o[h] = (j & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
            }
            *(u32x4*)(q.dst + j * K) = o;
        }
    };
    Piece cur = locate(t);
    u32x4 in[8];
    load(cur, in);
    for (;;) {
        const bool more = total - t > stride;  // t + stride < total, without overflow
        Piece nxt;
        u32x4 nin[8];
        if (more) {
            nxt = locate(t + stride);
            load(nxt, nin);
        }
        store(cur, in);
        if (!more) break;
        t += stride;
        cur = nxt;
#pragma unroll
        for (int i = 0; i < 8; ++i) in[i] = nin[i];
    }
}

__global__ void convert_kernel(const float* __restrict__ src, void* __restrict__ dst, int64_t n, int to) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = src[i];
        if (to == 3) ((__bf16*)dst)[i] = (__bf16)v;
        else ((_Float16*)dst)[i] = (_Float16)v;
    }
}

// Split-K reduction.  Thread t of the flattened grid takes 4 consecutive rows
// m of one column n (the W slices are read as 16-byte vectors), sums the S
// partials in slice order (deterministic), applies alpha / beta with C and
// writes D in its type; columns whose last group of 4 passes M store element
// by element.  HBM-bound: S * 16 bytes read per 4 outputs.
__global__ void __launch_bounds__(256) splitk_reduce_kernel(SplitReduceArgs a) {
    const int64_t mq = (a.M + 3) / 4;
    const int64_t total = mq * a.N;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int n = (int)(t / mq);
        const int m = (int)(t - (int64_t)n * mq) * 4;
        const float* w = a.W + (int64_t)n * a.ldw + m;
        f32x4_t s = *(const f32x4_t*)w;
        for (int k = 1; k < a.S; ++k) {
            const f32x4_t v = *(const f32x4_t*)(w + (int64_t)k * a.slice);
            s[0] += v[0]; s[1] += v[1]; s[2] += v[2]; s[3] += v[3];
        }
        const int nv = min(4, a.M - m);
        float r[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float c = 0.0f;
            if (a.C && e < nv) {
                const int64_t ci = (int64_t)n * a.ldc + m + e;
                c = a.out == kOutF32 ? ((const float*)a.C)[ci]
                  : a.out == kOutBF16 ? (float)((const __bf16*)a.C)[ci] : (float)((const _Float16*)a.C)[ci];
            }
            r[e] = a.C ? __builtin_fmaf(a.alpha, s[e], a.beta * c) : a.alpha * s[e];
        }
        const int64_t di = (int64_t)n * a.ldd + m;
        if (a.out == kOutF32) {
            float* d = (float*)a.D + di;
            if (nv == 4) *(f32x4_t*)d = f32x4_t{r[0], r[1], r[2], r[3]};
            else for (int e = 0; e < nv; ++e) d[e] = r[e];
        } else {
            uint32_t p0, p1;
            if (a.out == kOutBF16) {
                p0 = a.trunc ? pack_bf16x2_trunc(r[0], r[1]) : pack_bf16x2_rne(r[0], r[1]);
                p1 = a.trunc ? pack_bf16x2_trunc(r[2], r[3]) : pack_bf16x2_rne(r[2], r[3]);
            } else {
                p0 = pack_f16x2(r[0], r[1]);
                p1 = pack_f16x2(r[2], r[3]);
            }
            uint16_t* d = (uint16_t*)a.D + di;
            if (nv == 4) *(i32x2_t*)d = i32x2_t{(int)p0, (int)p1};
            else {
                const uint16_t h[4] = {(uint16_t)p0, (uint16_t)(p0 >> 16), (uint16_t)p1, (uint16_t)(p1 >> 16)};
                for (int e = 0; e < nv; ++e) d[e] = h[e];
            }
        }
    }
}

// FP64 oracle GEMM on v_mfma_f64_16x16x4_f64: 64 x 64 tile per 256-thread
// block, each wave a 32 x 32 quarter (2 x 2 MFMA tiles), K staged through LDS
// 16 at a time.  Same column-major problem and layouts as the fp32 path.
__global__ void __launch_bounds__(256) gemm_f64_kernel(F64Args a) {
    __shared__ double As[64][17];
    __shared__ double Bs[64][17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
    const int b = blockIdx.z;
    const double* A = a.A + (int64_t)b * a.bsA;
    const double* B = a.B + (int64_t)b * a.bsB;
    f64x4_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 4; ++e) acc[i][j][e] = 0.0;
    for (int64_t k0 = 0; k0 < a.K; k0 += 16) {
        for (int idx = threadIdx.x; idx < 64 * 16; idx += 256) {
            int r = idx & 63, kk = idx >> 6;
            int64_t m = m0 + r, n = n0 + r, k = k0 + kk;
            double av = 0.0, bv = 0.0;
            if (m < a.M && k < a.K) av = a.transA ? A[k + m * a.lda] : A[m + k * a.lda];
            if (n < a.N && k < a.K) bv = a.transB ? B[n + k * a.ldb] : B[k + n * a.ldb];
            As[r][kk] = av;
            Bs[r][kk] = bv;
        }
        __syncthreads();
        for (int ks = 0; ks < 16; ks += 4) {
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) {
                    // lane l supplies A[row l%16][k l/16] and B[k l/16][col l%16]
                    double av = As[wm * 32 + i * 16 + (lane & 15)][ks + (lane >> 4)];
                    double bv = Bs[wn * 32 + j * 16 + (lane & 15)][ks + (lane >> 4)];
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    double* D = a.D + (int64_t)b * a.bsD;
    const double* C = a.C ? a.C + (int64_t)b * a.bsC : nullptr;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 4; ++e) {
                // output row l/16 + 4*e, column l%16 of the 16 x 16 tile
                int64_t m = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * e;
                int64_t n = n0 + wn * 32 + j * 16 + (lane & 15);
                if (m < a.M && n < a.N) {
                    double r = a.alpha * acc[i][j][e];
                    if (a.beta != 0.0) r += a.beta * C[m + n * a.ldc];
                    D[m + n * a.ldd] = r;
                }
            }
}

}  // namespace ffgemm
