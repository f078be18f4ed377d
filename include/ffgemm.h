/*
 * ffgemm -- MI355X (gfx950) GEMM for the TensileLite gemmMfma problem.
 *
 * C ABI of libffgemm.so.  Plain pointers and sizes only; every matrix pointer
 * is a device pointer owned by the caller.  The library never allocates or
 * frees caller buffers; it owns only its grouped-GEMM plans (created and
 * destroyed explicitly).
 *
 * Problem (column-major, as the reference's gemmMfma problem):
 *     D = alpha * op(A) * op(B) + beta * C
 *     op(A) is M x K, op(B) is K x N, C and D are M x N with leading
 *     dimensions ldc / ldd.  Layout names follow BLAS: the first letter is
 *     op(A), the second op(B); "T" means the operand is stored transposed.
 *     The reference's only layout is TN: A is lda x M (K contiguous), B is
 *     ldb x N (K contiguous), C/D are ldc x N (M contiguous)
 *     (src/gemmMfma.cpp:210-217, 334-342).
 *
 * Status codes mirror E_ReturnState (inc/ff_utils.h:44-50).  No call exits
 * the process; the last error message is kept per thread
 * (ffgemm_last_error()).
 *
 * Every entry point taking a stream is asynchronous and stream-ordered,
 * performs no allocation and no host synchronisation, and may be captured
 * into a hipGraph.
 */
#ifndef FFGEMM_H
#define FFGEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* E_ReturnState: SUCCESS, RTN_WARN, RTN_ERR, RTN_FATAL (inc/ff_utils.h:44-50). */
typedef enum {
    FFGEMM_SUCCESS = 0,
    FFGEMM_WARN = 1,
    FFGEMM_ERR = 2,   /* invalid or unsupported problem; nothing launched */
    FFGEMM_FATAL = 3  /* HIP runtime failure */
} ffgemm_status;

typedef enum { FFGEMM_NN = 0, FFGEMM_NT = 1, FFGEMM_TN = 2, FFGEMM_TT = 3 } ffgemm_layout;

/* 1..3 are the reference's -d codes (src/gemmMfma.cpp:55-57). */
typedef enum {
    FFGEMM_F32 = 1,
    FFGEMM_F16 = 2,
    FFGEMM_BF16 = 3,
    FFGEMM_F8E4M3 = 4, /* OCP e4m3 (e4m3fn), TN only; see ffgemm_fp8_quantize */
    FFGEMM_F64 = 5     /* FP64 MFMA product, the oracle for error reports */
} ffgemm_dtype;

typedef enum {
    FFGEMM_COMPUTE_DEFAULT = 0, /* f16/bf16/f8: MFMA with fp32 accumulate; f32: fp32 MFMA */
    FFGEMM_COMPUTE_BF16X3 = 1   /* f32 in/out: hi/lo bf16 split, three MFMA products */
} ffgemm_compute;

/* Output rounding for bf16 D. */
typedef enum {
    FFGEMM_ROUND_RNE = 0,     /* round to nearest even (default) */
    FFGEMM_ROUND_TRUNCATE = 1 /* the reference's bf16 store (gemmMfmaKernelWriter.h:1661-1664) */
} ffgemm_round;

typedef struct ffgemm_problem {
    int32_t layout;    /* ffgemm_layout */
    int32_t dtype;     /* ffgemm_dtype of A and B */
    int32_t dtype_out; /* dtype of C and D; 0 = same as dtype (f8 -> f32) */
    int32_t compute;   /* ffgemm_compute */
    int32_t round;     /* ffgemm_round (bf16 output only) */
    int32_t reserved;
    int64_t M, N, K, batch; /* batch 0 is treated as 1 */
    int64_t lda, ldb, ldc, ldd;
    int64_t bsA, bsB, bsC, bsD; /* batch strides in elements */
    float alpha, beta;
    /* FP8 only: device pointers to the fp32 per-tensor quantisation scales of
     * A and B (q = x * scale); D is multiplied by 1/(scaleA*scaleB).
     * NULL means 1.0. */
    const float *scaleA;
    const float *scaleB;
} ffgemm_problem;

/* Tuning record; the fields the reference's -r -s -x -y -z -u -f flags set
 * (src/gemmMfmaKernelWriter.h:11-22) plus gfx950 knobs.  A zeroed record (or
 * a NULL pointer) asks the library to pick. */
typedef struct ffgemm_tile {
    int32_t mt_m;     /* macro tile along M (64, 128, 256) */
    int32_t mt_n;     /* macro tile along N (64, 128, 256) */
    int32_t depth_u;  /* K per main-loop iteration (fixed by dtype: 128 bytes of K) */
    int32_t waves_m;  /* waves along M (2) */
    int32_t waves_n;  /* waves along N (2) */
    int32_t mfma_mn;  /* MFMA M/N width: 32 (default), or 16 for 16x16 MFMAs
                       * (16x16x32 bf16/fp16, 16x16x4 fp32; at 64x64 and
                       * 128x128; -z of the reference) */
    int32_t group_m;  /* tile-raster group height (0 = default) */
    int32_t lds_buffers; /* 2 */
    int32_t split_k;  /* K slices (split-K), summed by a reduction pass: 0 = the
                       * library's choice, 1 = none, S >= 2 = S slices.  Slices
                       * need the fp32 partial planes in a workspace, so only
                       * ffgemm_run_ws splits (single problems, not bf16x3);
                       * ffgemm_run takes 0 or 1. */
    int32_t repack;   /* K-major repack (bf16/fp16, layouts NN, NT, TT): 0 = the
                       * library's choice, 1 = never, 2 = always.  An operand
                       * stored M/N-contiguous is copied K-contiguous into the
                       * workspace and the problem runs as TN, on the engine
                       * the K-contiguous layout has; only ffgemm_run_ws
                       * repacks (it needs K a multiple of 8, 16-byte aligned
                       * strides and pointers, and batch <= 65535). */
} ffgemm_tile;

/* Library / device bring-up.  Optional: every entry point initialises lazily. */
int ffgemm_init(void);
void ffgemm_shutdown(void);
const char *ffgemm_status_string(int status);
const char *ffgemm_last_error(void);
const char *ffgemm_version(void);

/* Checks a problem and the tile the library would use for it; no launch.
 * Replaces GemmMfmaKernelWriter::checkKernelParameters
 * (src/gemmMfmaKernelWriter.h:321-405). */
int ffgemm_query_tile(const ffgemm_problem *p, ffgemm_tile *out);

/* Maps the reference's tuning record onto a tile record that runs problem p:
 * MFMA tiles per wave r, s (-r -s), waves x, y (-x -y), MFMA width z (-z, 16
 * or 32), DepthU (-u) and LDS buffers (-f), i.e. MT0 = z*r*x, MT1 = z*s*y
 * (src/gemmMfma.cpp:30-66, src/gemmMfmaKernelWriter.h:11-22).  Returns
 * SUCCESS when a kernel has exactly that record; RTN_WARN when it runs as the
 * nearest supported record instead: the same macro tile with the engine's
 * own wave tiling, DepthU and LDS stages, or else the smallest macro tile
 * covering MT0 x MT1 (256 x 256 when none does) with 32 x 32 MFMAs.  *out
 * receives the record that runs, as ffgemm_query_tile reports it; the
 * message (ffgemm_last_error) names what changed.  RTN_ERR: a record out of
 * range, or a problem that cannot run.  Where the reference refuses a record
 * for this problem (checkKernelParameters, src/gemmMfmaKernelWriter.h:
 * 392-399), this build still runs it: the supported records are a superset. */
int ffgemm_tile_from_reference(const ffgemm_problem *p, int r, int s, int x, int y, int z, int depth_u,
                               int lds_buffers, ffgemm_tile *out);

/* Loads the tuning table that problems run without a tile consult: the
 * autotuner's best tile per (dtype, layout, M, N, K), one CSV line each,
 * `dtype,layout,M,N,K,mt_m,mt_n,group_m,waves_m,waves_n,mfma_mn,split_k,tflops`
 * (a line of 12 fields has no split_k: no split; of 11 fields no mfma_mn
 * either: 32 x 32 MFMAs; an entry's split_k applies on the workspace path).  It
 * replaces the table in use.  path NULL: the default, the file named by
 * FFGEMM_TUNE_DB, else tune_db.csv in the directory above the library's;
 * "" or a missing file: an empty table.  Without this call the default is
 * loaded on first use.  The table applies to single problems in the default
 * compute and output type; an entry that does not fit its problem is passed
 * over.  *entries (nullable) receives the entry count.  Replaces the
 * best-of selection of the reference's sweep (test/perf_test.py:58-167),
 * which picks a configuration by hand from the results sheet. */
int ffgemm_tune_db_load(const char *path, int *entries);

/* D = alpha*op(A)*op(B) + beta*C on `stream` (hipStream_t, NULL = null stream).
 * Replaces the path GemmMfmaAsmSolution::generateKernel -> StreamHip::Launch
 * (src/gemmMfma.cpp:24-185, ff_gpu_src/ff_hip_runtime.cpp:278-320).
 * `tile` may be NULL.  beta == 0 never reads C.  C may be the same buffer as
 * D when ldc == ldd and bsC == bsD (an in-place update: each element of C is
 * read by the thread that later writes it); other overlaps are not supported.
 * Leading dimensions and batch strides of any length, any K, and pointers
 * aligned to their element size are accepted.  Problems with every ld, batch
 * stride and pointer on the 16-byte grid (and K a multiple of 16 bytes of
 * elements for a K-contiguous operand) run the vector engines; any other runs
 * the general-stride kernel (element-wise loads and stores, MT64x64 or
 * MT128x128 with 2x2 waves and 2 LDS stages; fp8 MT128x128), which a tile
 * record can ask for and which never splits K.  Grouped plans take grid-
 * aligned problems only. */
int ffgemm_run(const ffgemm_problem *p, const void *A, const void *B, const void *C, void *D,
               const ffgemm_tile *tile, void *stream);

/* Device workspace (bytes) that ffgemm_run_ws can use for this problem; 0
 * when the problem has no workspace path.  Two uses, neither in the reference:
 *  - bf16x3: the fp32 operands are split once into bf16 hi/lo planes
 *    ((M + N) rows x 3*roundup(K, 64) bf16 per batch) and the three products
 *    run as one bf16 MFMA GEMM over K' = 3*roundup(K, 64); without a
 *    workspace they are split in registers inside the GEMM;
 *  - split-K: a single problem with too few output tiles to fill the CUs
 *    (by the library's cost model, or a tile record with split_k >= 2) cuts K
 *    into S slices whose fp32 partial products (S planes of roundup(M, 4) x N)
 *    are summed by a reduction pass; without a workspace it runs unsplit;
 *  - tail split: with no tile given (and no tuning-table entry), a problem
 *    whose 256 x 256 tiles fill whole passes over the CUs plus a short last
 *    pass runs as two sub-problems, the whole passes as one launch and the
 *    remaining tile columns (or rows) with split-K; the workspace holds the
 *    remainder's planes.  ffgemm_query_tile_ws then still reports the
 *    single-launch record;
 *  - K-major repack: a bf16/fp16 problem in layout NN, NT or TT copies its
 *    M/N-contiguous operand(s) K-contiguous (rows x K per batch) into the
 *    front of the workspace and runs as TN, through this same path, with the
 *    rest of the workspace for the TN problem's own plan.  Taken with no tile
 *    given (and no tuning-table entry) when the library's cost model puts it
 *    ahead, or when a tile record asks for it (repack 2).  No counterpart in
 *    the reference, whose only layout is TN. */
int ffgemm_workspace_size(const ffgemm_problem *p, size_t *bytes);

/* ffgemm_workspace_size for the plan ffgemm_run_ws follows with this tile
 * record (NULL: the library's choice, as ffgemm_workspace_size); a record
 * with repack 2 or split_k >= 2 and mt_m 0 keeps the library's tile. */
int ffgemm_workspace_size_ex(const ffgemm_problem *p, const ffgemm_tile *tile, size_t *bytes);

/* The tile record ffgemm_run_ws follows for this problem given a workspace of
 * ffgemm_workspace_size bytes, split_k and repack included (ffgemm_query_tile
 * reports the ffgemm_run plan, which never splits or repacks; a repacking
 * plan reports the TN problem's record with repack 2).  *out as input: a tile
 * to complete, or zeros. */
int ffgemm_query_tile_ws(const ffgemm_problem *p, ffgemm_tile *out);

/* ffgemm_run with a caller-owned device workspace (256-byte aligned).  Falls
 * back to ffgemm_run when the workspace is NULL or smaller than
 * ffgemm_workspace_size reports (RTN_ERR instead when the tile record asks
 * for split_k >= 2 or repack 2).  Calls on one workspace must not overlap. */
int ffgemm_run_ws(const ffgemm_problem *p, const void *A, const void *B, const void *C, void *D,
                  const ffgemm_tile *tile, void *workspace, size_t workspace_bytes, void *stream);

/* bf16x3 without any allocation: ffgemm_run_ws with the hi/lo planes in a
 * device array inside the library instead of a caller workspace.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The array holds FFGEMM_NOALLOC_SCRATCH_BYTES (64 MiB) per
 * device, from the library's load until the process ends; its address is
 * looked up once per device (ffgemm_init, or the first call on a device).
 * Only bf16x3 problems (F32 operands, COMPUTE_BF16X3); a problem whose planes,
 * (M + N) * 3 * roundup(K, 64) * 2 bytes per batch (each operand's planes
 * rounded up to 256 bytes), exceed the array is refused with RTN_ERR before
 * anything runs.  Problems with no planes path run the in-register split,
 * which needs no scratch.  Every call on one GPU uses the same array: calls
 * that overlap on one GPU (other streams, other threads) race; calls ordered
 * on one stream do not.  Allocates nothing, with the results of ffgemm_run_ws
 * for the same problem and tile. */
#define FFGEMM_NOALLOC_SCRATCH_BYTES (64u << 20)
int ffgemm_run_x3_noalloc(const ffgemm_problem *p, const void *A, const void *B, const void *C, void *D,
                          const ffgemm_tile *tile, void *stream);

/* The reference kernel's argument list, in its kernarg order
 * D, C, A, B, alpha, beta, ldd, bsD, ldc, bsC, lda, bsA, ldb, bsB
 * (src/gemmMfma.cpp:106-119, src/gemmMfmaKernelWriter.h:323-338), TN layout,
 * plus the sizes and dtype that the reference baked into the generated kernel.
 * dtype is the reference's -d code: 1 fp32, 2 fp16, 3 bf16. */
int ffgemm_run_tn_kernarg(void *D, const void *C, const void *A, const void *B, float alpha,
                          float beta, uint32_t ldd, uint32_t bsD, uint32_t ldc, uint32_t bsC,
                          uint32_t lda, uint32_t bsA, uint32_t ldb, uint32_t bsB, uint32_t M,
                          uint32_t N, uint32_t K, uint32_t batch, int dtype,
                          const ffgemm_tile *tile, void *stream);

/* The reference's Tensile-layout kernels (-t 1: Cijk_Alik_Bljk_{SB,HBH,BBH},
 * src/gemmMfma.cpp:72-102) take a 148-byte argument buffer: sizeC, sizeA,
 * sizeB (u64, elements), D, C, A, B, alpha, beta (32-bit slots), strideD0,
 * strideD1, StrideC0, StrideC1, StrideA0, StrideA1, StrideB0, StrideB1,
 * SizesFree0..2 (M, N, batch), SizesSum0 (K), then eight grid-mapping values
 * and a pad.  This entry runs the TN problem such a buffer describes, for a
 * caller that builds it: alpha/beta are fp16 in the low half of their slots
 * for F16 (as the reference's HBH kernels read them,
 * src/gemmMfmaKernelWriter.h:1845-1854) and fp32 otherwise; nonzero sizes
 * must cover what the problem addresses; the grid-mapping values are ignored
 * (the library maps its own grid).  dtype: FFGEMM_F32, F16 or BF16. */
int ffgemm_run_tensile_args(const void *kernarg, size_t kernarg_bytes, int dtype, const ffgemm_tile *tile,
                            void *stream);

/* FP8 e4m3 per-tensor quantisation: scale = 448 / max|x| (fp32, computed on
 * the device into *scale_out), q = rne_sat(x * scale).  x is an M x K (or
 * K x N) fp32 matrix given in `layout_of_x` terms: for operand 'A' with
 * trans_x = 0 the input is M-contiguous (lda x K), trans_x = 1 K-contiguous.
 * q is written K-contiguous ("T" for A, "N" for B) with leading dimension
 * ldq, as the FP8 GEMM (TN only) consumes it.  `work` is a device scratch of
 * at least ffgemm_fp8_workspace_size() bytes. */
size_t ffgemm_fp8_workspace_size(void);
int ffgemm_fp8_quantize(const float *x, int64_t rows, int64_t k, int64_t ldx, int trans_x,
                        uint8_t *q, int64_t ldq, float *scale_out, void *work, void *stream);

/* Elementwise fp32 -> bf16/fp16 conversion on the device (RNE), n elements. */
int ffgemm_convert(const float *src, void *dst, int64_t n, int dtype_dst, void *stream);

/* Grouped GEMM: `count` independent problems (same dtype, layout, compute)
 * run by one persistent launch.  The grouped kernels are 16-bit: bf16 or fp16
 * operands with D in the same type and the default compute; any other
 * problem (fp32, bf16x3, fp8, fp64, a different output type) or batch > 1 is
 * refused by ffgemm_grouped_plan_create with RTN_ERR, before it allocates.  A plan uploads the problem descriptors and
 * the tile schedule (every output tile dealt to a workgroup of the grid,
 * longest K first) once; this call allocates device memory and synchronises.
 * After it, ffgemm_grouped_plan_run only launches a kernel and can be
 * captured into a hipGraph.  `tile` picks the macro tile (NULL: the largest
 * with at least two tiles per CU); each tile runs the default engine for
 * it.  Pointer arrays are host arrays of device pointers. */
typedef struct ffgemm_grouped_plan_s *ffgemm_grouped_plan;
int ffgemm_grouped_plan_create(ffgemm_grouped_plan *plan, int count, const ffgemm_problem *probs,
                               const void *const *A, const void *const *B,
                               const void *const *C, void *const *D, const ffgemm_tile *tile);
int ffgemm_grouped_plan_run(ffgemm_grouped_plan plan, void *stream);
int ffgemm_grouped_plan_info(ffgemm_grouped_plan plan, int64_t *total_tiles, int32_t *grid,
                             ffgemm_tile *tile);
int ffgemm_grouped_plan_destroy(ffgemm_grouped_plan plan);

/* Number of compute units of the current device (for tile scheduling). */
int ffgemm_device_cu_count(void);

#ifdef __cplusplus
}
#endif
#endif
