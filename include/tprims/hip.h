/* MI355X extensions of libtprims: HIP executors, plan configuration
 * (complex method, split-BF16 / FP8 precision, tile override), plan reports,
 * grouped execution of many products in one persistent kernel, and the
 * vector-Jacobian product of a contraction.
 *
 * None of these has a C counterpart in the reference except where cited; the
 * cited item is the Rust interface the entry point stands in for. Streams
 * are passed as `void*` (a hipStream_t) so that a C consumer needs no HIP
 * headers. */
#ifndef TPRIMS_HIP_H
#define TPRIMS_HIP_H
#include <stddef.h>
#include <stdint.h>
#include "tprims/core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- executors ------------------------------------------------------------
 * Counterpart of crates/tprims-exec/src/exec.rs:41 (Exec) and the executor
 * handles of crates/tprims-capi/src/executor.rs:278: an executor owns the
 * device scratch arena and the stream the contraction kernels run on. */

/* An executor on `device` that runs on `stream` (a hipStream_t, borrowed, not
 * destroyed), or on a stream of its own when `stream` is NULL. */
TAPP_EXPORT TAPP_error tprims_tapp_executor_create_hip(TAPP_executor* out, int device, void* stream);
/* async != 0: TAPP_execute_product returns once the work is enqueued, when every
 * operand is device memory (host operands always synchronize). Default 0. */
TAPP_EXPORT TAPP_error tprims_tapp_executor_set_async(TAPP_executor exec, int async);
TAPP_EXPORT TAPP_error tprims_tapp_executor_synchronize(TAPP_executor exec);
TAPP_EXPORT TAPP_error tprims_tapp_executor_get_stream(TAPP_executor exec, void** stream);
TAPP_EXPORT TAPP_error tprims_tapp_executor_get_device(TAPP_executor exec, int* device);

/* ---- plan configuration (crates/tprims-contract/src/plan/config.rs:101) ---- */
enum {
  TPRIMS_CPLX_NATIVE = 0, /* four real MFMA products per complex block */
  TPRIMS_CPLX_1M = 1,     /* one real product over 1e/1r-expanded operands */
  TPRIMS_CPLX_3M = 2      /* three products, Karatsuba recombination */
};
enum {
  TPRIMS_PREC_STORAGE = 0,      /* f32 MFMA for f32/c32, f64 MFMA for f64/c64 */
  TPRIMS_PREC_SPLIT_BF16_2 = 1, /* f32/c32: hi*hi + hi*lo + lo*hi on bf16 MFMA */
  TPRIMS_PREC_SPLIT_BF16_3 = 2, /* f32/c32: six bf16 products, ~fp32 accuracy */
  TPRIMS_PREC_FP8_E4M3 = 3      /* f32/c32: e4m3 operands with per-panel scales */
};
typedef struct {
  int32_t complex_method; /* TPRIMS_CPLX_* */
  int32_t precision;      /* TPRIMS_PREC_* */
  int32_t tile;           /* 0: chosen by the planner's tuner; else a tile id (see report): storage
                             1 64x64, 2 32x32, 3 128x64, 4 128x128 (LDS-DMA), 5 32x32 on four K groups
                             (LDS-DMA; runs as 2 where its gathers do not fit), 6 32x32 and 7 32x48 on
                             eight K groups loading into registers (f64 in 16-byte runs; 7 needs B
                             gathered along K, else runs as 6; both run as 5 otherwise); split-BF16 11 128x128,
                             12 64x64, 13 256x128, 14 128x64; FP8 14 128x64 (c32), 15 256x256 (f32) */
  int32_t split_k;        /* 0: chosen by the planner; 1: never split K; n > 1: n K ranges
                             (storage precision; partial tiles reduced in split order) */
  int32_t reserved[4];
} tprims_plan_config;

/* TAPP_create_tensor_product with a configuration (NULL = defaults). A prec of
 * TAPP_BF16BF16_ACCUM_F32 selects TPRIMS_PREC_SPLIT_BF16_3 unless cfg says
 * otherwise. */
TAPP_EXPORT TAPP_error tprims_create_tensor_product_ex(TAPP_tensor_product* plan, TAPP_handle handle,
                                                       TAPP_element_op op_A, TAPP_tensor_info A, const int64_t* idx_A,
                                                       TAPP_element_op op_B, TAPP_tensor_info B, const int64_t* idx_B,
                                                       TAPP_element_op op_C, TAPP_tensor_info C, const int64_t* idx_C,
                                                       TAPP_element_op op_D, TAPP_tensor_info D, const int64_t* idx_D,
                                                       TAPP_prectype prec, const tprims_plan_config* cfg);

/* Execute with borrowed DLPack operands. Each view passes the operand
 * contract of tprims_tensor_layout (core.h; crates/tprims-capi/src/tensor.rs:145-230,
 * extended to ROCm memory: a kDLROCM view must be on the executor's device,
 * else TPRIMS_ERR_DEVICE), and its shape and strides must equal the extents
 * and strides of the info the plan was built from (strides of extent-1 modes
 * are free; else TPRIMS_ERR_SHAPE). A D flagged read-only is
 * TPRIMS_ERR_READ_ONLY. C.view may be NULL only with a zero beta. */
TAPP_EXPORT TAPP_error tprims_execute_product_dlpack(TAPP_tensor_product plan, TAPP_executor exec,
                                                     const void* alpha, tprims_tensor A, tprims_tensor B,
                                                     const void* beta, tprims_tensor C, tprims_tensor D);

/* ---- execution without allocation -------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results as
 * TAPP_execute_product, and it calls no hipMalloc, hipFree or allocating
 * library routine: the plan's scatter vectors and the split-K partial tiles
 * go to one static device array of TPRIMS_NOALLOC_SCRATCH_BYTES per GPU
 * (held from the library's first use of a GPU until the process ends), the
 * scatter vectors copied there on the executor's stream by every call.
 * Limits: operands in device memory of the executor's device (else
 * TPRIMS_ERR_DEVICE), the storage precision (else TPRIMS_ERR_UNSUPPORTED), and
 * scatter vectors plus split-K partials within the array (else
 * TPRIMS_ERR_UNSUPPORTED; a plan with split_k = 1 needs no partials). All
 * calls on one GPU share its array: calls that overlap in time on one GPU
 * (different streams or threads) race; calls on one stream are ordered. */
#define TPRIMS_NOALLOC_SCRATCH_BYTES (64u << 20)
TAPP_EXPORT TAPP_error tprims_execute_product_noalloc(TAPP_tensor_product plan, TAPP_executor exec,
                                                      TAPP_status* status, const void* alpha, const void* A,
                                                      const void* B, const void* beta, const void* C, void* D);

/* What the planner concluded (crates/tprims-contract/src/plan/report.rs:33). */
typedef struct {
  int64_t m, n, k, batch;   /* reduced matrix view after folding */
  int64_t flops;            /* 2 (real) or 8 (complex) per multiply-add */
  int32_t dtype;            /* TAPP_F32 .. TAPP_C64 */
  int32_t complex_method;
  int32_t precision;
  int32_t tile;             /* tile id used */
  int32_t tile_m, tile_n, tile_k;
  int32_t a_k_major, b_k_major; /* operand gathers run along K (1) or along M/N (0) */
  int32_t d_m_major;        /* write-back runs along M (1) or N (0) */
  int32_t m_axes, n_axes, k_axes, h_axes; /* folded axis counts */
  int32_t strategy;         /* 0 contraction kernel, 1 output pass only, 2 empty, 3 elementwise */
  int32_t source;           /* tile and K split from: 0 the planner's cost model, 1 the configuration,
                               2 the tuning cache, 3 tprims_product_autotune on this plan */
  char kernel[64];
} tprims_plan_report;
TAPP_EXPORT TAPP_error tprims_product_report(TAPP_tensor_product plan, tprims_plan_report* out);

/* ---- autotuning and the tuning cache ---------------------------------------
 * The reference resolves a kernel family and blocking once per plan and
 * freezes it there (crates/tprims-contract/src/select/resolve.rs:164). Here a
 * plan takes its tile and K split from the planner's cost model, or from the
 * process's tuning cache when that holds an entry for the plan's reduced
 * problem (dtype, complex method, precision and every folded axis with its
 * strides) and the plan asked for neither a tile nor a split.
 *
 * tprims_product_autotune times every tile (and, at the storage precision,
 * K split or none) the plan can run, `reps` executions each after one
 * warm-up (HIP events on the executor's stream), on the caller's operands
 * with the arguments of TAPP_execute_product, freezes the fastest into the
 * plan and records it in the tuning cache. On return D holds the result of
 * one execution on the chosen configuration: an output that is also read
 * (C overlapping D with beta != 0) is saved before the timing runs and
 * restored first. Operands must be device memory of the executor's device
 * (else TPRIMS_ERR_DEVICE). It changes the plan: no other call may use the
 * plan meanwhile. best_* (each may be NULL) receive the tile id, the number
 * of K ranges and the chosen configuration's mean time in ms.
 *
 * The cache is kept in the process; tprims_tuning_cache_save writes it to
 * `path` (replacing the file as a whole), tprims_tuning_cache_load merges a
 * saved file into it (all entries or, on any malformed line, none:
 * TPRIMS_ERR_INVALID_ARGUMENT with the line number in the last error). The
 * file is text: a header line, then one "tile split_k key" line per entry.
 * Entries apply to plans created after they are stored or loaded. Nothing is
 * read from the environment. */
TAPP_EXPORT TAPP_error tprims_product_autotune(TAPP_tensor_product plan, TAPP_executor exec, const void* alpha,
                                               const void* A, const void* B, const void* beta, const void* C,
                                               void* D, int32_t reps, int32_t* best_tile, int32_t* best_split,
                                               double* best_ms);
/* The plan's tuning-cache key: writes min(length, cap - 1) bytes plus a
 * terminating zero to out (null allowed when cap is 0) and the length to *len. */
TAPP_EXPORT TAPP_error tprims_product_tuning_key(TAPP_tensor_product plan, char* out, int64_t cap, int64_t* len);
TAPP_EXPORT TAPP_error tprims_tuning_cache_save(const char* path);
TAPP_EXPORT TAPP_error tprims_tuning_cache_load(const char* path);
TAPP_EXPORT TAPP_error tprims_tuning_cache_clear(void);
TAPP_EXPORT TAPP_error tprims_tuning_cache_size(int64_t* count);

/* ---- scatter vectors ----------------------------------------------------------
 * The reference exposes these as Rust API: build_scatter and
 * build_block_scatter (crates/tprims-kernel/src/pack/scatter.rs:70,130) and
 * the plan's reduced matrix view (crates/tprims-contract/src/plan/analysis.rs:210,
 * PackedPlan; pack/scatter.rs:24-31 names them the way for a caller to see the
 * traversal the engine makes). Each call writes min(length, cap) entries to
 * out (null allowed when cap is 0) and the full length to *len, so a first call
 * with cap 0 sizes the buffer. */
#define TPRIMS_SCATTER_IRREGULAR (-0x7fffffffffffffffLL - 1) /* block not an arithmetic run */
/* Offsets of the mixed-radix index tuples, fastest mode first; nmode 0 gives
 * the single offset 0, a zero extent the empty vector. */
TAPP_EXPORT TAPP_error tprims_build_scatter(const int64_t* extents, const int64_t* strides, int32_t nmode,
                                            int64_t* out, int64_t cap, int64_t* len);
/* Per block of blk entries of scat (the last block ragged): the common
 * difference, 0 for a one-entry block, or TPRIMS_SCATTER_IRREGULAR. blk must
 * be positive (the reference panics on 0). */
TAPP_EXPORT TAPP_error tprims_build_block_scatter(const int64_t* scat, int64_t n, int64_t blk, int64_t* out,
                                                  int64_t cap, int64_t* nblocks);
/* The plan's scatter vectors in elements, after role assignment, sorting and
 * folding (the K order is the planner's gather order). An empty role holds the
 * single offset 0. */
enum {
  TPRIMS_SCATTER_AM = 0, TPRIMS_SCATTER_AK, TPRIMS_SCATTER_BK, TPRIMS_SCATTER_BN,
  TPRIMS_SCATTER_CM, TPRIMS_SCATTER_CN, TPRIMS_SCATTER_DM, TPRIMS_SCATTER_DN,
  TPRIMS_SCATTER_HA, TPRIMS_SCATTER_HB, TPRIMS_SCATTER_HC, TPRIMS_SCATTER_HD
};
TAPP_EXPORT TAPP_error tprims_product_scatter(TAPP_tensor_product plan, int32_t which, int64_t* out, int64_t cap,
                                              int64_t* len);

/* ---- grouped execution (crates/tprims-contract/src/batch.rs:125) -----------
 * Many products of one dtype (each its own plan and shape) bound to their
 * operands and run as one persistent kernel over a device work list. alpha and
 * beta arrays hold one scalar per item (of the dtype). All pointers are device
 * memory. Items run at the same time, so no item may write bytes another
 * item reads or writes (span hulls; else TPRIMS_ERR_ALIASED); an item may
 * read its own output (C is D). The group owns its work list and holds a
 * reference to each plan (a plan destroyed with TAPP_destroy_tensor_product
 * stays alive until the group is destroyed); use_graph captures the launch
 * into a hipGraph that execute then replays. (TAPP_execute_batched_product, for
 * items of one plan, also runs as one launch when every operand is device
 * memory, the precision is the storage one and no item writes bytes another
 * item reads or writes; otherwise its items run in order.) */
typedef intptr_t tprims_group;
TAPP_EXPORT TAPP_error tprims_group_create(tprims_group* out, TAPP_executor exec, int num_items,
                                           const TAPP_tensor_product* plans, const void* alpha,
                                           const void* const* A, const void* const* B, const void* beta,
                                           const void* const* C, void* const* D);
TAPP_EXPORT TAPP_error tprims_group_execute(tprims_group group, TAPP_executor exec);
TAPP_EXPORT TAPP_error tprims_group_use_graph(tprims_group group, TAPP_executor exec, int enable);
/* Launch every item separately (one kernel per item), for comparison. */
TAPP_EXPORT TAPP_error tprims_group_execute_each(tprims_group group, TAPP_executor exec);
TAPP_EXPORT TAPP_error tprims_group_info(tprims_group group, int64_t* tiles, int64_t* flops, int32_t* tile_id);
/* Fix the tile shape of the persistent launch (a storage-precision tile id). */
TAPP_EXPORT TAPP_error tprims_group_set_tile(tprims_group group, int32_t tile_id);
/* The autotuner: times every storage tile shape on the group's own operands
 * (reps launches each, HIP events on the executor's stream), keeps the fastest
 * and reports it. On return every output holds the result of one
 * tprims_group_execute: outputs that accumulate into themselves (C is D,
 * beta != 0) are saved before the timing launches and restored before one
 * final execution on the chosen tile. */
TAPP_EXPORT TAPP_error tprims_group_autotune(tprims_group group, TAPP_executor exec, int reps, int32_t* best_tile,
                                             double* best_ms);
TAPP_EXPORT TAPP_error tprims_group_destroy(tprims_group group);

/* ---- differentiation --------------------------------------------------------
 * For D = op_D(alpha * op_A(A) * op_B(B) + beta * op_C(C)) and a cotangent dD
 * (in D's layout), writes the cotangents of A and B in A's and B's layouts,
 * in the convention of conjugate Wirtinger derivatives (PyTorch's):
 *   dA = op_A(conj(alpha) * op_D(dD) * conj(op_B(B)))  summed over the labels not in A,
 * and symmetrically dB. A repeated label of A writes only A's diagonal (the
 * other addressed elements are zeroed); a label of A absent from B and D
 * receives the same value at every index (broadcast). accumulate != 0 adds into
 * dA/dB instead of overwriting. dA or dB may be NULL to skip it. */
TAPP_EXPORT TAPP_error tprims_product_vjp(TAPP_tensor_product plan, TAPP_executor exec, const void* alpha,
                                          const void* A, const void* B, const void* dD, void* dA, void* dB,
                                          int accumulate);
/* The plan of the adjoint contraction (which = 0: dA, 1: dB). */
TAPP_EXPORT TAPP_error tprims_product_vjp_report(TAPP_tensor_product plan, int which, tprims_plan_report* out);

#ifdef __cplusplus
}
#endif
#endif
