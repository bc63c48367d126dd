/* TAPP (Tensor Algebra Processing Primitives) C interface, as exported by the
 * MI355X build of libtprims.
 *
 * Every prototype below is the standard TAPP prototype, re-declared from the
 * published interface so that a C consumer written against upstream
 * `<tapp.h>` compiles and links against this library unchanged. The
 * replaced reference interface is crates/tprims-capi/include/tapp.h and
 * crates/tprims-capi/include/tapp/{util,error,attributes,datatype,handle,
 * executor,status,tensor,product}.h; each group cites the Rust definition it
 * mirrors.
 *
 * Data pointers passed to TAPP_execute_product may be device (HIP) or host
 * memory: device memory is used in place, host memory is staged through the
 * executor's device scratch. All arithmetic runs on the GPU.
 */
#ifndef TAPP_TAPP_H_
#define TAPP_TAPP_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define TAPP_EXPORT extern "C" __attribute__((visibility("default")))
extern "C" {
#else
#define TAPP_EXPORT __attribute__((visibility("default")))
#endif

/* ---- errors (crates/tprims-capi/src/status.rs:69,81) ---------------------- */
typedef int TAPP_error;
/* True only for 0. */
TAPP_EXPORT bool TAPP_check_success(TAPP_error error);
/* snprintf-like: writes at most maxlen-1 characters plus a NUL, returns the
 * untruncated length. */
TAPP_EXPORT size_t TAPP_explain_error(TAPP_error error, size_t maxlen, char* message);

/* ---- attributes (crates/tprims-capi/src/handle.rs:84,101,121) -------------- */
typedef intptr_t TAPP_attr;
typedef int TAPP_key;
/* No keys are defined: all three refuse with TPRIMS_ERR_UNSUPPORTED. */
TAPP_EXPORT TAPP_error TAPP_attr_set(TAPP_attr attr, TAPP_key key, void* value);
TAPP_EXPORT TAPP_error TAPP_attr_get(TAPP_attr attr, TAPP_key key, void** value);
TAPP_EXPORT TAPP_error TAPP_attr_clear(TAPP_attr attr, TAPP_key key);

/* ---- datatypes and precisions (crates/tprims-capi/src/abi.rs:16-50) ------- */
typedef int TAPP_datatype;
enum {
  TAPP_F32 = 0,
  TAPP_F64 = 1,
  TAPP_C32 = 2, /* interleaved float re, im */
  TAPP_C64 = 3, /* interleaved double re, im */
  TAPP_F16 = 4, /* rejected with TPRIMS_ERR_DTYPE */
  TAPP_BF16 = 5, /* rejected with TPRIMS_ERR_DTYPE */
  TAPP_FLOAT = TAPP_F32,
  TAPP_DOUBLE = TAPP_F64,
  TAPP_SCOMPLEX = TAPP_C32,
  TAPP_DCOMPLEX = TAPP_C64
};

typedef int TAPP_prectype;
enum {
  TAPP_DEFAULT_PREC = -1,
  TAPP_F32F32_ACCUM_F32 = TAPP_F32,
  TAPP_F64F64_ACCUM_F64 = TAPP_F64,
  TAPP_F16F16_ACCUM_F16 = TAPP_F16,
  TAPP_F16F16_ACCUM_F32 = 5,
  /* For f32/c32 storage this library computes the products as split-BF16
   * (three bf16 terms per operand, six MFMA products, fp32 accumulation). */
  TAPP_BF16BF16_ACCUM_F32 = 6
};

/* ---- handles (crates/tprims-capi/src/handle.rs:35,50) -------------------- */
typedef intptr_t TAPP_handle;
TAPP_EXPORT TAPP_error TAPP_create_handle(TAPP_handle* handle);
TAPP_EXPORT TAPP_error TAPP_destroy_handle(TAPP_handle handle);

/* ---- executors (crates/tprims-capi/src/executor.rs:278,299) --------------- */
typedef intptr_t TAPP_executor;
/* A GPU executor on the current HIP device with a stream of its own.
 * Executor 0 is the library's default executor of the current device. */
TAPP_EXPORT TAPP_error TAPP_create_executor(TAPP_executor* exec);
TAPP_EXPORT TAPP_error TAPP_destroy_executor(TAPP_executor exec);

/* ---- status (crates/tprims-capi/src/handle.rs:134) ----------------------- */
typedef intptr_t TAPP_status;
TAPP_EXPORT TAPP_error TAPP_destroy_status(TAPP_status status);

/* ---- tensor infos (crates/tprims-capi/src/tensor_info.rs:58-221) ---------- */
typedef intptr_t TAPP_tensor_info;
TAPP_EXPORT TAPP_error TAPP_create_tensor_info(TAPP_tensor_info* info, TAPP_datatype type, int nmode,
                                               const int64_t* extents, const int64_t* strides);
TAPP_EXPORT TAPP_error TAPP_destroy_tensor_info(TAPP_tensor_info info);
TAPP_EXPORT int TAPP_get_nmodes(TAPP_tensor_info info);
TAPP_EXPORT TAPP_error TAPP_set_nmodes(TAPP_tensor_info info, int nmodes);
TAPP_EXPORT void TAPP_get_extents(TAPP_tensor_info info, int64_t* extents);
TAPP_EXPORT TAPP_error TAPP_set_extents(TAPP_tensor_info info, const int64_t* extents);
TAPP_EXPORT void TAPP_get_strides(TAPP_tensor_info info, int64_t* strides);
TAPP_EXPORT TAPP_error TAPP_set_strides(TAPP_tensor_info info, const int64_t* strides);

/* ---- products (crates/tprims-capi/src/product.rs:83,193;
 *                crates/tprims-capi/src/execute.rs:192,235) ---------------- */
typedef int TAPP_element_op;
enum { TAPP_IDENTITY = 0, TAPP_CONJUGATE = 1 };
#define TAPP_IN_PLACE NULL

typedef intptr_t TAPP_tensor_product;
TAPP_EXPORT TAPP_error TAPP_create_tensor_product(TAPP_tensor_product* plan, TAPP_handle handle,
                                                  TAPP_element_op op_A, TAPP_tensor_info A, const int64_t* idx_A,
                                                  TAPP_element_op op_B, TAPP_tensor_info B, const int64_t* idx_B,
                                                  TAPP_element_op op_C, TAPP_tensor_info C, const int64_t* idx_C,
                                                  TAPP_element_op op_D, TAPP_tensor_info D, const int64_t* idx_D,
                                                  TAPP_prectype prec);
TAPP_EXPORT TAPP_error TAPP_destroy_tensor_product(TAPP_tensor_product plan);
TAPP_EXPORT TAPP_error TAPP_execute_product(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status,
                                            const void* alpha, const void* A, const void* B, const void* beta,
                                            const void* C, void* D);
TAPP_EXPORT TAPP_error TAPP_execute_batched_product(TAPP_tensor_product plan, TAPP_executor exec,
                                                    TAPP_status* status, int num_batches, const void* alpha,
                                                    const void** A, const void** B, const void* beta,
                                                    const void** C, void** D);

#ifdef __cplusplus
}
#endif
#endif /* TAPP_TAPP_H_ */
