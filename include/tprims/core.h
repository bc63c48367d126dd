/* libtprims core extensions: status codes, last error, library identity and
 * borrowed DLPack operands. Mirrors crates/tprims-capi/include/tprims/core.h
 * (status table crates/tprims-capi/src/status.rs:11-42, identity
 * crates/tprims-capi/src/handle.rs:144,162, version/parts
 * crates/tprims-capi/src/lib.rs:137,150, borrow crates/tprims-capi/src/tensor.rs:26,55).
 *
 * MI355X build: operands may live in device memory (kDLROCM) as well as in
 * host memory (kDLCPU, kDLROCMHost). */
#ifndef TPRIMS_CORE_H
#define TPRIMS_CORE_H

#include <stddef.h>
#include <stdint.h>

#include "tapp.h"
#include "dlpack/dlpack.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TPRIMS_ABI_VERSION 100 /* 0.1.0 */

typedef int32_t tprims_status;
#define TPRIMS_OK 0
#define TPRIMS_ERR_INVALID_ARGUMENT 1
#define TPRIMS_ERR_SHAPE 2
#define TPRIMS_ERR_DTYPE 3
#define TPRIMS_ERR_DEVICE 4
#define TPRIMS_ERR_READ_ONLY 5
#define TPRIMS_ERR_ALIASED 6
#define TPRIMS_ERR_WOULD_MATERIALIZE 7
#define TPRIMS_ERR_NUMERICAL 8
#define TPRIMS_BUSY 9
#define TPRIMS_ERR_WOULD_DEADLOCK 10
/* 11 retired */
#define TPRIMS_ERR_PANIC 12 /* a C++ exception was caught at the ABI boundary */
#define TPRIMS_ERR_INTERNAL 13
#define TPRIMS_ERR_LABELS 14
#define TPRIMS_ERR_UNSUPPORTED 15

/* A borrowed operand, valid for the duration of a call; never freed here. */
typedef struct {
  DLTensor* view;
  uint64_t flags;
} tprims_tensor;

TAPP_EXPORT tprims_tensor tprims_tensor_borrow_versioned(struct DLManagedTensorVersioned* m);
TAPP_EXPORT tprims_tensor tprims_tensor_borrow_raw(DLTensor* t, uint64_t flags);

/* The operand contract of a borrowed tensor, as every DLPack entry point
 * applies it. Rust API counterparts: tprims::tensor::dtype_of
 * (crates/tprims-capi/src/tensor.rs:96) and tprims::tensor::layout (:145)
 * plus view_mut's read-only check (:281-297).
 * dtype: a null view is TPRIMS_ERR_INVALID_ARGUMENT; lanes != 1 or a type
 * other than float32/64, complex64/128 TPRIMS_ERR_DTYPE.
 * layout, in this order: null view (INVALID_ARGUMENT); a device other than
 * kDLCPU, kDLCUDAHost, kDLROCM, kDLROCMHost (DEVICE); dtype as above or
 * different from `dtype` (DTYPE); negative ndim (SHAPE); null shape with
 * ndim > 0 (INVALID_ARGUMENT); a negative extent (SHAPE); NULL strides are
 * compact row-major (overflow: SHAPE); a span that overflows int64 elements or
 * bytes (SHAPE); null data for a non-empty tensor (INVALID_ARGUMENT);
 * data + byte_offset not aligned to the real type (4 bytes f32/c32, 8 f64/c64:
 * INVALID_ARGUMENT); writable != 0 and the read-only flag (READ_ONLY). An
 * empty tensor's data is never read: origin is then a placeholder address.
 * dims/strides receive min(ndim, cap) entries. */
typedef struct {
  int32_t ndim;
  int32_t empty;   /* some extent is zero */
  void* origin;    /* data + byte_offset */
  int64_t span_lo; /* smallest and largest element offset addressed, from origin */
  int64_t span_hi;
} tprims_tensor_layout_info;
TAPP_EXPORT TAPP_error tprims_tensor_dtype(tprims_tensor t, TAPP_datatype* out);
TAPP_EXPORT TAPP_error tprims_tensor_layout(tprims_tensor t, TAPP_datatype dtype, int writable, int64_t* dims,
                                            int64_t* strides, int32_t cap, tprims_tensor_layout_info* out);

/* Message of the last failed call on this thread; "" after a success. */
TAPP_EXPORT const char* tprims_last_error(void);
TAPP_EXPORT uint32_t tprims_abi_version(void);
/* 1 for "core", "tapp" and "hip"; 0 otherwise. */
TAPP_EXPORT int32_t tprims_has_part(const char* part);

TAPP_EXPORT const char* TAPP_implementation_name(void);
TAPP_EXPORT const char* TAPP_implementation_version(void);

#ifdef __cplusplus
}
#endif
#endif
