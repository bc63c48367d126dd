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
