/* Executor extensions of the reference (crates/tprims-capi/include/tprims/tapp_ext.h,
 * crates/tprims-capi/src/executor.rs:318,358,392), kept for source
 * compatibility. The MI355X executor computes on the GPU: "threads" are the
 * host-side budget a caller asked for and are recorded and reported, not
 * used to partition work. */
#ifndef TPRIMS_TAPP_EXT_H
#define TPRIMS_TAPP_EXT_H
#include <stddef.h>
#include "tprims/core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  size_t stack_size;
} tprims_rayon_opts;

/* nthreads == 0 is TPRIMS_ERR_INVALID_ARGUMENT. Creates a GPU executor on the
 * current device; get_threads reports (nthreads, budget) for nthreads >= 2
 * and (0, 1) for nthreads == 1, as the reference's serial executor does. */
TAPP_EXPORT TAPP_error tprims_tapp_executor_create_rayon(TAPP_executor* out, size_t nthreads,
                                                         const tprims_rayon_opts* opts);
TAPP_EXPORT TAPP_error tprims_tapp_executor_set_budget(TAPP_executor exec, size_t budget);
TAPP_EXPORT TAPP_error tprims_tapp_executor_get_threads(TAPP_executor exec, size_t* pool_size, size_t* budget);

#ifdef __cplusplus
}
#endif
#endif
