// Grouped execution and differentiation entry points (placeholders).
#include "capi_util.h"
#include "executor.h"
extern "C" {
#include "tprims/tprims.h"
}
using namespace tp;
extern "C" {
TAPP_error tprims_group_create(tprims_group*, TAPP_executor, int, const TAPP_tensor_product*, const void*,
                               const void* const*, const void* const*, const void*, const void* const*,
                               void* const*) {
  return ffi([] { fail(ERR_UNSUPPORTED, "grouped execution is not available yet"); });
}
TAPP_error tprims_group_execute(tprims_group, TAPP_executor) { return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); }); }
TAPP_error tprims_group_use_graph(tprims_group, TAPP_executor, int) { return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); }); }
TAPP_error tprims_group_execute_each(tprims_group, TAPP_executor) { return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); }); }
TAPP_error tprims_group_info(tprims_group, int64_t*, int64_t*, int32_t*) { return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); }); }
TAPP_error tprims_group_destroy(tprims_group) { return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); }); }
TAPP_error tprims_product_vjp(TAPP_tensor_product, TAPP_executor, const void*, const void*, const void*, const void*,
                              void*, void*, int) {
  return ffi([] { fail(ERR_UNSUPPORTED, "n/a"); });
}
}
