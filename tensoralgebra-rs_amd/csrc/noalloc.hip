// An execution entry point that allocates nothing: TAPP_execute_product for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). It has no counterpart in
// the reference (whose CPU path leases its workspace from the Exec,
// crates/tprims-exec/src/workspace.rs).
//
// What TAPP_execute_product allocates on the device (the plan's scatter
// vectors on first use per device, the split-K partial tiles in the
// executor's arena, staging for host operands, split-BF16 / FP8 panels) comes
// here from one file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES
// per GPU: the plan's scatter vectors are copied to its start on the
// executor's stream (from the plan's own host image), the split-K partials
// follow. Operands must be device memory (no staging) and the precision the
// storage one (split-BF16 and FP8 panels do not fit a modest array).
//
// Host code never uses the array's name as an address (it is a host-side
// stand-in): the device address comes from hipGetSymbolAddress, once per
// device, with that device current, and only that address is passed to the
// copies and kernels. Nothing in the array must start at zero: the scatter
// vectors are rewritten by every call before its kernels run, and every split
// writes all of its in-range partial elements before the reduction reads
// them. Every call on one GPU shares that GPU's array: calls that overlap in
// time on one GPU (different streams or threads without ordering) race;
// calls on one stream are ordered by it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "capi_util.h"
#include "executor.h"
#include "kernels.h"

extern "C" {
#include "tprims/tprims.h"
}

namespace tp {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)TPRIMS_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;  // device -> the device address of `scratch` there

// The device address of `scratch` on `device` (which is current), looked up
// on the first call per device and kept.
char* scratch_on(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    fail(ERR_INTERNAL, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  g_scratch_ptr[device] = (char*)ptr;
  return (char*)ptr;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace
}  // namespace tp

using namespace tp;

extern "C" TAPP_error tprims_execute_product_noalloc(TAPP_tensor_product plan, TAPP_executor exec,
                                                     TAPP_status* status, const void* alpha, const void* A,
                                                     const void* B, const void* beta, const void* C, void* D) {
  return ffi([&] {
    if (status) *status = 0;
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    Plan& p = *(Plan*)plan;
    double al[2], be[2];
    read_scalar(alpha, p.prob.dtype, "alpha", al);
    read_scalar(beta, p.prob.dtype, "beta", be);
    validate_item(p, Item{A, B, C, D}, be[0] == 0.0 && be[1] == 0.0);
    if (p.strategy == 2) return;  // empty output: nothing to do
    if (!C) be[0] = be[1] = 0.0;
    const bool read_c = be[0] != 0.0 || be[1] != 0.0;
    const bool product = (p.strategy == 0 || p.strategy == 3) && (al[0] != 0.0 || al[1] != 0.0);
    if (product && p.cfg.precision != PREC_STORAGE)
      fail(ERR_UNSUPPORTED, "the no-allocation entry point runs the storage precision only (split-BF16 and FP8 "
                            "panels need TAPP_execute_product)");
    // Scratch layout: the scatter vectors, then the split-K partial tiles.
    const std::vector<int64_t>& tables = p.host_tables();
    const size_t table_bytes = tables.size() * sizeof(int64_t);
    const size_t work_bytes = product && p.strategy == 0 && p.ksplit > 1 ? splitk_work_bytes(p, p.ksplit, p.H) : 0;
    const size_t need = round256(table_bytes) + work_bytes;
    if (need > NOALLOC_SCRATCH_BYTES)
      fail(ERR_UNSUPPORTED, "this product needs " + std::to_string(need) + " bytes of device scratch (scatter "
                            "vectors " + std::to_string(table_bytes) + ", split-K partials " +
                            std::to_string(work_bytes) + "); the no-allocation entry point has " +
                            std::to_string(NOALLOC_SCRATCH_BYTES) + " (a plan with split_k = 1 needs no partials)");
    Executor* ex = resolve_executor(exec);
    DeviceGuard dg(ex->device);
    Where wa, wb, wc, wd = classify(D, ex->device);
    if (product) {
      wa = classify(A, ex->device);
      wb = classify(B, ex->device);
    }
    if (read_c) wc = classify(C, ex->device);
    if (wa.host || wb.host || wc.host || wd.host)
      fail(ERR_DEVICE, "the no-allocation entry point takes device memory only (host operands would need staging)");
    ex->inflight++;
    struct Release {
      Executor* e;
      ~Release() { e->inflight--; }
    } rel{ex};
    hipStream_t s = ex->get_stream();
    char* base = scratch_on(ex->device);
    hip_check(hipMemcpyAsync(base, tables.data(), table_bytes, hipMemcpyHostToDevice, s),
              "copying the scatter vectors to the scratch array");
    ContractArgs a =
        p.args_with(p.tables_at((int64_t*)base), wa.dev, wb.dev, read_c ? wc.dev : nullptr, const_cast<void*>(wd.dev),
                    al, be, read_c);
    launch_plan(p, *ex, a, product, p.ksplit, p.kslabs, s, work_bytes ? base + round256(table_bytes) : nullptr);
    ex->finish(s, !ex->async.load());
  });
}
