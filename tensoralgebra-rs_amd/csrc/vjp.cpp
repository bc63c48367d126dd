// Vector-Jacobian products of a contraction. Not in the reference (automatic
// differentiation is out of its scope, docs/design-principles.md:85); each
// cotangent is itself a contraction, lowered by the same lowering
// (crates/tprims-contract/src/api/validate.rs:207) with two extensions the
// reference refuses for products: an output-only label (a broadcast over an
// index A reduced) and a repeated output label writing a diagonal.
//
// For D = op_D(alpha * op_A(A) * op_B(B) + beta * op_C(C)) and cotangent g in
// D's layout, with conjugate-Wirtinger derivatives:
//   dA = op_A( conj(alpha) * op_D(g) * conj(op_B(B)) )   over the labels of A
//   dB = op_B( conj(alpha) * conj(op_A(A)) * op_D(g) )   over the labels of B
// A repeated label of A makes dA live on A's diagonal: the other addressed
// elements are zeroed first (or left as they are when accumulating, which is
// the scatter-add into the diagonal). A label of A that is in neither B nor D
// (a reduction) receives the same value at each of its indices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "capi_util.h"
#include "executor.h"
#include "kernels.h"

extern "C" {
#include "tprims/tprims.h"
}

using namespace tp;

namespace tp {
void fill_report(const Plan* p, tprims_plan_report* out);
}

namespace {

bool has_repeats(const std::vector<int64_t>& l) {
  for (size_t i = 0; i < l.size(); i++)
    for (size_t j = i + 1; j < l.size(); j++)
      if (l[i] == l[j]) return true;
  return false;
}

// Builds (once) the adjoint plan of operand `which` (0: A, 1: B).
Plan* adjoint(Plan* fwd, int which) {
  std::lock_guard<std::mutex> g(fwd->mu);
  if (fwd->adj[which]) return fwd->adj[which];
  const Problem& p = fwd->prob;
  Operand gD{p.d.layout, p.d.conj};  // op_D(g)
  Problem q;
  if (which == 0) {
    Operand y{p.b.layout, !p.b.conj};  // conj(op_B(B))
    Operand out{p.a.layout, p.a.conj};
    q = lower(p.dtype, gD, p.ld, y, p.lb, &out, &p.la, out, p.la, /*allow_broadcast=*/true);
  } else {
    Operand x{p.a.layout, !p.a.conj};  // conj(op_A(A))
    Operand out{p.b.layout, p.b.conj};
    q = lower(p.dtype, x, p.la, gD, p.ld, &out, &p.lb, out, p.lb, /*allow_broadcast=*/true);
  }
  PlanConfig cfg = fwd->cfg;
  cfg.tile = 0;
  Plan* a = make_plan(q, cfg);
  fwd->adj[which] = a;
  fwd->adj_zero_fill[which] = has_repeats(which == 0 ? p.la : p.lb);
  return a;
}


bool on_device(const void* p, int device) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (at.type == hipMemoryTypeDevice && at.device == device) || at.type == hipMemoryTypeManaged;
}

}  // namespace

extern "C" {

TAPP_error tprims_product_vjp(TAPP_tensor_product plan, TAPP_executor exec, const void* alpha, const void* A,
                              const void* B, const void* dD, void* dA, void* dB, int accumulate) {
  return ffi([&] {
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    Plan* fwd = (Plan*)plan;
    if (!dD) fail(ERR_INVALID_ARGUMENT, "the cotangent dD is null");
    if (dA && !B) fail(ERR_INVALID_ARGUMENT, "B is needed for dA");
    if (dB && !A) fail(ERR_INVALID_ARGUMENT, "A is needed for dB");
    double al[2];
    read_scalar(alpha, fwd->prob.dtype, "alpha", al);
    const double alc[2] = {al[0], -al[1]};  // conj(alpha)
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    Plan* adjA = dA ? adjoint(fwd, 0) : nullptr;
    Plan* adjB = dB ? adjoint(fwd, 1) : nullptr;
    // Validate both items before writing either.
    if (adjA) validate_item(*adjA, Item{dD, B, accumulate ? dA : nullptr, dA}, !accumulate);
    if (adjB) validate_item(*adjB, Item{A, dD, accumulate ? dB : nullptr, dB}, !accumulate);
    if (dA && dB && dA == dB) fail(ERR_ALIASED, "dA and dB are the same buffer");
    Executor* ex = resolve_executor(exec);
    for (const void* q : {A, B, dD, (const void*)dA, (const void*)dB})
      if (q && !on_device(q, ex->device)) fail(ERR_DEVICE, "VJP operands must be device memory of the executor");
    ex->inflight++;
    struct Release {
      Executor* e;
      ~Release() { e->inflight--; }
    } rel{ex};
    DeviceGuard dg(ex->device);
    hipStream_t s = ex->get_stream();
    // The adjoint launches are queued back to back; the call waits once at
    // the end (a synchronous executor) or records its completion.
    const bool sync = !ex->async.load();
    for (int w = 0; w < 2; w++) {
      Plan* adj = w == 0 ? adjA : adjB;
      if (!adj) continue;
      void* out = w == 0 ? dA : dB;
      const Operand& own = w == 0 ? fwd->prob.a : fwd->prob.b;
      if (!accumulate && fwd->adj_zero_fill[w]) {
        if (own.layout.rank() > 8) fail(ERR_UNSUPPORTED, "zero fill supports rank <= 8");
        hip_check(launch_zero_layout(out, fwd->prob.dtype, own.layout.rank(), own.layout.dims.data(),
                                     own.layout.strides.data(), s),
                  "zero-fill launch");
      }
      const void* x = w == 0 ? dD : A;
      const void* y = w == 0 ? B : dD;
      run_item(*adj, *ex, alc, accumulate ? one : zero, Item{x, y, accumulate ? out : nullptr, out}, false);
    }
    ex->finish(s, sync);
  });
}

// The adjoint plan's report (which = 0: dA, 1: dB), for inspection.
TAPP_error tprims_product_vjp_report(TAPP_tensor_product plan, int which, tprims_plan_report* out) {
  return ffi([&] {
    if (plan == 0 || !out || which < 0 || which > 1) fail(ERR_INVALID_ARGUMENT, "bad plan, report or operand");
    fill_report(adjoint((Plan*)plan, which), out);
  });
}

}  // extern "C"
