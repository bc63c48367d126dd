// The GPU executor: the MI355X counterpart of crates/tprims-exec (Exec, the
// workspace provider) and of the executor handles of
// crates/tprims-capi/src/executor.rs. It owns a stream, a device scratch arena
// (staging of host operands) and the in-flight count that makes destruction
// refuse with TPRIMS_BUSY while calls run.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "plan.h"

namespace tp {

struct Executor {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = false;  // create our own stream lazily
  bool stream_ready = false;
  bool is_default = false;
  std::atomic<int> async{0};
  size_t pool_size = 0, budget = 1;  // recorded reference thread settings
  std::atomic<int> inflight{0};
  std::mutex mu;  // stream creation and scratch
  char* scratch = nullptr;
  size_t scratch_cap = 0;
  // Split-BF16 operand panels (reused call after call in stream order).
  std::mutex panel_mu;
  char* panels = nullptr;
  size_t panels_cap = 0;
  // Batch-entry offset vectors of fused batched calls: a device buffer reused
  // in stream order, filled from a pinned host buffer whose previous upload
  // (batch_ev) has completed before it is rewritten.
  std::mutex batch_mu;
  char* batch_vecs = nullptr;
  size_t batch_cap = 0;
  char* batch_host = nullptr;
  size_t batch_host_cap = 0;
  hipEvent_t batch_ev = nullptr;
  bool batch_ev_pending = false;

  hipStream_t get_stream();
  char* ensure_scratch(size_t bytes);  // caller holds mu
  char* ensure_panels(size_t bytes);   // caller holds panel_mu
  char* ensure_batch(size_t bytes);       // caller holds batch_mu
  char* ensure_batch_host(size_t bytes);  // caller holds batch_mu; waits for the previous upload
  // Ends a call's stream work: waits for it when sync.
  void finish(hipStream_t s, bool sync);
  // Work still queued or running on the executor's stream (an asynchronous
  // call's, or, on a caller's stream, anyone's): destruction refuses then.
  bool busy();
  ~Executor();
};

// Zero selects the default executor of the current device (null stream).
Executor* resolve_executor(intptr_t handle);
Executor* new_executor(int device, hipStream_t stream);

struct Item {
  const void* a;
  const void* b;
  const void* c;
  void* d;
};

// Byte addresses [lo, hi) an operand of a plan addresses from its origin;
// absent for a null origin or an empty layout.
struct Range {
  bool present = false;
  __int128 lo = 0, hi = 0;
};
Range byte_range(const Plan& p, int op, const void* origin);
bool overlap(const Range& x, const Range& y);

// The byte ranges one item of a batch or group writes (d) and reads (in:
// A and B when the product runs, C when beta != 0).
struct ItemSpans {
  Range d;
  Range in[3];
};
// True when running the items at once equals running them in order: no
// item's output meets another item's output or inputs (span hulls, so
// conservative). An item may read its own output (C is D).
bool items_independent(const std::vector<ItemSpans>& items);

// Where an operand lives, as seen from an executor's device: device (or
// managed, or pinned host) memory usable in place at `dev`, or pageable host
// memory (`host`), which run_item stages. Device memory of another device is
// TPRIMS_ERR_DEVICE.
struct Where {
  const void* dev = nullptr;
  bool host = false;
};
Where classify(const void* p, int device);

// The null rules of the C ABI (crates/tprims-capi/src/execute.rs:30-47) and the
// plan's alias preflight (crates/tprims-contract/src/plan/mod.rs:384-428).
void validate_item(const Plan& p, const Item& it, bool beta_zero);

// Runs one validated item (plan/mod.rs:487-530 semantics). sync: wait for
// the work before returning (staged host operands always wait).
void run_item(Plan& p, Executor& ex, const double alpha[2], const double beta[2], const Item& it, bool sync);

// Runs validated items of one plan as ONE launch when that gives the same
// result as running them in order: every operand in device memory, storage
// precision, element-aligned origins, and no item writing memory another item
// reads or writes. The items become extra batch entries (their origins folded
// into the batch offset vectors). Returns false, having done nothing, when the
// items must run one by one.
bool run_items_fused(Plan& p, Executor& ex, const double alpha[2], const double beta[2],
                     const std::vector<Item>& items);

// The launches of one planned call on device-resident arguments: output pass,
// elementwise strategy, FP8, split-BF16, split K (partials then the ordered
// reduction) or the contraction kernel. ksplit <= 1 runs K in one range.
// Split K writes its partial tiles to `work` (splitk_work_bytes) when given,
// else to the executor's panel arena (grown as needed).
void launch_plan(Plan& p, Executor& ex, ContractArgs& a, bool product, int ksplit, int64_t kslabs, hipStream_t s,
                 void* work = nullptr);
size_t splitk_work_bytes(const Plan& p, int ksplit, int64_t H);

void hip_check(hipError_t e, const char* what);

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int d);
  ~DeviceGuard();
};

}  // namespace tp
