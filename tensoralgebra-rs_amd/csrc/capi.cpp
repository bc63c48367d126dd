// The C ABI of libtprims (MI355X build): the standard TAPP entry points and
// the tprims extensions. Restates the behaviour of crates/tprims-capi/src
// (handle.rs, tensor_info.rs, product.rs, execute.rs, executor.rs, status.rs,
// tensor.rs, lib.rs): the same null rules, status codes, validation order and
// last-error contract, with the contraction running on the GPU.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_util.h"
#include "executor.h"
#include "kernels.h"
#include "plan.h"

extern "C" {
#include "tprims/tprims.h"
}

using namespace tp;

namespace tp {

static thread_local std::string g_last_error;

void set_last_error(const std::string& m) { g_last_error = m; }
void clear_last_error() {
  if (!g_last_error.empty()) g_last_error.clear();
}

// Reads alpha/beta as the plan's element type.
void read_scalar(const void* p, DType t, const char* name, double out[2]) {
  if (!p) fail(ERR_INVALID_ARGUMENT, std::string(name) + " is null or zero");
  out[0] = out[1] = 0.0;
  switch (t) {
    case DType::F32: {
      float v;
      memcpy(&v, p, 4);
      out[0] = v;
      break;
    }
    case DType::F64: memcpy(out, p, 8); break;
    case DType::C32: {
      float v[2];
      memcpy(v, p, 8);
      out[0] = v[0];
      out[1] = v[1];
      break;
    }
    case DType::C64: memcpy(out, p, 16); break;
  }
}

}  // namespace tp

namespace {

struct HandleState {
  uint64_t reserved = 0;
};

struct TensorInfo {
  int dtype;
  std::vector<int64_t> extents, strides;
};

DType dtype_of(int t) {
  switch (t) {
    case TAPP_F32: return DType::F32;
    case TAPP_F64: return DType::F64;
    case TAPP_C32: return DType::C32;
    case TAPP_C64: return DType::C64;
  }
  fail(ERR_DTYPE, "unsupported datatype");
}

bool op_of(int op, const char* name) {
  if (op == TAPP_IDENTITY) return false;
  if (op == TAPP_CONJUGATE) return true;
  fail(ERR_UNSUPPORTED, std::string("element operation ") + std::to_string(op) + " on " + name + " is not supported");
}

Operand operand_of(const TensorInfo& t, bool conj) {
  Operand o;
  o.layout.dims = t.extents;
  o.layout.strides = t.strides;
  o.layout.offset = 0;
  o.conj = conj;
  return o;
}

std::vector<int64_t> labels_of(const int64_t* p, const TensorInfo& t, const char* name) {
  size_t n = t.extents.size();
  if (n == 0) return {};
  if (!p) fail(ERR_INVALID_ARGUMENT, std::string("labels of ") + name + " is null or zero");
  return std::vector<int64_t>(p, p + n);
}

int create_product(TAPP_tensor_product* plan_out, TAPP_handle handle, TAPP_element_op op_a, TAPP_tensor_info a,
                   const int64_t* idx_a, TAPP_element_op op_b, TAPP_tensor_info b, const int64_t* idx_b,
                   TAPP_element_op op_c, TAPP_tensor_info c, const int64_t* idx_c, TAPP_element_op op_d,
                   TAPP_tensor_info d, const int64_t* idx_d, TAPP_prectype prec, const tprims_plan_config* cfgp) {
  return ffi([&] {
    if (!plan_out) fail(ERR_INVALID_ARGUMENT, "plan out-parameter is null or zero");
    if (handle == 0) fail(ERR_INVALID_ARGUMENT, "library handle is null or zero");
    auto *ta = (TensorInfo*)a, *tb = (TensorInfo*)b, *tc = (TensorInfo*)c, *td = (TensorInfo*)d;
    if (!ta || !tb || !tc || !td) fail(ERR_INVALID_ARGUMENT, "tensor info is null or zero");
    if (ta->dtype != tb->dtype || ta->dtype != tc->dtype || ta->dtype != td->dtype)
      fail(ERR_DTYPE, "operands have different storage types");
    const bool single = ta->dtype == TAPP_F32 || ta->dtype == TAPP_C32;
    const int storage_prec = single ? TAPP_F32F32_ACCUM_F32 : TAPP_F64F64_ACCUM_F64;
    PlanConfig cfg;
    if (cfgp) {
      cfg.complex_method = cfgp->complex_method;
      cfg.precision = cfgp->precision;
      cfg.tile = cfgp->tile;
      cfg.split_k = cfgp->split_k;
    }
    if (prec == TAPP_BF16BF16_ACCUM_F32 && single) {
      if (cfg.precision == PREC_STORAGE) cfg.precision = PREC_SPLIT_BF16_3;
    } else if (prec != TAPP_DEFAULT_PREC && prec != storage_prec) {
      fail(ERR_UNSUPPORTED, "computational precision " + std::to_string(prec) + " differs from the storage precision");
    }
    bool oa = op_of(op_a, "A"), ob = op_of(op_b, "B"), oc = op_of(op_c, "C"), od = op_of(op_d, "D");
    auto la = labels_of(idx_a, *ta, "A");
    auto lb = labels_of(idx_b, *tb, "B");
    auto lc = labels_of(idx_c, *tc, "C");
    auto ld = labels_of(idx_d, *td, "D");
    DType dt = dtype_of(ta->dtype);
    Operand A = operand_of(*ta, oa), B = operand_of(*tb, ob), C = operand_of(*tc, oc), D = operand_of(*td, od);
    Problem pr = lower(dt, A, la, B, lb, &C, &lc, D, ld);
    Plan* p = make_plan(pr, cfg);
    *plan_out = (TAPP_tensor_product)p;
  });
}

int execute_items(Plan* p, TAPP_executor exec, const void* alpha, const void* beta, const std::vector<Item>& items) {
  double al[2], be[2];
  read_scalar(alpha, p->prob.dtype, "alpha", al);
  read_scalar(beta, p->prob.dtype, "beta", be);
  const bool beta_zero = be[0] == 0.0 && be[1] == 0.0;
  for (auto& it : items) validate_item(*p, it, beta_zero);
  if (items.empty() || p->strategy == 2) return 0;
  Executor* ex = resolve_executor(exec);
  ex->inflight++;
  struct Release {
    Executor* e;
    ~Release() { e->inflight--; }
  } rel{ex};
  // A batch runs as one launch when that equals running its items in order
  // (TAPP runs them in order, crates/tprims-capi/src/execute.rs:235).
  if (items.size() > 1 && run_items_fused(*p, *ex, al, be, items)) return 0;
  const bool sync = !ex->async.load();
  for (auto& it : items) run_item(*p, *ex, al, be, it, sync);
  return 0;
}

}  // namespace

namespace tp {

void fill_report(const Plan* p, tprims_plan_report* out) {
  memset(out, 0, sizeof *out);
  out->m = p->M;
  out->n = p->N;
  out->k = p->K;
  out->batch = p->H;
  out->flops = p->strategy == 0 || p->strategy == 3 ? p->flops() : 0;
  out->dtype = (int32_t)p->prob.dtype;
  out->complex_method = p->cfg.complex_method;
  out->precision = p->cfg.precision;
  out->tile = p->tile;
  if (const TileCfg* t = tile_cfg(p->tile)) {
    out->tile_m = t->bm;
    out->tile_n = t->bn;
    out->tile_k = t->bk;
  }
  out->a_k_major = p->a_kmajor;
  out->b_k_major = p->b_kmajor;
  out->d_m_major = p->d_mmajor;
  out->m_axes = (int32_t)p->m_ax.size();
  out->n_axes = (int32_t)p->n_ax.size();
  out->k_axes = (int32_t)p->k_ax.size();
  out->h_axes = (int32_t)p->h_ax.size();
  out->strategy = p->strategy;
  out->source = p->source;
  std::string k = p->kernel_name();
  strncpy(out->kernel, k.c_str(), sizeof out->kernel - 1);
}

}  // namespace tp

extern "C" {

// ---- status (status.rs:69-94, 165-171) ----------------------------------------
bool TAPP_check_success(TAPP_error error) { return error == 0; }

size_t TAPP_explain_error(TAPP_error error, size_t maxlen, char* message) {
  const char* s = describe(error);
  size_t len = strlen(s);
  if (message && maxlen > 0) {
    size_t n = len < maxlen - 1 ? len : maxlen - 1;
    memcpy(message, s, n);
    message[n] = 0;
  }
  return len;
}

const char* tprims_last_error(void) { return g_last_error.c_str(); }

uint32_t tprims_abi_version(void) { return TPRIMS_ABI_VERSION; }

int32_t tprims_has_part(const char* part) {
  if (!part) return 0;
  return strcmp(part, "core") == 0 || strcmp(part, "tapp") == 0 || strcmp(part, "hip") == 0;
}

const char* TAPP_implementation_name(void) {
  return "tprims-mi355x: HIP block-scatter contraction for gfx950 (MFMA f64/f32, split-BF16)";
}
const char* TAPP_implementation_version(void) { return "0.1.0"; }

// ---- handles and attributes (handle.rs:35-137) --------------------------------
TAPP_error TAPP_create_handle(TAPP_handle* handle) {
  return ffi([&] {
    if (!handle) fail(ERR_INVALID_ARGUMENT, "handle out-parameter is null or zero");
    *handle = (TAPP_handle) new HandleState;
  });
}

TAPP_error TAPP_destroy_handle(TAPP_handle handle) {
  return ffi([&] {
    if (handle == 0) fail(ERR_INVALID_ARGUMENT, "handle is null or zero");
    delete (HandleState*)handle;
  });
}

TAPP_error TAPP_attr_set(TAPP_attr, TAPP_key, void*) {
  return ffi([] { fail(ERR_UNSUPPORTED, "no attribute keys are defined"); });
}
TAPP_error TAPP_attr_get(TAPP_attr, TAPP_key, void** value) {
  if (value) *value = nullptr;
  return ffi([] { fail(ERR_UNSUPPORTED, "no attribute keys are defined"); });
}
TAPP_error TAPP_attr_clear(TAPP_attr, TAPP_key) {
  return ffi([] { fail(ERR_UNSUPPORTED, "no attribute keys are defined"); });
}

TAPP_error TAPP_destroy_status(TAPP_status) {
  return ffi([] {});
}

// ---- tensor infos (tensor_info.rs:58-221) --------------------------------------
TAPP_error TAPP_create_tensor_info(TAPP_tensor_info* info, TAPP_datatype type, int nmode, const int64_t* extents,
                                   const int64_t* strides) {
  return ffi([&] {
    if (!info) fail(ERR_INVALID_ARGUMENT, "info out-parameter is null or zero");
    if (nmode < 0) fail(ERR_SHAPE, "negative mode count");
    if (type != TAPP_F32 && type != TAPP_F64 && type != TAPP_C32 && type != TAPP_C64)
      fail(ERR_DTYPE, "unsupported datatype");
    if (nmode > 0 && (!extents || !strides)) fail(ERR_INVALID_ARGUMENT, "extents or strides is null or zero");
    auto* t = new TensorInfo;
    t->dtype = type;
    if (nmode > 0) {
      t->extents.assign(extents, extents + nmode);
      t->strides.assign(strides, strides + nmode);
    }
    *info = (TAPP_tensor_info)t;
  });
}

TAPP_error TAPP_destroy_tensor_info(TAPP_tensor_info info) {
  return ffi([&] {
    if (info == 0) fail(ERR_INVALID_ARGUMENT, "info is null or zero");
    delete (TensorInfo*)info;
  });
}

int TAPP_get_nmodes(TAPP_tensor_info info) {
  if (info == 0) return -1;
  return (int)((TensorInfo*)info)->extents.size();
}

TAPP_error TAPP_set_nmodes(TAPP_tensor_info info, int nmodes) {
  return ffi([&] {
    if (info == 0) fail(ERR_INVALID_ARGUMENT, "info is null or zero");
    if (nmodes < 0) fail(ERR_SHAPE, "negative mode count");
    auto* t = (TensorInfo*)info;
    t->extents.resize(nmodes, 1);
    t->strides.resize(nmodes, 0);
  });
}

void TAPP_get_extents(TAPP_tensor_info info, int64_t* extents) {
  if (info == 0 || !extents) return;
  auto* t = (TensorInfo*)info;
  if (!t->extents.empty()) memcpy(extents, t->extents.data(), t->extents.size() * 8);
}

TAPP_error TAPP_set_extents(TAPP_tensor_info info, const int64_t* extents) {
  return ffi([&] {
    if (info == 0) fail(ERR_INVALID_ARGUMENT, "info is null or zero");
    if (!extents) fail(ERR_INVALID_ARGUMENT, "extents is null or zero");
    auto* t = (TensorInfo*)info;
    std::copy(extents, extents + t->extents.size(), t->extents.begin());
  });
}

void TAPP_get_strides(TAPP_tensor_info info, int64_t* strides) {
  if (info == 0 || !strides) return;
  auto* t = (TensorInfo*)info;
  if (!t->strides.empty()) memcpy(strides, t->strides.data(), t->strides.size() * 8);
}

TAPP_error TAPP_set_strides(TAPP_tensor_info info, const int64_t* strides) {
  return ffi([&] {
    if (info == 0) fail(ERR_INVALID_ARGUMENT, "info is null or zero");
    if (!strides) fail(ERR_INVALID_ARGUMENT, "strides is null or zero");
    auto* t = (TensorInfo*)info;
    std::copy(strides, strides + t->strides.size(), t->strides.begin());
  });
}

// ---- products (product.rs:83-207) ---------------------------------------------
TAPP_error TAPP_create_tensor_product(TAPP_tensor_product* plan, TAPP_handle handle, TAPP_element_op op_A,
                                      TAPP_tensor_info A, const int64_t* idx_A, TAPP_element_op op_B,
                                      TAPP_tensor_info B, const int64_t* idx_B, TAPP_element_op op_C,
                                      TAPP_tensor_info C, const int64_t* idx_C, TAPP_element_op op_D,
                                      TAPP_tensor_info D, const int64_t* idx_D, TAPP_prectype prec) {
  return create_product(plan, handle, op_A, A, idx_A, op_B, B, idx_B, op_C, C, idx_C, op_D, D, idx_D, prec, nullptr);
}

TAPP_error tprims_create_tensor_product_ex(TAPP_tensor_product* plan, TAPP_handle handle, TAPP_element_op op_A,
                                           TAPP_tensor_info A, const int64_t* idx_A, TAPP_element_op op_B,
                                           TAPP_tensor_info B, const int64_t* idx_B, TAPP_element_op op_C,
                                           TAPP_tensor_info C, const int64_t* idx_C, TAPP_element_op op_D,
                                           TAPP_tensor_info D, const int64_t* idx_D, TAPP_prectype prec,
                                           const tprims_plan_config* cfg) {
  return create_product(plan, handle, op_A, A, idx_A, op_B, B, idx_B, op_C, C, idx_C, op_D, D, idx_D, prec, cfg);
}

TAPP_error TAPP_destroy_tensor_product(TAPP_tensor_product plan) {
  return ffi([&] {
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    Plan::release((Plan*)plan);  // a group built on it keeps it alive
  });
}

TAPP_error tprims_product_report(TAPP_tensor_product plan, tprims_plan_report* out) {
  return ffi([&] {
    if (plan == 0 || !out) fail(ERR_INVALID_ARGUMENT, "plan or report is null or zero");
    fill_report((Plan*)plan, out);
  });
}

// ---- scatter vectors (crates/tprims-kernel/src/pack/scatter.rs) -------------
namespace {
void copy_scatter(const std::vector<int64_t>& v, int64_t* out, int64_t cap, int64_t* len) {
  if (!len) fail(ERR_INVALID_ARGUMENT, "length pointer is null");
  if (cap < 0) fail(ERR_INVALID_ARGUMENT, "negative capacity");
  if (cap > 0 && !out) fail(ERR_INVALID_ARGUMENT, "output buffer is null");
  const int64_t n = (int64_t)v.size(), k = n < cap ? n : cap;
  if (k > 0) std::memcpy(out, v.data(), (size_t)k * sizeof(int64_t));
  *len = n;
}
}  // namespace

TAPP_error tprims_build_scatter(const int64_t* extents, const int64_t* strides, int32_t nmode, int64_t* out,
                                int64_t cap, int64_t* len) {
  return ffi([&] {
    if (nmode < 0) fail(ERR_INVALID_ARGUMENT, "negative mode count");
    if (nmode > 0 && (!extents || !strides)) fail(ERR_INVALID_ARGUMENT, "extents or strides are null");
    std::vector<int64_t> e(extents, extents + nmode), s(strides, strides + nmode);
    bool empty = false;
    for (int64_t x : e) {
      if (x < 0) fail(ERR_SHAPE, "negative extent");
      empty = empty || x == 0;
    }
    if (!empty) {
      __int128 total = 1;
      for (int64_t x : e)
        if ((total *= x) > ((__int128)1 << 32)) fail(ERR_SHAPE, "scatter vector longer than 2^32 entries");
    }
    copy_scatter(empty ? std::vector<int64_t>{} : build_scatter(e, s), out, cap, len);
  });
}

TAPP_error tprims_build_block_scatter(const int64_t* scat, int64_t n, int64_t blk, int64_t* out, int64_t cap,
                                      int64_t* nblocks) {
  return ffi([&] {
    if (blk <= 0) fail(ERR_INVALID_ARGUMENT, "block size must be positive");
    if (n < 0) fail(ERR_INVALID_ARGUMENT, "negative scatter length");
    if (n > 0 && !scat) fail(ERR_INVALID_ARGUMENT, "scatter vector is null");
    copy_scatter(build_block_scatter(scat, n, blk), out, cap, nblocks);
  });
}

TAPP_error tprims_product_scatter(TAPP_tensor_product plan, int32_t which, int64_t* out, int64_t cap,
                                  int64_t* len) {
  return ffi([&] {
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    const Plan* p = (const Plan*)plan;
    const std::vector<int64_t>* v[12] = {&p->am, &p->ak, &p->bk, &p->bn, &p->cm, &p->cn,
                                         &p->dm, &p->dn, &p->ha, &p->hb, &p->hc, &p->hd};
    if (which < 0 || which >= 12) fail(ERR_INVALID_ARGUMENT, "scatter id out of range");
    copy_scatter(*v[which], out, cap, len);
  });
}

// ---- execution (execute.rs:192-279) -------------------------------------------
TAPP_error TAPP_execute_product(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status, const void* alpha,
                                const void* A, const void* B, const void* beta, const void* C, void* D) {
  return ffi([&] {
    if (status) *status = 0;
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    execute_items((Plan*)plan, exec, alpha, beta, {Item{A, B, C, D}});
  });
}

TAPP_error TAPP_execute_batched_product(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status,
                                        int num_batches, const void* alpha, const void** A, const void** B,
                                        const void* beta, const void** C, void** D) {
  return ffi([&] {
    if (status) *status = 0;
    if (num_batches < 0) fail(ERR_SHAPE, "negative batch count");
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    if (!A || !B || !D) fail(ERR_INVALID_ARGUMENT, "A, B or D pointer array is null or zero");
    std::vector<Item> items;
    for (int i = 0; i < num_batches; i++) items.push_back(Item{A[i], B[i], C ? C[i] : nullptr, D[i]});
    execute_items((Plan*)plan, exec, alpha, beta, items);
  });
}

// ---- executors (executor.rs:278-392) -------------------------------------------
static int current_device_or_zero() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) {
    (void)hipGetLastError();
    d = 0;
  }
  return d;
}

TAPP_error TAPP_create_executor(TAPP_executor* exec) {
  return ffi([&] {
    if (!exec) fail(ERR_INVALID_ARGUMENT, "null executor out-parameter");
    *exec = (TAPP_executor)new_executor(current_device_or_zero(), nullptr);
  });
}

TAPP_error TAPP_destroy_executor(TAPP_executor exec) {
  return ffi([&] {
    if (exec == 0) return;
    Executor* e = (Executor*)exec;
    if (e->inflight.load() > 0) fail(BUSY, "executor has calls in flight");
    // Work still queued on the executor's stream (an asynchronous call's)
    // counts as in flight.
    if (e->busy())
      fail(BUSY, "work is still queued on the executor's stream; synchronize it first "
                 "(tprims_tapp_executor_synchronize)");
    delete e;
  });
}

TAPP_error tprims_tapp_executor_create_rayon(TAPP_executor* out, size_t nthreads, const tprims_rayon_opts*) {
  return ffi([&] {
    if (!out) fail(ERR_INVALID_ARGUMENT, "null executor out-parameter");
    if (nthreads == 0) fail(ERR_INVALID_ARGUMENT, "nthreads must be at least 1");
    Executor* e = new_executor(current_device_or_zero(), nullptr);
    if (nthreads >= 2) {
      e->pool_size = nthreads;
      e->budget = nthreads;
    }
    *out = (TAPP_executor)e;
  });
}

TAPP_error tprims_tapp_executor_set_budget(TAPP_executor exec, size_t budget) {
  return ffi([&] {
    if (budget == 0) fail(ERR_INVALID_ARGUMENT, "budget must be at least 1");
    if (exec == 0) return;
    Executor* e = (Executor*)exec;
    if (e->pool_size > 0) e->budget = budget < e->pool_size ? budget : e->pool_size;
  });
}

TAPP_error tprims_tapp_executor_get_threads(TAPP_executor exec, size_t* pool_size, size_t* budget) {
  return ffi([&] {
    size_t p = 0, b = 1;
    if (exec != 0) {
      Executor* e = (Executor*)exec;
      p = e->pool_size;
      b = e->pool_size > 0 ? e->budget : 1;
    }
    if (pool_size) *pool_size = p;
    if (budget) *budget = b;
  });
}

TAPP_error tprims_tapp_executor_create_hip(TAPP_executor* out, int device, void* stream) {
  return ffi([&] {
    if (!out) fail(ERR_INVALID_ARGUMENT, "null executor out-parameter");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
      (void)hipGetLastError();
      fail(ERR_DEVICE, "no HIP device " + std::to_string(device));
    }
    *out = (TAPP_executor)new_executor(device, (hipStream_t)stream);
  });
}

TAPP_error tprims_tapp_executor_set_async(TAPP_executor exec, int async) {
  return ffi([&] {
    if (exec == 0) fail(ERR_INVALID_ARGUMENT, "the default executor is always synchronous");
    ((Executor*)exec)->async.store(async != 0);
  });
}

TAPP_error tprims_tapp_executor_synchronize(TAPP_executor exec) {
  return ffi([&] {
    Executor* e = resolve_executor(exec);
    DeviceGuard g(e->device);
    hip_check(hipStreamSynchronize(e->get_stream()), "hipStreamSynchronize");
  });
}

TAPP_error tprims_tapp_executor_get_stream(TAPP_executor exec, void** stream) {
  return ffi([&] {
    if (!stream) fail(ERR_INVALID_ARGUMENT, "stream out-parameter is null");
    Executor* e = resolve_executor(exec);
    *stream = (void*)e->get_stream();
  });
}

TAPP_error tprims_tapp_executor_get_device(TAPP_executor exec, int* device) {
  return ffi([&] {
    if (!device) fail(ERR_INVALID_ARGUMENT, "device out-parameter is null");
    *device = exec == 0 ? current_device_or_zero() : ((Executor*)exec)->device;
  });
}

// ---- DLPack operands (tensor.rs:96-230 operand contract) ----------------------
namespace {

// A validated borrowed operand (crates/tprims-capi/src/tensor.rs:128-141 Layout).
struct DlLayout {
  std::vector<int64_t> dims, strides;
  const char* origin = nullptr;  // data + byte_offset; a never-read placeholder when empty
  int64_t lo = 0, hi = 0;        // smallest and largest element offset addressed
  bool empty = false;
};

// Address handed out as the origin of an empty tensor: non-null, aligned for
// every element type, never dereferenced (the reference uses a dangling pointer).
alignas(16) const char k_empty_origin[16] = {};

DType dl_dtype(const tprims_tensor& t) {  // tensor.rs:96-115
  if (!t.view) fail(ERR_INVALID_ARGUMENT, "null DLTensor");
  const DLDataType d = t.view->dtype;
  if (d.lanes != 1) fail(ERR_DTYPE, "lanes = " + std::to_string(d.lanes) + " (must be 1)");
  if (d.code == kDLFloat && d.bits == 32) return DType::F32;
  if (d.code == kDLFloat && d.bits == 64) return DType::F64;
  if (d.code == kDLComplex && d.bits == 64) return DType::C32;
  if (d.code == kDLComplex && d.bits == 128) return DType::C64;
  fail(ERR_DTYPE, "unsupported dtype code " + std::to_string(d.code) + " bits " + std::to_string(d.bits));
}

// Validates device, dtype, shape, strides and offset with checked arithmetic
// and a typed status per rejected field, in the order of tensor.rs:145-230;
// `writable` adds the read-only check of an output (tensor.rs:292-297).
// ROCm device memory is accepted here; the executions compare its device
// with the executor's.
DlLayout dl_layout(const tprims_tensor& t, DType dt, bool writable) {
  if (!t.view) fail(ERR_INVALID_ARGUMENT, "null DLTensor");
  const DLTensor& v = *t.view;
  switch (v.device.device_type) {
    case kDLCPU:
    case kDLCUDAHost:
    case kDLROCMHost:
    case kDLROCM: break;
    default: fail(ERR_DEVICE, "device type " + std::to_string(v.device.device_type) + " is not CPU or ROCm");
  }
  if (dl_dtype(t) != dt) fail(ERR_DTYPE, "operand dtypes differ");
  if (v.ndim < 0) fail(ERR_SHAPE, "negative ndim");
  const int nd = v.ndim;
  if (nd > 0 && !v.shape) fail(ERR_INVALID_ARGUMENT, "null shape");
  DlLayout l;
  l.dims.assign(nd, 0);
  l.strides.assign(nd, 0);
  for (int i = 0; i < nd; i++) {
    if (v.shape[i] < 0) fail(ERR_SHAPE, "negative extent");
    l.dims[i] = v.shape[i];
    l.empty = l.empty || v.shape[i] == 0;
  }
  if (!v.strides) {  // compact row-major (the DLPack convention)
    __int128 acc = 1;
    for (int i = nd - 1; i >= 0; i--) {
      l.strides[i] = (int64_t)acc;
      acc *= l.dims[i] > 1 ? l.dims[i] : 1;
      if (acc > (__int128)INT64_MAX) fail(ERR_SHAPE, "stride overflow");
    }
  } else {
    for (int i = 0; i < nd; i++) l.strides[i] = v.strides[i];
  }
  const int64_t es = elem_size(dt);
  if (!l.empty) {
    __int128 lo = 0, hi = 0;
    for (int i = 0; i < nd; i++) {
      const __int128 reach = (__int128)(l.dims[i] - 1) * l.strides[i];
      if (reach < 0) lo += reach;
      else hi += reach;
      if (lo < (__int128)INT64_MIN || hi > (__int128)INT64_MAX) fail(ERR_SHAPE, "span overflow");
    }
    if (!v.data) fail(ERR_INVALID_ARGUMENT, "null data for a non-empty tensor");
    if ((hi - lo + 1) * es > (__int128)INT64_MAX) fail(ERR_SHAPE, "span overflow");
    l.lo = (int64_t)lo;
    l.hi = (int64_t)hi;
    if (v.byte_offset > (uint64_t)INT64_MAX) fail(ERR_SHAPE, "byte_offset overflow");
    l.origin = (const char*)v.data + v.byte_offset;
    const uintptr_t align = is_double(dt) ? 8 : 4;  // a complex element aligns like its real part
    if ((uintptr_t)l.origin % align != 0)
      fail(ERR_INVALID_ARGUMENT, "data + byte_offset is not aligned to " + std::to_string(align) + " bytes");
  } else {
    l.origin = k_empty_origin;
  }
  if (writable && (t.flags & DLPACK_FLAG_BITMASK_READ_ONLY)) fail(ERR_READ_ONLY, "output tensor is read-only");
  return l;
}

// The operand's view must describe the layout the plan was built from.
const void* dlpack_origin(const tprims_tensor& t, const Operand& o, DType dt, const char* name, bool output) {
  DlLayout l;
  try {
    l = dl_layout(t, dt, output);
  } catch (const Error& e) {
    fail(e.status, std::string(name) + ": " + e.what());
  }
  if ((int)l.dims.size() != o.layout.rank()) fail(ERR_SHAPE, std::string(name) + ": rank differs from the plan's");
  for (size_t i = 0; i < l.dims.size(); i++)
    if (l.dims[i] != o.layout.dims[i] || (l.dims[i] > 1 && l.strides[i] != o.layout.strides[i]))
      fail(ERR_SHAPE, std::string(name) + ": extents or strides differ from the plan's");
  return l.origin;
}

}  // namespace

TAPP_error tprims_tensor_dtype(tprims_tensor t, TAPP_datatype* out) {
  return ffi([&] {
    if (!out) fail(ERR_INVALID_ARGUMENT, "dtype out-parameter is null");
    *out = (TAPP_datatype)dl_dtype(t);
  });
}

TAPP_error tprims_tensor_layout(tprims_tensor t, TAPP_datatype dtype, int writable, int64_t* dims, int64_t* strides,
                                int32_t cap, tprims_tensor_layout_info* out) {
  return ffi([&] {
    if (!out) fail(ERR_INVALID_ARGUMENT, "layout out-parameter is null");
    if (cap < 0 || (cap > 0 && (!dims || !strides))) fail(ERR_INVALID_ARGUMENT, "dims/strides buffers are null");
    DlLayout l = dl_layout(t, dtype_of(dtype), writable != 0);
    const int nd = (int)l.dims.size();
    for (int i = 0; i < nd && i < cap; i++) {
      dims[i] = l.dims[i];
      strides[i] = l.strides[i];
    }
    out->ndim = nd;
    out->empty = l.empty;
    out->origin = (void*)l.origin;
    out->span_lo = l.lo;
    out->span_hi = l.hi;
  });
}

TAPP_error tprims_execute_product_dlpack(TAPP_tensor_product plan, TAPP_executor exec, const void* alpha,
                                         tprims_tensor A, tprims_tensor B, const void* beta, tprims_tensor C,
                                         tprims_tensor D) {
  return ffi([&] {
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    Plan* p = (Plan*)plan;
    const DType dt = p->prob.dtype;
    const void* a = dlpack_origin(A, p->prob.a, dt, "A", false);
    const void* b = dlpack_origin(B, p->prob.b, dt, "B", false);
    const void* c = C.view ? dlpack_origin(C, p->prob.c, dt, "C", false) : nullptr;
    void* d = const_cast<void*>(dlpack_origin(D, p->prob.d, dt, "D", true));
    // Device memory must be the executor's (resolved only now, so that every
    // field error above is reported without touching the GPU runtime).
    const tprims_tensor* ops[4] = {&A, &B, &C, &D};
    const char* names[4] = {"A", "B", "C", "D"};
    for (int i = 0; i < 4; i++)
      if (ops[i]->view && ops[i]->view->device.device_type == kDLROCM) {
        const int device = resolve_executor(exec)->device;
        if (ops[i]->view->device.device_id != device)
          fail(ERR_DEVICE, std::string(names[i]) + ": ROCm tensor on device " +
                               std::to_string(ops[i]->view->device.device_id) + ", the executor on device " +
                               std::to_string(device));
      }
    execute_items(p, exec, alpha, beta, {Item{a, b, c, d}});
  });
}

// ---- DLPack borrow (tensor.rs:26-55) -------------------------------------------
tprims_tensor tprims_tensor_borrow_versioned(struct DLManagedTensorVersioned* m) {
  tprims_tensor t{nullptr, 0};
  if (!m || m->version.major != 1) return t;
  t.view = &m->dl_tensor;
  t.flags = m->flags;
  return t;
}

tprims_tensor tprims_tensor_borrow_raw(DLTensor* v, uint64_t flags) { return tprims_tensor{v, flags}; }

}  // extern "C"
