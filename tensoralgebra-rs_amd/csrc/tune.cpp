// Per-plan autotuning and the tuning cache. The reference resolves a kernel
// family and blocking once per plan and freezes it there
// (crates/tprims-contract/src/select/resolve.rs:164); here a plan's tile and
// K split come from the planner's cost model unless a tuning cache entry
// exists for its reduced problem, and tprims_product_autotune measures the
// candidates on the caller's operands and freezes the fastest into the plan.
// The cache lives in the process and is saved to / loaded from a file whose
// path the caller passes (nothing is read from the environment).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "capi_util.h"
#include "executor.h"
#include "kernels.h"
#include "plan.h"
#include "tprims/tprims.h"

namespace tp {
namespace {

struct Entry {
  int tile = 0;
  int split_k = 0;
};

std::mutex g_mu;
std::map<std::string, Entry>& cache() {
  static std::map<std::string, Entry> m;
  return m;
}

const char* kHeader = "# tprims tuning cache 1";

void axes_to(std::ostringstream& o, char role, const std::vector<Axis>& v) {
  o << ';' << role << '=';
  for (size_t i = 0; i < v.size(); i++) {
    if (i) o << ',';
    o << v[i].extent << ':' << v[i].sa << ':' << v[i].sb << ':' << v[i].sc << ':' << v[i].sd;
  }
}

}  // namespace

// The reduced problem exactly: dtype, complex method, precision, strategy
// and every folded axis of every role with its extent and four strides. Two
// plans with equal keys run the same kernels on the same offsets.
std::string tuning_key(const Plan& p, const PlanConfig& cfg) {
  std::ostringstream o;
  o << "dt=" << (int)p.prob.dtype << ";cm=" << cfg.complex_method << ";pr=" << cfg.precision << ";st=" << p.strategy;
  axes_to(o, 'm', p.m_ax);
  axes_to(o, 'n', p.n_ax);
  axes_to(o, 'k', p.k_ax);
  axes_to(o, 'h', p.h_ax);
  axes_to(o, 'r', p.r_ax);
  return o.str();
}

bool tuning_lookup(const std::string& key, int* tile, int* split_k) {
  std::lock_guard<std::mutex> g(g_mu);
  auto it = cache().find(key);
  if (it == cache().end()) return false;
  *tile = it->second.tile;
  *split_k = it->second.split_k;
  return true;
}

void tuning_store(const std::string& key, int tile, int split_k) {
  std::lock_guard<std::mutex> g(g_mu);
  cache()[key] = Entry{tile, split_k};
}

namespace {

// Re-plans `p` under `cfg` and drops its device copies of the scatter
// vectors when they no longer match (the K order can change with the tile).
void replan(Plan& p, const PlanConfig& cfg) {
  const std::vector<int64_t> before = p.flat_tables();
  configure(&p, cfg);
  if (p.flat_tables() == before) return;
  std::lock_guard<std::mutex> g(p.mu);
  for (auto& [d, s] : p.dev) {
    DeviceGuard dg(d);
    (void)hipFree(s.base);
  }
  p.dev.clear();
  p.host_image.clear();
}

// The tiles a plan of this precision can run (tile_allowed filters per dtype
// and complex method).
std::vector<int> tile_candidates(int prec) {
  if (prec == PREC_STORAGE) return {TILE_128x128, TILE_128x64, TILE_64x64, TILE_32x32, TILE_32x32K, TILE_32x32R, TILE_32x48R};
  if (prec == PREC_FP8) return {STILE_256x256, STILE_128x64};
  return {STILE_256x128, STILE_128x128, STILE_128x64, STILE_64x64};
}

}  // namespace
}  // namespace tp

using namespace tp;

extern "C" {

TAPP_error tprims_product_autotune(TAPP_tensor_product plan, TAPP_executor exec, const void* alpha, const void* A,
                                   const void* B, const void* beta, const void* C, void* D, int32_t reps,
                                   int32_t* best_tile, int32_t* best_split, double* best_ms) {
  return ffi([&] {
    if (plan == 0) fail(ERR_INVALID_ARGUMENT, "plan is null or zero");
    if (reps < 1 || reps > 1000) fail(ERR_INVALID_ARGUMENT, "reps must be in [1, 1000]");
    Plan& p = *(Plan*)plan;
    double al[2], be[2];
    read_scalar(alpha, p.prob.dtype, "alpha", al);
    read_scalar(beta, p.prob.dtype, "beta", be);
    const bool beta_zero = be[0] == 0.0 && be[1] == 0.0;
    const Item it{A, B, C, D};
    validate_item(p, it, beta_zero);
    Executor* ex = resolve_executor(exec);
    ex->inflight++;
    struct Release {
      Executor* e;
      ~Release() { e->inflight--; }
    } rel{ex};
    DeviceGuard dg(ex->device);
    const bool read_c = C != nullptr && !beta_zero;
    for (const void* q : {A, B, read_c ? C : nullptr, (const void*)D})
      if (q && classify(q, ex->device).host)
        fail(ERR_DEVICE, "autotuning runs on device operands of the executor's device");

    const PlanConfig base{p.cfg.complex_method, p.cfg.precision, 0, 0};
    const std::string key = tuning_key(p, base);
    hipStream_t s = ex->get_stream();

    // An output that is also read (C overlapping D, beta != 0) accumulates
    // on every timing run: its span is saved first and restored before the
    // one execution that leaves the result.
    const Range rd = byte_range(p, OP_D, D);
    const Range rc = read_c ? byte_range(p, OP_C, C) : Range{};
    const bool keep = read_c && rd.present && overlap(rd, rc);
    char* saved = nullptr;
    const size_t dbytes = keep ? (size_t)(rd.hi - rd.lo) : 0;
    struct Free {
      char*& q;
      ~Free() {
        if (q) (void)hipFree(q);
      }
    } fr{saved};
    if (keep) {
      hip_check(hipMalloc((void**)&saved, dbytes), "hipMalloc(autotune snapshot)");
      hip_check(hipMemcpyAsync(saved, (const void*)(intptr_t)rd.lo, dbytes, hipMemcpyDeviceToDevice, s),
                "saving the in-place output");
    }

    PlanConfig best = p.cfg;
    float best_t = 1e30f;
    if (p.strategy == 0) {
      hipEvent_t e0, e1;
      hip_check(hipEventCreate(&e0), "hipEventCreate");
      hip_check(hipEventCreate(&e1), "hipEventCreate");
      struct Events {
        hipEvent_t a, b;
        ~Events() {
          (void)hipEventDestroy(a);
          (void)hipEventDestroy(b);
        }
      } evs{e0, e1};
      std::vector<std::string> seen;
      for (int t : tile_candidates(base.precision)) {
        if (!tile_allowed(t, p.prob.dtype, base.precision, base.complex_method)) continue;
        for (int sk : {0, 1}) {
          if (sk == 1 && base.precision != PREC_STORAGE) continue;
          PlanConfig c = base;
          c.tile = t;
          c.split_k = sk;
          replan(p, c);
          // Candidates that plan to the same kernels are timed once.
          const std::string id = p.kernel_name();
          if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
          seen.push_back(id);
          run_item(p, *ex, al, be, it, false);  // warm-up: uploads, first-launch costs
          hip_check(hipEventRecord(e0, s), "hipEventRecord");
          for (int r = 0; r < reps; r++) run_item(p, *ex, al, be, it, false);
          hip_check(hipEventRecord(e1, s), "hipEventRecord");
          hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
          float ms = 0;
          hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
          ms /= (float)reps;
          if (ms < best_t) {
            best_t = ms;
            best = c;
          }
        }
      }
      replan(p, best);
      p.source = 3;
      tuning_store(key, best.tile, best.split_k);
    }
    if (keep)
      hip_check(hipMemcpyAsync((void*)(intptr_t)rd.lo, saved, dbytes, hipMemcpyDeviceToDevice, s),
                "restoring the in-place output");
    // One execution on the chosen configuration leaves D as one call does.
    run_item(p, *ex, al, be, it, true);
    if (best_tile) *best_tile = p.tile;
    if (best_split) *best_split = p.ksplit;
    if (best_ms) *best_ms = p.strategy == 0 ? (double)best_t : 0.0;
  });
}

TAPP_error tprims_product_tuning_key(TAPP_tensor_product plan, char* out, int64_t cap, int64_t* len) {
  return ffi([&] {
    if (plan == 0 || !len) fail(ERR_INVALID_ARGUMENT, "plan or length pointer is null or zero");
    if (cap < 0 || (cap > 0 && !out)) fail(ERR_INVALID_ARGUMENT, "output buffer is null");
    const Plan& p = *(const Plan*)plan;
    const std::string k = tuning_key(p, PlanConfig{p.cfg.complex_method, p.cfg.precision, 0, 0});
    *len = (int64_t)k.size();
    if (cap > 0) {
      const size_t n = std::min((size_t)cap - 1, k.size());
      memcpy(out, k.data(), n);
      out[n] = 0;
    }
  });
}

TAPP_error tprims_tuning_cache_save(const char* path) {
  return ffi([&] {
    if (!path) fail(ERR_INVALID_ARGUMENT, "path is null");
    std::ostringstream o;
    o << kHeader << "\n# tile split_k key\n";
    {
      std::lock_guard<std::mutex> g(g_mu);
      for (auto& [k, e] : cache()) o << e.tile << ' ' << e.split_k << ' ' << k << '\n';
    }
    // Written beside the target and renamed over it, so a reader never sees
    // a partial file.
    const std::string tmp = std::string(path) + ".tmp";
    {
      std::ofstream f(tmp, std::ios::trunc);
      if (!f) fail(ERR_INVALID_ARGUMENT, std::string("cannot write ") + tmp);
      f << o.str();
      if (!f.good()) fail(ERR_INVALID_ARGUMENT, std::string("cannot write ") + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) fail(ERR_INVALID_ARGUMENT, std::string("cannot replace ") + path);
  });
}

TAPP_error tprims_tuning_cache_load(const char* path) {
  return ffi([&] {
    if (!path) fail(ERR_INVALID_ARGUMENT, "path is null");
    std::ifstream f(path);
    if (!f) fail(ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    std::string line;
    if (!std::getline(f, line) || line != kHeader)
      fail(ERR_INVALID_ARGUMENT, std::string(path) + ": not a tuning cache (first line must be '" + kHeader + "')");
    std::map<std::string, Entry> got;
    int64_t no = 1;
    while (std::getline(f, line)) {
      no++;
      if (line.empty() || line[0] == '#') continue;
      std::istringstream in(line);
      Entry e;
      std::string key, extra;
      if (!(in >> e.tile >> e.split_k >> key) || (in >> extra) || e.split_k < 0 || e.split_k > 64 || e.tile < 0)
        fail(ERR_INVALID_ARGUMENT, std::string(path) + ":" + std::to_string(no) + ": expected 'tile split_k key'");
      got[key] = e;
    }
    // All or nothing: the entries are merged only once the whole file parsed.
    std::lock_guard<std::mutex> g(g_mu);
    for (auto& [k, e] : got) cache()[k] = e;
  });
}

TAPP_error tprims_tuning_cache_clear(void) {
  return ffi([&] {
    std::lock_guard<std::mutex> g(g_mu);
    cache().clear();
  });
}

TAPP_error tprims_tuning_cache_size(int64_t* n) {
  return ffi([&] {
    if (!n) fail(ERR_INVALID_ARGUMENT, "count pointer is null");
    std::lock_guard<std::mutex> g(g_mu);
    *n = (int64_t)cache().size();
  });
}

}  // extern "C"
