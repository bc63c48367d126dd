// Executor runtime: stream and scratch ownership, the pre-write validation
// of an execution, host-operand staging and the launch of the planned kernel.
#include "executor.h"

#include <algorithm>
#include <climits>
#include <map>
#include <string>
#include <vector>

#include "kernels.h"

namespace tp {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(ERR_INTERNAL, std::string(what) + ": " + hipGetErrorString(e));
  }
}

DeviceGuard::DeviceGuard(int d) {
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) fail(ERR_DEVICE, "no HIP device is available");
  if (cur != d) {
    hip_check(hipSetDevice(d), "hipSetDevice");
    prev = cur;
  }
}
DeviceGuard::~DeviceGuard() {
  if (prev >= 0) (void)hipSetDevice(prev);
}

hipStream_t Executor::get_stream() {
  std::lock_guard<std::mutex> g(mu);
  if (!stream_ready) {
    if (owns_stream) {
      DeviceGuard dg(device);
      hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    stream_ready = true;
  }
  return stream;
}

char* Executor::ensure_scratch(size_t bytes) {
  if (bytes <= scratch_cap) return scratch;
  if (scratch) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_check(hipFree(scratch), "hipFree");
    scratch = nullptr;
    scratch_cap = 0;
  }
  size_t cap = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
  hip_check(hipMalloc((void**)&scratch, cap), "hipMalloc(scratch)");
  scratch_cap = cap;
  return scratch;
}

char* Executor::ensure_panels(size_t bytes) {
  if (bytes <= panels_cap) return panels;
  if (panels) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_check(hipFree(panels), "hipFree");
    panels = nullptr;
    panels_cap = 0;
  }
  size_t cap = bytes + bytes / 8;
  hip_check(hipMalloc((void**)&panels, cap), "hipMalloc(panels)");
  panels_cap = cap;
  return panels;
}

char* Executor::ensure_batch(size_t bytes) {
  if (bytes <= batch_cap) return batch_vecs;
  if (batch_vecs) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_check(hipFree(batch_vecs), "hipFree");
    batch_vecs = nullptr;
    batch_cap = 0;
  }
  size_t cap = bytes < ((size_t)1 << 16) ? ((size_t)1 << 16) : bytes + bytes / 8;
  hip_check(hipMalloc((void**)&batch_vecs, cap), "hipMalloc(batch offsets)");
  batch_cap = cap;
  return batch_vecs;
}

char* Executor::ensure_batch_host(size_t bytes) {
  if (batch_ev_pending) {
    hip_check(hipEventSynchronize(batch_ev), "hipEventSynchronize");
    batch_ev_pending = false;
  }
  if (!batch_ev) hip_check(hipEventCreateWithFlags(&batch_ev, hipEventDisableTiming), "hipEventCreate");
  if (bytes <= batch_host_cap) return batch_host;
  if (batch_host) {
    hip_check(hipHostFree(batch_host), "hipHostFree");
    batch_host = nullptr;
    batch_host_cap = 0;
  }
  size_t cap = bytes < ((size_t)1 << 16) ? ((size_t)1 << 16) : bytes + bytes / 8;
  hip_check(hipHostMalloc((void**)&batch_host, cap, hipHostMallocDefault), "hipHostMalloc(batch offsets)");
  batch_host_cap = cap;
  return batch_host;
}

void Executor::finish(hipStream_t s, bool sync) {
  if (sync) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

bool Executor::busy() {
  std::lock_guard<std::mutex> g(mu);
  if (!stream_ready || (!stream && !owns_stream)) return false;
  DeviceGuard dg(device);
  const hipError_t e = hipStreamQuery(stream);
  if (e == hipErrorNotReady) return true;
  if (e != hipSuccess) (void)hipGetLastError();
  return false;
}

Executor::~Executor() {
  if (stream_ready && stream) (void)hipStreamSynchronize(stream);
  if (scratch) (void)hipFree(scratch);
  if (panels) (void)hipFree(panels);
  if (batch_vecs) (void)hipFree(batch_vecs);
  if (batch_host) (void)hipHostFree(batch_host);
  if (batch_ev) (void)hipEventDestroy(batch_ev);
  if (owns_stream && stream_ready && stream) (void)hipStreamDestroy(stream);
}

Executor* new_executor(int device, hipStream_t stream) {
  Executor* e = new Executor;
  e->device = device;
  e->stream = stream;
  e->owns_stream = stream == nullptr;
  e->stream_ready = stream != nullptr;
  return e;
}

Executor* resolve_executor(intptr_t handle) {
  if (handle != 0) return reinterpret_cast<Executor*>(handle);
  static std::mutex mu;
  static std::map<int, Executor*> defaults;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    fail(ERR_DEVICE, "no HIP device is available");
  }
  std::lock_guard<std::mutex> g(mu);
  auto it = defaults.find(dev);
  if (it != defaults.end()) return it->second;
  Executor* e = new Executor;  // lives until the process ends
  e->device = dev;
  e->stream = nullptr;  // the null stream
  e->owns_stream = false;
  e->stream_ready = true;
  e->is_default = true;
  defaults[dev] = e;
  return e;
}

Range byte_range(const Plan& p, int op, const void* origin) {
  Range r;
  if (!origin) return r;
  const Span& s = p.prob.spans[op];
  if (!s.present) return r;
  int64_t off = 0;
  switch (op) {
    case OP_A: off = p.prob.a.layout.offset; break;
    case OP_B: off = p.prob.b.layout.offset; break;
    case OP_C: off = p.prob.c.layout.offset; break;
    case OP_D: off = p.prob.d.layout.offset; break;
  }
  const __int128 base = (__int128)(uintptr_t)origin, es = elem_size(p.prob.dtype);
  r.present = true;
  r.lo = base + ((__int128)s.lo - off) * es;
  r.hi = base + ((__int128)s.hi - off + 1) * es;
  return r;
}

bool overlap(const Range& x, const Range& y) { return x.present && y.present && x.lo < y.hi && y.lo < x.hi; }

bool items_independent(const std::vector<ItemSpans>& items) {
  const size_t n = items.size();
  std::vector<std::pair<Range, size_t>> outs;
  outs.reserve(n);
  for (size_t i = 0; i < n; i++)
    if (items[i].d.present) outs.push_back({items[i].d, i});
  std::sort(outs.begin(), outs.end(), [](const auto& x, const auto& y) { return x.first.lo < y.first.lo; });
  for (size_t i = 1; i < outs.size(); i++)
    if (outs[i].first.lo < outs[i - 1].first.hi) return false;
  // Does r meet an output of an item other than `owner`? Outputs are now
  // disjoint and sorted, so only those starting below r.hi and ending above
  // r.lo, found by binary search and walking down, can.
  auto hits_other_output = [&](const Range& r, size_t owner) {
    if (!r.present) return false;
    size_t lo = 0, hi = outs.size();
    while (lo < hi) {
      const size_t mid = (lo + hi) / 2;
      if (outs[mid].first.lo < r.hi) lo = mid + 1;
      else hi = mid;
    }
    for (size_t k = lo; k-- > 0;) {
      if (outs[k].first.hi <= r.lo) break;
      if (outs[k].second != owner) return true;
    }
    return false;
  };
  for (size_t i = 0; i < n; i++)
    for (const Range& r : items[i].in)
      if (hits_other_output(r, i)) return false;
  return true;
}

Where classify(const void* p, int device) {
  Where w;
  if (!p) return w;
  hipPointerAttribute_t at{};
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    w.host = true;
    return w;
  }
  switch (at.type) {
    case hipMemoryTypeDevice:
    case hipMemoryTypeManaged:
    case hipMemoryTypeArray:
    case hipMemoryTypeUnified:
      if (at.type == hipMemoryTypeDevice && at.device != device)
        fail(ERR_DEVICE, "an operand lives on device " + std::to_string(at.device) + ", the executor on device " +
                             std::to_string(device));
      w.dev = p;
      return w;
    case hipMemoryTypeHost:
      w.dev = at.devicePointer ? at.devicePointer : p;
      return w;
    default:
      w.host = true;
      return w;
  }
}

namespace {

bool is_zero(const double v[2]) { return v[0] == 0.0 && v[1] == 0.0; }

}  // namespace

void launch_plan(Plan& p, Executor& ex, ContractArgs& a, bool product, int ksplit, int64_t kslabs, hipStream_t s,
                 void* work) {
  if (!product) {
    hip_check(launch_output_pass(a, p.prob.dtype, s), "output-pass launch");
  } else if (p.strategy == 3) {
    hip_check(launch_elementwise(a, p.prob.dtype, s), "elementwise launch");
  } else if (p.cfg.precision == PREC_FP8) {
    std::lock_guard<std::mutex> g(ex.panel_mu);
    char* panels = ex.ensure_panels(fp8_panel_bytes(a, p.prob.dtype));
    hip_check(launch_contract_fp8(a, p.prob.dtype, panels, s), "FP8 contraction launch");
  } else if (p.cfg.precision == PREC_SPLIT_BF16_2 || p.cfg.precision == PREC_SPLIT_BF16_3) {
    const int terms = p.cfg.precision == PREC_SPLIT_BF16_2 ? 2 : 3;
    std::lock_guard<std::mutex> g(ex.panel_mu);
    char* panels = ex.ensure_panels(split_panel_bytes(a, p.prob.dtype, terms, p.tile));
    hip_check(launch_contract_split(a, p.prob.dtype, terms, p.tile, panels, s), "split-BF16 contraction launch");
  } else if (ksplit > 1) {
    std::unique_lock<std::mutex> g(ex.panel_mu, std::defer_lock);
    a.ksplit = ksplit;
    a.kslabs = kslabs;
    if (work) {
      a.work = work;
    } else {
      g.lock();
      a.work = ex.ensure_panels(splitk_work_bytes(p, ksplit, a.H));
    }
    hip_check(launch_contract(a, p.prob.dtype, p.cfg.complex_method, p.tile, s), "contraction launch");
    hip_check(launch_splitk_reduce(a, p.prob.dtype, s), "split-K reduction launch");
  } else {
    hip_check(launch_contract(a, p.prob.dtype, p.cfg.complex_method, p.tile, s), "contraction launch");
  }
}

size_t splitk_work_bytes(const Plan& p, int ksplit, int64_t H) {
  return (size_t)ksplit * (size_t)(p.M * p.N * H) * (size_t)elem_size(p.prob.dtype);
}

void validate_item(const Plan& p, const Item& it, bool beta_zero) {
  if (!it.a || !it.b || !it.d) fail(ERR_INVALID_ARGUMENT, "A, B or D data pointer is null or zero");
  if (!it.c && !beta_zero)
    fail(ERR_UNSUPPORTED,
         "a null C (TAPP_IN_PLACE) with a nonzero beta is ambiguous in TAPP and not supported; pass D as C to "
         "accumulate in place");
  if (p.prob.out_empty) return;
  Range rd = byte_range(p, OP_D, it.d);
  if (overlap(rd, byte_range(p, OP_A, it.a))) fail(ERR_ALIASED, "the output overlaps operand A");
  if (overlap(rd, byte_range(p, OP_B, it.b))) fail(ERR_ALIASED, "the output overlaps operand B");
  if (p.prob.c_separate && it.c) {
    bool same = it.c == it.d && p.prob.c_matches_d;
    if (!same && overlap(rd, byte_range(p, OP_C, it.c)))
      fail(ERR_ALIASED, "C and D overlap without being the same mapping at the same origin");
  }
}

void run_item(Plan& p, Executor& ex, const double alpha[2], const double beta_in[2], const Item& it, bool sync) {
  if (p.strategy == 2) return;
  double beta[2] = {beta_in[0], beta_in[1]};
  if (!it.c) beta[0] = beta[1] = 0.0;
  const bool read_c = !is_zero(beta);
  const bool product = (p.strategy == 0 || p.strategy == 3) && !is_zero(alpha);
  DeviceGuard dg(ex.device);
  hipStream_t s = ex.get_stream();
  const int64_t es = elem_size(p.prob.dtype);

  Where wa, wb, wc, wd;
  if (product) {
    wa = classify(it.a, ex.device);
    wb = classify(it.b, ex.device);
  }
  if (read_c) wc = classify(it.c, ex.device);
  wd = classify(it.d, ex.device);

  const bool staged = wa.host || wb.host || wc.host || wd.host;
  std::unique_lock<std::mutex> lock(ex.mu, std::defer_lock);
  struct Stage {
    int op;
    const void* host;
    size_t bytes;
    int64_t lo;
    size_t at;
  };
  std::vector<Stage> stages;
  if (staged) {
    lock.lock();
    size_t total = 0;
    auto plan_stage = [&](int op, const void* host) {
      const Span& sp = p.prob.spans[op];
      if (!sp.present) return;
      size_t bytes = (size_t)(sp.hi - sp.lo + 1) * es;
      stages.push_back({op, host, bytes, sp.lo, total});
      total += (bytes + 255) & ~(size_t)255;
    };
    if (wa.host) plan_stage(OP_A, it.a);
    if (wb.host) plan_stage(OP_B, it.b);
    const bool c_is_d = it.c == it.d;
    if (wc.host && !c_is_d) plan_stage(OP_C, it.c);
    if (wd.host) plan_stage(OP_D, it.d);
    char* base = ex.ensure_scratch(total);
    for (auto& st : stages) {
      char* dst = base + st.at;
      const char* src = (const char*)st.host + st.lo * es;
      hip_check(hipMemcpyAsync(dst, src, st.bytes, hipMemcpyHostToDevice, s), "staging copy to the device");
      const void* origin = dst - st.lo * es;
      switch (st.op) {
        case OP_A: wa.dev = origin; break;
        case OP_B: wb.dev = origin; break;
        case OP_C: wc.dev = origin; break;
        case OP_D: wd.dev = origin; break;
      }
    }
    if (wc.host && c_is_d) wc.dev = wd.dev;
  }

  ContractArgs a = p.args(ex.device, wa.dev, wb.dev, wc.dev, const_cast<void*>(wd.dev), alpha, beta, read_c);
  launch_plan(p, ex, a, product, p.ksplit, p.kslabs, s);

  if (staged) {
    for (auto& st : stages)
      if (st.op == OP_D) {
        char* dst = (char*)st.host + st.lo * es;
        hip_check(hipMemcpyAsync(dst, (const char*)wd.dev + st.lo * es, st.bytes, hipMemcpyDeviceToHost, s),
                  "staging copy to the host");
      }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  } else {
    ex.finish(s, sync);
  }
}

bool run_items_fused(Plan& p, Executor& ex, const double alpha[2], const double beta_in[2],
                     const std::vector<Item>& items) {
  const size_t n = items.size();
  if (n < 2 || p.strategy == 2) return false;
  double beta[2] = {beta_in[0], beta_in[1]};
  for (const Item& it : items)
    if (!it.c) beta[0] = beta[1] = 0.0;  // a null C needs beta 0 (validated), so C is never read
  const bool read_c = !is_zero(beta);
  const bool product = (p.strategy == 0 || p.strategy == 3) && !is_zero(alpha);
  if (product && p.cfg.precision != PREC_STORAGE) return false;  // per-call operand packing
  const int64_t es = elem_size(p.prob.dtype);
  const int64_t H = p.H;
  if ((int64_t)n > (INT64_MAX / 8) / std::max<int64_t>(H, 1)) return false;

  // Every operand the launch touches must be device memory of this device, and
  // every origin an element multiple away from the first item's.
  DeviceGuard dg(ex.device);
  // Device allocations already seen ([base, end) by base): one pointer query
  // per allocation instead of one per operand of every item.
  std::map<uintptr_t, uintptr_t> known;
  auto on_device = [&](const void* q) {
    const uintptr_t u = (uintptr_t)q;
    auto it = known.upper_bound(u);
    if (it != known.begin() && u < std::prev(it)->second) return true;
    Where w = classify(q, ex.device);
    if (w.host || w.dev != q) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)q) == hipSuccess && size > 0)
      known[(uintptr_t)base] = (uintptr_t)base + size;
    else
      (void)hipGetLastError();
    return true;
  };
  auto device_origin = [&](const void* q, const void* q0, int64_t& delta) {
    if (!on_device(q)) return false;
    const intptr_t d = (intptr_t)q - (intptr_t)q0;
    if (d % es != 0) return false;
    delta = d / es;
    return true;
  };
  std::vector<int64_t> da(n, 0), db(n, 0), dc(n, 0), dd(n, 0);
  bool a16 = true, b16 = true;
  for (size_t i = 0; i < n; i++) {
    const Item& it = items[i];
    if (product && !(device_origin(it.a, items[0].a, da[i]) && device_origin(it.b, items[0].b, db[i]))) return false;
    if (read_c && !device_origin(it.c, items[0].c, dc[i])) return false;
    if (!device_origin(it.d, items[0].d, dd[i])) return false;
    a16 = a16 && (uintptr_t)it.a % 16 == 0;
    b16 = b16 && (uintptr_t)it.b % 16 == 0;
  }

  // Running the items at once equals running them in order only if no item
  // writes bytes that another item reads or writes (span hulls, conservative).
  if (!p.prob.out_empty) {
    std::vector<ItemSpans> spans(n);
    for (size_t i = 0; i < n; i++) {
      spans[i].d = byte_range(p, OP_D, items[i].d);
      if (product) {
        spans[i].in[0] = byte_range(p, OP_A, items[i].a);
        spans[i].in[1] = byte_range(p, OP_B, items[i].b);
      }
      if (read_c) spans[i].in[2] = byte_range(p, OP_C, items[i].c);
    }
    if (!items_independent(spans)) return false;
  }

  // Batch offsets of item i, entry h: the item's origin delta plus the plan's
  // own offset of entry h, written to pinned memory and uploaded in stream
  // order.
  hipStream_t s = ex.get_stream();
  std::lock_guard<std::mutex> g(ex.batch_mu);
  const size_t bytes = 4 * n * H * sizeof(int64_t);
  int64_t* ha = (int64_t*)ex.ensure_batch_host(bytes);
  int64_t* hb = ha + n * H;
  int64_t* hc = hb + n * H;
  int64_t* hd = hc + n * H;
  for (size_t i = 0; i < n; i++)
    for (int64_t h = 0; h < H; h++) {
      const size_t e = i * H + h;
      ha[e] = da[i] + (p.ha.empty() ? 0 : p.ha[h]);
      hb[e] = db[i] + (p.hb.empty() ? 0 : p.hb[h]);
      hc[e] = dc[i] + (p.hc.empty() ? 0 : p.hc[h]);
      hd[e] = dd[i] + (p.hd.empty() ? 0 : p.hd[h]);
    }

  int64_t* dev = (int64_t*)ex.ensure_batch(bytes);
  hip_check(hipMemcpyAsync(dev, ha, bytes, hipMemcpyHostToDevice, s), "uploading batch offsets");
  hip_check(hipEventRecord(ex.batch_ev, s), "hipEventRecord");
  ex.batch_ev_pending = true;
  const Item& i0 = items[0];
  ContractArgs a = p.args(ex.device, product ? i0.a : nullptr, product ? i0.b : nullptr, read_c ? i0.c : nullptr,
                          i0.d, alpha, beta, read_c);
  a.H = (int64_t)n * H;
  a.ha = dev;
  a.hb = dev + n * H;
  a.hc = dev + 2 * n * H;
  a.hd = dev + 3 * n * H;
  a.h_affine = 0;
  a.a_kvec = a16 ? a.a_kvec : 0;
  a.b_kvec = b16 ? a.b_kvec : 0;
  a.ab_vec = a.ab_vec && a16 && b16;  // every item's origin, not only the first
  a.a_vec16 = a.a_vec16 && a16;
  // The items multiply the tile count: K is split only while the whole batch
  // would still leave the GPU underfilled.
  int ksplit = p.ksplit;
  if (ksplit > 1) {
    const TileCfg* t = tile_cfg(p.tile);
    const int64_t tiles = ((p.M + t->bm - 1) / t->bm) * ((p.N + t->bn - 1) / t->bn) * a.H;
    if (tiles >= 128) ksplit = 1;
  }
  launch_plan(p, ex, a, product, ksplit, p.kslabs, s);
  ex.finish(s, !ex.async.load());
  return true;
}

}  // namespace tp
