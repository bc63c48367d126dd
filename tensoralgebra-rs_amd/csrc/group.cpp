// Grouped execution: many independent contractions of one dtype, each with its
// own plan and shape, run as ONE persistent kernel over a device work list
// (the GPU counterpart of crates/tprims-contract/src/batch.rs:125
// contract_batched, which validates all items before running any, and of
// crates/tprims-capi/src/execute.rs:235 TAPP_execute_batched_product).
//
// The work list is one record per item (its kernel arguments and where its
// tiles start in the flattened tile space); workgroups stride over the tile
// space, so items of very different size share one launch. Items are ordered
// by descending K (the cost of one tile) so the longest tiles start first.
// The tile shape of the launch is chosen by measurement (autotune) among the
// storage-precision tiles; the launch can be captured once into a hipGraph and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "capi_util.h"
#include "executor.h"
#include "kernels.h"

extern "C" {
#include "tprims/tprims.h"
}

using namespace tp;

namespace {

struct Group {
  Executor* ex = nullptr;
  DType dtype = DType::F64;
  int method = 0;
  int tile = TILE_64x64;
  bool all_off32 = true;  // every item's gathers fit 32-bit offsets: the DMA tiles apply
  int grid = 0;
  int64_t total_tiles = 0;
  int64_t flops = 0;
  std::vector<Plan*> plans;          // held (Plan::retain) until the group is destroyed
  std::vector<ContractArgs> args;   // per contraction item, tile fields unset
  std::vector<int> own_tile;        // the tile each item's own plan chose
  std::vector<ContractArgs> passes; // items that only need the output pass
  GroupItem* d_items = nullptr;
  int n_items = 0;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t capture = nullptr;  // the legacy null stream cannot be captured
  // Output byte ranges of the items that read their own output (C is D,
  // beta != 0): repeated launches would accumulate there, so the autotuner
  // saves and restores them.
  std::vector<std::pair<uintptr_t, size_t>> inplace;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~Group() {
    drop_graph();
    if (capture) (void)hipStreamDestroy(capture);
    if (d_items) (void)hipFree(d_items);
    for (Plan* p : plans) Plan::release(p);
  }
};

void read_scalars(const void* p, DType t, int i, double out[2]) {
  out[0] = out[1] = 0;
  switch (t) {
    case DType::F32: out[0] = ((const float*)p)[i]; break;
    case DType::F64: out[0] = ((const double*)p)[i]; break;
    case DType::C32:
      out[0] = ((const float*)p)[2 * i];
      out[1] = ((const float*)p)[2 * i + 1];
      break;
    case DType::C64:
      out[0] = ((const double*)p)[2 * i];
      out[1] = ((const double*)p)[2 * i + 1];
      break;
  }
}

bool device_pointer(const void* p, int device) {
  if (!p) return true;
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (at.type == hipMemoryTypeDevice && at.device == device) || at.type == hipMemoryTypeManaged;
}

// Builds the device work list for the group's current tile.
void build_work_list(Group& g) {
  const TileCfg* t = tile_cfg(g.tile);
  std::vector<GroupItem> items;
  int64_t total = 0;
  for (const ContractArgs& a : g.args) {
    GroupItem it{};
    it.args = a;
    it.args.tiles_m = (int)((a.M + t->bm - 1) / t->bm);
    it.args.tiles_n = (int)((a.N + t->bn - 1) / t->bn);
    it.tile_begin = total;
    total += (int64_t)it.args.tiles_m * it.args.tiles_n * a.H;
    items.push_back(it);
  }
  g.drop_graph();
  if (g.d_items) {
    hip_check(hipFree(g.d_items), "hipFree(work list)");
    g.d_items = nullptr;
  }
  g.n_items = (int)items.size();
  g.total_tiles = total;
  if (!items.empty()) {
    hip_check(hipMalloc((void**)&g.d_items, items.size() * sizeof(GroupItem)), "hipMalloc(work list)");
    hip_check(hipMemcpy(g.d_items, items.data(), items.size() * sizeof(GroupItem), hipMemcpyHostToDevice),
              "uploading the work list");
  }
  // Persistent grid: enough workgroups to fill every CU a few times over.
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g.ex->device) == hipSuccess) cus = prop.multiProcessorCount;
  else (void)hipGetLastError();
  const int per_cu = g.tile == TILE_128x128 ? 1 : (t->threads >= 256 ? 2 : 8);
  g.grid = (int)std::max<int64_t>(1, std::min<int64_t>(total, (int64_t)cus * per_cu));
}

void launch_all(Group& g, hipStream_t s) {
  for (const ContractArgs& a : g.passes) hip_check(launch_output_pass(a, g.dtype, s), "output-pass launch");
  if (g.n_items == 0) return;
  if (g.all_off32 && dma_supported(true, g.dtype, g.method, g.tile))
    hip_check(launch_grouped_dma(g.d_items, g.n_items, g.total_tiles, g.dtype, g.method, g.tile, g.grid, s),
              "grouped launch");
  else
    hip_check(launch_grouped(g.d_items, g.n_items, g.total_tiles, g.dtype, g.method, g.tile, g.grid, s),
              "grouped launch");
}

void run_group(Group& g, Executor& ex) {
  DeviceGuard dg(ex.device);
  hipStream_t s = ex.get_stream();
  if (g.use_graph) {
    if (!g.gexec) {
      // Capture on a private stream (the executor's may be the null stream);
      // the graph is then launched on the executor's stream.
      if (!g.capture) hip_check(hipStreamCreateWithFlags(&g.capture, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipStreamBeginCapture(g.capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        launch_all(g, g.capture);
      } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(g.capture, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
      }
      hip_check(hipStreamEndCapture(g.capture, &g.graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&g.gexec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(g.gexec, s), "hipGraphLaunch");
  } else {
    launch_all(g, s);
  }
  ex.finish(s, !ex.async.load());
}

// Storage tiles a group can run: 128x64 is not offered for c64, the
// DMA-only 128x128 tile needs 32-bit gather offsets in every item and is not
// offered for c64 (any schedule) or the c32 1m schedule, whose grouped bodies
// do not fit the registers of a 16-wave workgroup (spills inside the K loop),
// and the K-group 32x32 tile has no grouped form.
bool group_tile_ok(const Group& g, int id) {
  if (id == TILE_128x128)
    return g.all_off32 && dma_supported(true, g.dtype, g.method, id) && g.dtype != DType::C64 &&
           !(g.dtype == DType::C32 && g.method == CPLX_1M);
  return tile_cfg(id) != nullptr && id < 10 && id != TILE_32x32K && id != TILE_32x32R && id != TILE_32x48R && !(id == TILE_128x64 && g.dtype == DType::C64);
}

Group* group_of(tprims_group h) {
  if (h == 0) fail(ERR_INVALID_ARGUMENT, "group is null or zero");
  return (Group*)h;
}

}  // namespace

extern "C" {

TAPP_error tprims_group_create(tprims_group* out, TAPP_executor exec, int num_items, const TAPP_tensor_product* plans,
                               const void* alpha, const void* const* A, const void* const* B, const void* beta,
                               const void* const* C, void* const* D) {
  return ffi([&] {
    if (!out) fail(ERR_INVALID_ARGUMENT, "group out-parameter is null");
    if (num_items < 0) fail(ERR_SHAPE, "negative item count");
    if (num_items > 0 && (!plans || !alpha || !beta || !A || !B || !D))
      fail(ERR_INVALID_ARGUMENT, "a plan, scalar or pointer array is null");
    std::unique_ptr<Group> g(new Group);
    // Validate every item, and the items against each other, before the GPU
    // runtime is touched.
    for (int i = 0; i < num_items; i++) {
      if (plans[i] == 0) fail(ERR_INVALID_ARGUMENT, "plan " + std::to_string(i) + " is null");
      Plan* p = (Plan*)plans[i];
      if (i == 0) {
        g->dtype = p->prob.dtype;
        g->method = p->cfg.complex_method;
      }
      if (p->prob.dtype != g->dtype) fail(ERR_DTYPE, "grouped items must share one dtype");
      if (p->cfg.complex_method != g->method) fail(ERR_UNSUPPORTED, "grouped items must share one complex method");
      if (p->cfg.precision != PREC_STORAGE) fail(ERR_UNSUPPORTED, "grouped items run at the storage precision");
      double al[2], be[2];
      read_scalars(alpha, g->dtype, i, al);
      read_scalars(beta, g->dtype, i, be);
      const void* c = C ? C[i] : nullptr;
      validate_item(*p, Item{A[i], B[i], c, D[i]}, be[0] == 0 && be[1] == 0);
    }
    // The items run at the same time: none may write bytes another item reads
    // or writes (the reference's contract_batched gets this from disjoint
    // mutable borrows, crates/tprims-contract/src/batch.rs:125).
    std::vector<ItemSpans> spans(num_items);
    for (int i = 0; i < num_items; i++) {
      const Plan* p = (const Plan*)plans[i];
      if (p->prob.out_empty) continue;
      double al[2], be[2];
      read_scalars(alpha, g->dtype, i, al);
      read_scalars(beta, g->dtype, i, be);
      const void* c = C ? C[i] : nullptr;
      const bool product = (p->strategy == 0 || p->strategy == 3) && (al[0] != 0 || al[1] != 0);
      const bool read_c = c && (be[0] != 0 || be[1] != 0);
      spans[i].d = byte_range(*p, OP_D, D[i]);
      if (product) {
        spans[i].in[0] = byte_range(*p, OP_A, A[i]);
        spans[i].in[1] = byte_range(*p, OP_B, B[i]);
      }
      if (read_c) {
        spans[i].in[2] = byte_range(*p, OP_C, c);
        if (overlap(spans[i].d, spans[i].in[2]))
          g->inplace.push_back({(uintptr_t)spans[i].d.lo, (size_t)(spans[i].d.hi - spans[i].d.lo)});
      }
    }
    if (!items_independent(spans))
      fail(ERR_ALIASED, "grouped items overlap: an item's output meets another item's operands or output");
    Executor* ex = resolve_executor(exec);
    g->ex = ex;
    for (int i = 0; i < num_items; i++)
      for (const void* q : {A[i], B[i], C ? C[i] : nullptr, (const void*)D[i]})
        if (!device_pointer(q, ex->device)) fail(ERR_DEVICE, "grouped operands must be device memory of the executor");
    DeviceGuard dg(ex->device);
    std::vector<std::pair<int64_t, int>> order;
    for (int i = 0; i < num_items; i++) order.push_back({((Plan*)plans[i])->K, i});
    std::stable_sort(order.begin(), order.end(), [](auto& x, auto& y) { return x.first > y.first; });
    for (auto& [k, i] : order) {
      Plan* p = (Plan*)plans[i];
      p->retain();
      g->plans.push_back(p);
      if (p->strategy == 2) continue;
      double al[2], be[2];
      read_scalars(alpha, g->dtype, i, al);
      read_scalars(beta, g->dtype, i, be);
      const void* c = C ? C[i] : nullptr;
      if (!c) be[0] = be[1] = 0;
      const bool read_c = be[0] != 0 || be[1] != 0;
      ContractArgs a = p->args(ex->device, A[i], B[i], c, D[i], al, be, read_c);
      if (p->strategy == 1 || (al[0] == 0 && al[1] == 0)) {
        g->passes.push_back(a);
      } else {
        g->all_off32 = g->all_off32 && a.off32;
        g->args.push_back(a);
        g->own_tile.push_back(p->tile);
        g->flops += p->flops();
      }
    }
    // Default tile: the one most items' plans chose (the DMA-only 128x128 tile
    // when every item can take it).
    int votes[16] = {0};
    for (Plan* p : g->plans)
      if (p->tile > 0 && p->tile <= TILE_128x128) votes[p->tile]++;
    g->tile = (int)(std::max_element(votes, votes + TILE_128x128 + 1) - votes);
    if (g->tile == 0 || !group_tile_ok(*g, g->tile)) g->tile = TILE_64x64;
    build_work_list(*g);
    *out = (tprims_group)g.release();
  });
}

TAPP_error tprims_group_execute(tprims_group group, TAPP_executor exec) {
  return ffi([&] {
    Group* g = group_of(group);
    Executor* ex = exec ? (Executor*)exec : g->ex;
    if (ex->device != g->ex->device) fail(ERR_DEVICE, "the group was built for another device");
    ex->inflight++;
    struct Release {
      Executor* e;
      ~Release() { e->inflight--; }
    } rel{ex};
    run_group(*g, *ex);
  });
}

TAPP_error tprims_group_use_graph(tprims_group group, TAPP_executor, int enable) {
  return ffi([&] {
    Group* g = group_of(group);
    g->use_graph = enable != 0;
    if (!g->use_graph) g->drop_graph();
  });
}

TAPP_error tprims_group_execute_each(tprims_group group, TAPP_executor exec) {
  return ffi([&] {
    Group* g = group_of(group);
    Executor* ex = exec ? (Executor*)exec : g->ex;
    DeviceGuard dg(ex->device);
    hipStream_t s = ex->get_stream();
    for (const ContractArgs& a : g->passes) hip_check(launch_output_pass(a, g->dtype, s), "output-pass launch");
    // Each item on the tile its own plan chose.
    for (size_t i = 0; i < g->args.size(); i++)
      hip_check(launch_contract(g->args[i], g->dtype, g->method, g->own_tile[i], s), "contraction launch");
    ex->finish(s, !ex->async.load());
  });
}

TAPP_error tprims_group_info(tprims_group group, int64_t* tiles, int64_t* flops, int32_t* tile_id) {
  return ffi([&] {
    Group* g = group_of(group);
    if (tiles) *tiles = g->total_tiles;
    if (flops) *flops = g->flops;
    if (tile_id) *tile_id = g->tile;
  });
}

TAPP_error tprims_group_set_tile(tprims_group group, int32_t tile_id) {
  return ffi([&] {
    Group* g = group_of(group);
    if (!group_tile_ok(*g, tile_id)) fail(ERR_UNSUPPORTED, "tile id not available for grouped execution");
    DeviceGuard dg(g->ex->device);
    g->tile = tile_id;
    build_work_list(*g);
  });
}

TAPP_error tprims_group_autotune(tprims_group group, TAPP_executor exec, int reps, int32_t* best_tile,
                                 double* best_ms) {
  return ffi([&] {
    Group* g = group_of(group);
    Executor* ex = exec ? (Executor*)exec : g->ex;
    if (reps < 1) reps = 1;
    DeviceGuard dg(ex->device);
    hipStream_t s = ex->get_stream();
    const bool graph = g->use_graph;
    g->use_graph = false;
    // Items that accumulate into their own output are saved first and
    // restored after the timing launches, so that on return every output
    // holds the result of exactly one execution.
    size_t saved_bytes = 0;
    for (auto& r : g->inplace) saved_bytes += (r.second + 255) & ~(size_t)255;
    char* saved = nullptr;
    struct FreeSaved {
      char*& p;
      ~FreeSaved() {
        if (p) (void)hipFree(p);
      }
    } free_saved{saved};
    if (saved_bytes) {
      hip_check(hipMalloc((void**)&saved, saved_bytes), "hipMalloc(autotune snapshot)");
      size_t at = 0;
      for (auto& r : g->inplace) {
        hip_check(hipMemcpyAsync(saved + at, (const void*)r.first, r.second, hipMemcpyDeviceToDevice, s),
                  "saving an in-place output");
        at += (r.second + 255) & ~(size_t)255;
      }
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int best = g->tile;
    float best_t = 1e30f;
    for (int id : {TILE_128x128, TILE_64x64, TILE_32x32, TILE_128x64}) {
      if (!group_tile_ok(*g, id)) continue;
      g->tile = id;
      build_work_list(*g);
      launch_all(*g, s);  // warm-up
      hip_check(hipEventRecord(e0, s), "hipEventRecord");
      for (int r = 0; r < reps; r++) launch_all(*g, s);
      hip_check(hipEventRecord(e1, s), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      ms /= reps;
      if (ms < best_t) {
        best_t = ms;
        best = id;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->tile = best;
    build_work_list(*g);
    g->use_graph = graph;
    // One execution on the chosen tile, from the saved in-place outputs.
    if (saved_bytes) {
      size_t at = 0;
      for (auto& r : g->inplace) {
        hip_check(hipMemcpyAsync((void*)r.first, saved + at, r.second, hipMemcpyDeviceToDevice, s),
                  "restoring an in-place output");
        at += (r.second + 255) & ~(size_t)255;
      }
    }
    launch_all(*g, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (best_tile) *best_tile = best;
    if (best_ms) *best_ms = best_t;
  });
}

TAPP_error tprims_group_destroy(tprims_group group) {
  return ffi([&] {
    Group* g = group_of(group);
    delete g;
  });
}

}  // extern "C"
