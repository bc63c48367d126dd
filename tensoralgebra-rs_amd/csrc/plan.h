// The GPU plan of one contraction: folded role axes, scatter vectors, gather
// orientation, K order and tile choice. Everything here depends only on shapes,
// strides and labels (crates/tprims-contract/src/plan/analysis.rs:210 PackedPlan,
// crates/tprims-contract/src/plan/orientation.rs, crates/tprims-contract/src/resolve.rs:42).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kargs.h"
#include "problem.h"

namespace tp {

// One folded axis (analysis.rs:63-76): its extent and strides in A, B, C, D.
struct Axis {
  int64_t extent;
  int64_t sa, sb, sc, sd;
};

struct PlanConfig {
  int32_t complex_method = CPLX_NATIVE;
  int32_t precision = PREC_STORAGE;
  int32_t tile = 0;  // 0: chosen by the planner
  int32_t split_k = 0;  // 0: chosen by the planner, 1: never, n > 1: n K ranges
};

// Scatter vectors of a plan resident on one device (one allocation).
struct DeviceScatter {
  int64_t* base = nullptr;
  const int64_t *am, *ak, *bk, *bn, *cm, *cn, *dm, *dn, *ha, *hb, *hc, *hd, *dr;
  const uint32_t *ak32, *bk32;
};

struct Plan {
  Problem prob;
  PlanConfig cfg;
  std::vector<Axis> m_ax, n_ax, k_ax, h_ax, r_ax;
  int64_t M = 1, N = 1, K = 1, H = 1, R = 1;
  std::vector<int64_t> am, ak, bk, bn, cm, cn, dm, dn, ha, hb, hc, hd, dr;
  bool a_kmajor = true, b_kmajor = true, d_mmajor = true;
  // Host half of ContractArgs::a_kvec / b_kvec (the pointer alignment is
  // checked per call).
  int a_kvec = 0, b_kvec = 0;
  bool off32 = false;  // ContractArgs::off32
  bool ab_vec = false;  // host half of ContractArgs::ab_vec (pointer alignment checked per call)
  bool a_vec16 = false;  // host half of ContractArgs::a_vec16
  // Split K (ContractArgs::ksplit, kslabs): 1, or the number of K ranges a
  // tile's work is divided into when the tiles alone would leave most of the
  // GPU idle; the partials are reduced in a fixed order.
  int ksplit = 1;
  int64_t kslabs = 0;
  // With off32: the K scatters of A and B as byte offsets, padded with the
  // out-of-range sentinel to a multiple of 64 entries (ContractArgs::ak32).
  std::vector<uint32_t> ak32, bk32;
  int tile = 0;
  // Where the tile and K split came from: 0 the planner's cost model, 1 the
  // configuration, 2 the tuning cache, 3 autotuning of this plan.
  int source = 0;
  // 0 contraction (MFMA tiles), 1 output pass only (K empty), 2 output empty,
  // 3 elementwise (all-batch: tiny M x N and K, many batch entries)
  int strategy = 0;

  // Device copies, uploaded on first use per device.
  std::mutex mu;
  std::map<int, DeviceScatter> dev;
  std::vector<int64_t> host_image;  // host_tables()

  // Owners: the TAPP handle and every group built on the plan. The plan is
  // deleted when the last of them lets go (release).
  std::atomic<int> refs{1};
  void retain() { refs.fetch_add(1); }
  static void release(Plan* p) {
    if (p && p->refs.fetch_sub(1) == 1) delete p;
  }

  // Adjoint plans (cotangents of A and B), built on the first VJP.
  Plan* adj[2] = {nullptr, nullptr};
  bool adj_zero_fill[2] = {false, false};

  int64_t flops() const { return (is_complex(prob.dtype) ? 8 : 2) * M * N * K * H; }
  const DeviceScatter& on_device(int device);
  // The scatter vectors as one int64 image (the layout on_device uploads):
  // the 13 vectors back to back, then the 32-bit K offsets of A and B.
  std::vector<int64_t> flat_tables() const;
  // The same layout's pointers into an image at `base` (device memory).
  DeviceScatter tables_at(int64_t* base) const;
  // flat_tables(), built once and kept with the plan (a stable host source
  // for copies into caller-provided device memory).
  const std::vector<int64_t>& host_tables();
  // Kernel arguments for data pointers (device addresses) and scalars, with
  // the scatter vectors of this device (uploaded on first use) or of `s`.
  ContractArgs args(int device, const void* A, const void* B, const void* C, void* D, const double alpha[2],
                    const double beta[2], bool read_c);
  ContractArgs args_with(const DeviceScatter& s, const void* A, const void* B, const void* C, void* D,
                         const double alpha[2], const double beta[2], bool read_c) const;
  std::string kernel_name() const;
  ~Plan();
};

// Builds the plan of a validated problem.
Plan* make_plan(const Problem& p, const PlanConfig& cfg);
// Re-runs the tile choice and K ordering for a new configuration. With no
// tile and no K split asked for, a tuning cache entry for the plan's reduced
// problem supplies them.
void configure(Plan* plan, const PlanConfig& cfg);
// Whether tile `id` exists for this dtype, precision and complex method.
bool tile_allowed(int id, DType t, int prec, int method);

// The tuning cache (tune.cpp): the tile and K split configuration that
// autotuning chose for a reduced problem, keyed by tuning_key (dtype,
// complex method, precision, strategy and every folded axis with its
// strides; the tile and split of `cfg` are not part of it).
std::string tuning_key(const Plan& p, const PlanConfig& cfg);
bool tuning_lookup(const std::string& key, int* tile, int* split_k);
void tuning_store(const std::string& key, int tile, int split_k);

// The scatter vector of a mixed-radix index group, fastest axis first
// (crates/tprims-kernel/src/pack/scatter.rs:70).
std::vector<int64_t> build_scatter(const std::vector<int64_t>& extents, const std::vector<int64_t>& strides);
// The block scatter of scat[0, len) at block size blk > 0: per block of blk
// consecutive entries (the last one ragged) the common difference of the run,
// 0 for a run of one entry, or SCATTER_IRREGULAR when the run is not an
// arithmetic progression (crates/tprims-kernel/src/pack/scatter.rs:130-176).
constexpr int64_t SCATTER_IRREGULAR = INT64_MIN;
std::vector<int64_t> build_block_scatter(const int64_t* scat, int64_t len, int64_t blk);
// Folding of adjacent compatible axes (analysis.rs:299-317).
std::vector<Axis> fold_axes(const std::vector<Axis>& axes);

}  // namespace tp
