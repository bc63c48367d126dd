// Host-side entry points of the HIP kernels (implemented in *.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "kargs.h"

namespace tp {

// A workgroup tile of the contraction kernels.
struct TileCfg {
  int id;       // stable id (plan reports, configuration override)
  int bm, bn, bk;
  int threads;
  const char* name;
};

// Tile ids of the storage-precision kernels (16x16x4 MFMA family).
enum TileId : int {
  TILE_64x64 = 1,  // 4 waves of 32x32
  TILE_32x32 = 2,  // 1 wave of 32x32
  TILE_128x64 = 3, // 4 waves of 64x32
  TILE_128x128 = 4, // LDS-DMA kernel only (16 waves of 32x32, or 8 of 64x32: by dtype, schedule, majors)
  TILE_32x32K = 5,  // LDS-DMA kernel only: 4 waves of 32x32, each over every fourth K slab
  TILE_32x32R = 6,  // f64 in 16-byte runs: 8 waves of 32x32, each over every eighth K slab, loading into registers
  TILE_32x48R = 7,  // the same on 32x48 tiles (B gathered along K)
};
// Tile ids of the split-BF16 kernels (16x16x32 bf16 MFMA family).
enum SplitTileId : int {
  STILE_128x128 = 11,  // 4 waves of 64x64 (f32)
  STILE_64x64 = 12,    // 4 waves of 32x32
  STILE_256x128 = 13,  // 8 waves of 64x64 (f32)
  STILE_128x64 = 14,   // 8 waves of 32x32 (c32)
  STILE_256x256 = 15,  // FP8 f32: 8 waves of 128x64
};

const TileCfg* tile_cfg(int id);

// Largest workgroup count of a one-dimensional launch of `threads`-wide
// workgroups (work-items below 2^32), rounded down to a multiple of 8 XCDs.
inline int64_t max_grid(int threads) { return (int64_t)((0xffffffffull / (unsigned)threads) & ~7ull); }

// The contraction kernel of a storage dtype, complex method and tile.
// Returns hipSuccess or the launch error.
hipError_t launch_contract(const ContractArgs& a, DType t, int method, int tile_id, hipStream_t s);
// The LDS-DMA variant of the storage kernels (every dtype and complex method,
// tiles 64x64, 128x64 (not c64) and 128x128; 32-bit gather offsets), which
// launch_contract picks whenever it applies, split K included.
bool dma_supported(bool off32, DType t, int method, int tile_id);
hipError_t launch_contract_dma(const ContractArgs& a, DType t, int method, int tile_id, hipStream_t s);
// The 32x32 tile on four K groups (TILE_32x32K): 32-bit gather offsets, one
// K range (no split K), and K-major A and B unless both gather in 16-byte
// runs of f64 or c32 elements (any majors) or f32 elements (K-major).
bool kgroup_supported(bool off32, bool ab_vec, bool a_kmajor, bool b_kmajor, DType t);
hipError_t launch_contract_kgroup(const ContractArgs& a, DType t, int method, hipStream_t s);
// The 32x32 tile on eight K groups with operands loaded straight into
// registers (TILE_32x32R; wide: 32x48, TILE_32x48R, B gathered along K):
// f64, 32-bit gather offsets, A and B in 16-byte runs, one K range.
bool direct_supported(bool off32, bool ab_vec, DType t);
hipError_t launch_contract_direct(const ContractArgs& a, DType t, bool wide, hipStream_t s);
// Split-BF16 for f32/c32 storage (nterms 2 or 3): A and B are split once into
// bf16 fragment panels (device scratch of split_panel_bytes), then multiplied.
size_t split_panel_bytes(const ContractArgs& a, DType t, int nterms, int tile_id);
// Whether split-BF16 splits A inside the GEMM rather than packing A panels
// (with 32-bit gather offsets and a 16-byte aligned A).
bool split_a_in_gemm(bool cplx, bool a_kmajor, bool a_vec16, int64_t N, int tile_id);
hipError_t launch_contract_split(const ContractArgs& a, DType t, int nterms, int tile_id, void* panels,
                                 hipStream_t s);
// FP8 (e4m3, one power-of-two scale per 128-k panel of every row) for f32/c32
// storage: packs A and B into fp8 fragment panels plus scale words (device
// scratch of fp8_panel_bytes), then multiplies on the scaled fp8 MFMA. Tiles:
// 256x256 (f32), 128x64 (c32).
size_t fp8_panel_bytes(const ContractArgs& a, DType t);
hipError_t launch_contract_fp8(const ContractArgs& a, DType t, void* panels, hipStream_t s);
// The elementwise (all-batch) strategy: one thread per output element, K
// summed serially; for tiny M x N and K with many batch entries.
hipError_t launch_elementwise(const ContractArgs& a, DType t, hipStream_t s);
// Split-K reduction: sums the a.ksplit partial tiles in a.work in split order
// and writes D = op_D(alpha * sum + beta * op_C(C)).
hipError_t launch_splitk_reduce(const ContractArgs& a, DType t, hipStream_t s);
// D = op_D(beta * op_C(C)) over the output (empty K or alpha == 0).
hipError_t launch_output_pass(const ContractArgs& a, DType t, hipStream_t s);
// Zero every element a layout addresses (rank <= 8; used by the adjoints).
hipError_t launch_zero_layout(void* base, DType t, int rank, const int64_t* dims, const int64_t* strides,
                              hipStream_t s);
// Persistent grouped kernel over `items` (device array) covering `total_tiles`.
hipError_t launch_grouped(const GroupItem* items, int num_items, int64_t total_tiles, DType t, int method,
                          int tile_id, int grid, hipStream_t s);
// The same on the LDS-DMA tiles (every item with 32-bit gather offsets and a
// tile dma_supported accepts).
hipError_t launch_grouped_dma(const GroupItem* items, int num_items, int64_t total_tiles, DType t, int method,
                              int tile_id, int grid, hipStream_t s);

}  // namespace tp
