// Kernel-argument records shared by the host planner and the HIP kernels.
// Everything is plain data: device pointers, extents and flags.
#pragma once
#include <cstdint>

namespace tp {

// One contraction as a batched matrix product over scatter vectors
// (the GPU form of crates/tprims-contract/src/plan/analysis.rs PackedPlan):
//   D[hd[h] + dm[m] + dn[n] + dr[r]] =
//     op_D(alpha * sum_k op_A(A[ha[h] + am[m] + ak[k]]) * op_B(B[hb[h] + bk[k] + bn[n]])
//          + beta * op_C(C[hc[h] + cm[m] + cn[n]]))      for every r in [0, R)
// Offsets are in elements of the storage type (complex elements for c32/c64).
struct ContractArgs {
  const void* A;
  const void* B;
  const void* C;
  void* D;
  const int64_t* am;
  const int64_t* ak;
  const int64_t* bk;
  const int64_t* bn;
  const int64_t* cm;
  const int64_t* cn;
  const int64_t* dm;
  const int64_t* dn;
  const int64_t* ha;
  const int64_t* hb;
  const int64_t* hc;
  const int64_t* hd;
  const int64_t* dr;
  int64_t M, N, K, H, R;
  double alpha[2];
  double beta[2];
  int32_t conj_a, conj_b, conj_c, conj_d;
  int32_t a_kmajor;  // A gathers run along K (1) or along M (0)
  int32_t b_kmajor;  // B gathers run along K (1) or along N (0)
  int32_t d_mmajor;  // the write-back runs along M (1) or along N (0)
  int32_t read_c;    // beta != 0: C is read (never read otherwise)
  int32_t tiles_m, tiles_n;
  // K % 8 == 0, every row and batch offset of A (B) is 16-byte aligned and
  // the K scatter runs contiguously over each group of 8 k starting at a
  // multiple of 8, from a 16-byte aligned offset: the 8 k of a group of one
  // row load as whole 16-byte vectors (split-BF16 and FP8 packers). 2: the K
  // scatter is the identity (offsets computed, no K table read); 1: group
  // starts read from the K table; 0: neither.
  int32_t a_kvec, b_kvec;
  // Every gather offset of A and of B (rows plus k, in bytes from the batch
  // entry's base) is nonnegative and below 2^30: gathers use 32-bit buffer
  // offsets.
  int32_t off32;
  // f32, f64, c32 with off32: A and B gather in 16-byte runs. Along its major
  // index (K, or M for A / N for B) each operand's scatter consists of runs
  // of 16 / es consecutive elements starting at multiples of 16 / es, its
  // extent there is a multiple of 16 / es, and every run start, k offset,
  // row offset and batch offset is a multiple of 16 bytes, as are A and B.
  int32_t ab_vec;
  // A alone gathers in such 16-byte runs (off32, f32/f64/c32; its base
  // 16-byte aligned): what the split-BF16 kernel's in-kernel split of an
  // M-major A needs.
  int32_t a_vec16;
  // With off32: the K scatters as 32-bit byte offsets, entries from K up to a
  // multiple of 64 holding the out-of-range sentinel 0x40000000.
  const uint32_t* ak32;
  const uint32_t* bk32;
  // The batch role is one folded axis: batch offsets are h * h_stride[A, B,
  // C, D] (used by the elementwise kernel instead of the scatter vectors).
  int64_t h_stride[4];
  int32_t h_affine;
  // The M and N roles are each at most one folded axis: C and D offsets are
  // m * cd_stride[0] + n * cd_stride[1] (C) and m * cd_stride[2] + n *
  // cd_stride[3] (D), with no scatter-vector loads in the write-back.
  int32_t mn_affine;
  int64_t cd_stride[4];
  // Split K: ksplit workgroups per output tile, each over kslabs K slabs,
  // writing its unscaled partial tile to work[split][h][m][n] (storage type);
  // a reduction pass sums the splits in order and does the write-back.
  int32_t ksplit;
  int64_t kslabs;
  void* work;
};

// One item of a grouped launch: the contraction plus where its tiles start in
// the group's flattened tile space.
struct GroupItem {
  ContractArgs args;
  int64_t tile_begin;  // first global tile index of this item
};

}  // namespace tp
