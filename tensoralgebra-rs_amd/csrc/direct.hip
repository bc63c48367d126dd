// Register-direct K-group tile for f64 products with small outputs and a long
// K (crates/tprims-contract/src/plan/analysis.rs PackedPlan executed as one
// 32 x 32 tile per workgroup): every wave is a K group that computes the
// whole 32 x 32 tile over slabs g, g + KG, g + 2 KG, ... (16 k each), with
// its operands loaded straight from memory into the MFMA's registers as
// 16-byte buffer loads, DEPTH slabs in flight. The K loop has no LDS staging
// and no barrier, so a wave's only waits are for its own loads; the groups
// meet once, to add their partial tiles through LDS in group order (the sum
// order is fixed, so repeated calls are bitwise equal) before the fused
// write-back.
//
// Operands gather in 16-byte runs (ContractArgs::ab_vec): pairs of f64
// consecutive along the operand's major index, every pair start, row, k and
// batch offset a multiple of 16 bytes. A K-major pair is two k of one row,
// an M-major pair two rows at one k. v_mfma_f64_16x16x4f64 takes row (or
// column) lane % 16 and k lane / 16; the 16 k of a slab are assigned to its
// four MFMA steps s as k(s, q) = 8 (s / 2) + 2 q + s % 2 for lane group
// q = lane / 16, so the two k of a K-major pair feed steps 2p and 2p + 1 of
// one row. The two rows of an M-major pair feed the tile's two row blocks:
// row block i, lane row r holds tile row 2 r + i (columns alike for an
// N-major B), which the staging for the write-back undoes.
//
// The K offsets of a slab are the 16 consecutive entries of the padded
// 32-bit K scatter (ContractArgs::ak32 / bk32) at its first k: one scalar
// load per operand (the scalar cache and its counter, so the waits for them
// do not drain the vector loads in flight), from which each lane picks its
// entries.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

typedef double v2d __attribute__((ext_vector_type(2)));
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

// 16 consecutive entries of a K scatter at entry k0, as scalar loads (the
// index is uniform across the wave).
__device__ inline u32x16 koff16(const uint32_t* tab, int64_t k0) {
  return *(const __attribute__((address_space(4))) u32x16*)(tab + k0);
}

__device__ inline v2d load16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

// NB: column blocks of 16 (BN = 16 NB); three only with B gathered along K
// (an N-major pair feeds two blocks). MINW: waves per SIMD the registers must
// allow (KG / 2: two workgroups per CU; KG / 4: one).
template <bool AKM_, bool BKM_, int KG_, int DEPTH_, int NB_, int MINW_>
struct RCfg {
  static constexpr bool AKM = AKM_, BKM = BKM_;
  static constexpr int KG = KG_, DEPTH = DEPTH_, NB = NB_, MINW = MINW_;
  static constexpr int BM = 32, BN = 16 * NB, BK = 16, LDD = BN + 1;
  static constexpr int NT = 64 * KG;
  static_assert(NB == 2 || (NB == 3 && BKM), "three column blocks need B gathered along K");
};

// One slab's operands as a lane holds them: four 16-byte loads of A and
// 2 NB of B (K-major: [2 i + p] = rows of block i, k pair p; M/N-major:
// [s] = the row pair of step s).
template <int NB>
struct Slab {
  v2d a[4], b[2 * NB];
};

template <class C>
__global__ __launch_bounds__(C::NT, C::MINW) void k_contract_direct(ContractArgs p) {
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, KG = C::KG, DEPTH = C::DEPTH, LDD = C::LDD;
  __shared__ __attribute__((aligned(16))) double red[KG * BM * LDD];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  int64_t h;
  int tm, tn;
  dev::decode_tile((int64_t)blockIdx.x, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int kg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t M = p.M, N = p.N, K = p.K;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const __amdgpu_buffer_rsrc_t rsA = dev::make_rsrc((const double*)p.A + p.ha[h]);
  const __amdgpu_buffer_rsrc_t rsB = dev::make_rsrc((const double*)p.B + p.hb[h]);

  // Row byte offsets: K-major, the lane's row of each block; M/N-major, the
  // first row of the lane's pair (the extent is even, so a pair is wholly in
  // or out of range).
  constexpr int NB = C::NB;
  uint32_t a_row[2], b_row[NB];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const int64_t m = C::AKM ? m0 + 16 * i + lr : m0 + 2 * lr;
    a_row[i] = m < M ? (uint32_t)(p.am[m] * 8) : dev::ROW_OOB;
  }
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int64_t n = C::BKM ? n0 + 16 * j + lr : n0 + 2 * lr;
    b_row[j] = n < N ? (uint32_t)(p.bn[n] * 8) : dev::ROW_OOB;
  }

  // Entry base + 2 q of a slab's K offsets for lane group q, as masks (a
  // select chain on the lane group would become branches).
  const uint32_t q0 = lq == 0 ? ~0u : 0u, q1 = lq == 1 ? ~0u : 0u, q2 = lq == 2 ? ~0u : 0u, q3 = lq == 3 ? ~0u : 0u;
  auto pick = [&](const u32x16& v, int base) -> uint32_t {
    return (v[base] & q0) | (v[base + 2] & q1) | (v[base + 4] & q2) | (v[base + 6] & q3);
  };
  struct KOff {
    u32x16 a, b;
  };
  // This group's slabs: gs = kg + KG t for t in [0, nk). The K offsets of
  // slab t (slab 0's for t past the group's last, so the table read stays in
  // the padded vector).
  const int64_t nslabs = (K + BK - 1) / BK;
  const int64_t nk = kg < nslabs ? (nslabs - kg + KG - 1) / KG : 0;
  auto koffs = [&](int64_t t) {
    const int64_t k0 = t < nk ? (kg + (int64_t)KG * t) * BK : 0;
    return KOff{koff16(p.ak32, k0), koff16(p.bk32, k0)};
  };
  // The loads of a slab from its K offsets; a slab past the group's last
  // loads zeros from the out-of-range sentinel (no memory access), so every
  // path through the loop issues the same loads and the waits stay exact.
  auto issue = [&](const KOff& ko, bool valid, Slab<NB>& sl) {
    const uint32_t oob = valid ? 0u : dev::K_OOB;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if constexpr (C::AKM) sl.a[t] = load16(rsA, a_row[t >> 1] + (pick(ko.a, 8 * (t & 1)) | oob));
      else sl.a[t] = load16(rsA, a_row[0] + (pick(ko.a, 8 * (t >> 1) + (t & 1)) | oob));
    }
#pragma unroll
    for (int t = 0; t < 2 * NB; t++) {
      if constexpr (C::BKM) sl.b[t] = load16(rsB, b_row[t >> 1] + (pick(ko.b, 8 * (t & 1)) | oob));
      else sl.b[t] = load16(rsB, b_row[0] + (pick(ko.b, 8 * (t >> 1) + (t & 1)) | oob));
    }
  };

  dev::v4d acc[2][NB];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < NB; j++) acc[i][j] = dev::v4d{0.0, 0.0, 0.0, 0.0};
  auto compute = [&](const Slab<NB>& sl) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      double av[2], bv[NB];
#pragma unroll
      for (int i = 0; i < 2; i++) av[i] = C::AKM ? sl.a[2 * i + (s >> 1)][s & 1] : sl.a[s][i];
#pragma unroll
      for (int j = 0; j < NB; j++) bv[j] = C::BKM ? sl.b[2 * j + (s >> 1)][s & 1] : sl.b[s][j];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  };

  // A ring of DEPTH + 1 slabs: each step issues the loads of slab t + DEPTH
  // into the slot freed by slab t - 1 and the scalar load of the next K
  // offsets, then runs the MFMAs of slab t, whose loads were issued DEPTH
  // steps earlier. The steady-state loop has no branch inside (every path
  // issues the same loads in the same order, so the compiler's waits count
  // exactly the loads still in flight); the last slabs, already loaded, run
  // after it.
  constexpr int RING = DEPTH + 1;
  Slab<NB> buf[RING];
  KOff ko = koffs(0);
#pragma unroll
  for (int d = 0; d < DEPTH; d++) {
    issue(ko, d < nk, buf[d]);
    ko = koffs(d + 1);
  }
  int64_t t0 = 0;
  for (; t0 + RING <= nk; t0 += RING) {
#pragma unroll
    for (int d = 0; d < RING; d++) {
      const int64_t t = t0 + d;
      issue(ko, t + DEPTH < nk, buf[(d + DEPTH) % RING]);
      ko = koffs(t + DEPTH + 1);
      compute(buf[d]);
    }
  }
  // At most DEPTH slabs remain, all in flight, from ring slot 0 on.
#pragma unroll
  for (int d = 0; d < DEPTH; d++)
    if (t0 + d < nk) compute(buf[d]);

  // Partial tiles into LDS (group g at red + g BM LDD), undoing the row and
  // column pairing; the groups' sums in group order into group 0's part.
  double* mine = red + kg * BM * LDD;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = lq + 4 * q;  // the output row within the block
        const int row = C::AKM ? 16 * i + r : 2 * r + i;
        const int col = C::BKM ? 16 * j + lr : 2 * lr + j;
        mine[row * LDD + col] = acc[i][j][q];
      }
  __syncthreads();
  if constexpr (KG > 1) {
    for (int e = threadIdx.x; e < BM * BN; e += C::NT) {
      const int o = (e / BN) * LDD + e % BN;
      double v = red[o];
#pragma unroll
      for (int g = 1; g < KG; g++) v += red[g * BM * LDD + o];
      red[o] = v;
    }
    __syncthreads();
  }
  dev::write_back<double, double, false, BM, BN, C::NT, LDD, true>(p, h, m0, n0, red, nullptr);
}

template <class C>
hipError_t launch_rcfg(const ContractArgs& a, hipStream_t s) {
  ContractArgs p = a;
  p.tiles_m = (int)((p.M + C::BM - 1) / C::BM);
  p.tiles_n = (int)((p.N + C::BN - 1) / C::BN);
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  if (nwg == 0) return hipSuccess;
  if (nwg > max_grid(C::NT)) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_contract_direct<C>), dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
  return hipGetLastError();
}

// 32 x 32: two workgroups per CU, one slab in flight per wave; 32 x 48 (B
// gathered along K): one workgroup per CU, two slabs in flight.
template <bool AKM>
hipError_t launch_direct_b(const ContractArgs& a, bool wide, hipStream_t s) {
  if (wide) return launch_rcfg<RCfg<AKM, true, 8, 2, 3, 2>>(a, s);
  if (a.b_kmajor) return launch_rcfg<RCfg<AKM, true, 8, 1, 2, 4>>(a, s);
  return launch_rcfg<RCfg<AKM, false, 8, 1, 2, 4>>(a, s);
}

}  // namespace

bool direct_supported(bool off32, bool ab_vec, DType t) { return off32 && ab_vec && t == DType::F64; }

hipError_t launch_contract_direct(const ContractArgs& a, DType t, bool wide, hipStream_t s) {
  if (a.ksplit > 1 || !direct_supported(a.off32, a.ab_vec, t) || (wide && !a.b_kmajor)) return hipErrorInvalidValue;
  return a.a_kmajor ? launch_direct_b<true>(a, wide, s) : launch_direct_b<false>(a, wide, s);
}

}  // namespace tp
