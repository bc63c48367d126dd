// Storage-precision contraction kernels whose gathers are LDS DMA: every
// lane's buffer load writes its element straight into LDS (buffer_load ... lds),
// so the slabs never pass through registers and no VALU or LDS store
// instruction is spent on them. Selected for f32, f64, c32 and c64 (native, 1m
// and 3m schedules) when every gather offset fits the 32-bit buffer range
// (ContractArgs::off32). Other cases run the register-staged kernels of
// contract.hip, with the same
// semantics (crates/tprims-kernel/src/pack/pack.rs:57-265 packing,
// kernels/reference/scalar.rs:31 real, :64 planar, :110 1m, :136 3m
// microkernels, pack/writeback.rs:114-495
// write-back, driver/tile.rs:142 K accumulation).
//
// A DMA instruction writes lane i's element at LDS slot i of a 64-element
// block, so the LDS layout is fixed by which element each lane fetches:
//  - K-major operands (contiguous along K): a block holds RPI = 64 / BK rows
//    x BK k; lane i fetches row i / BK, k (i % BK) ^ (SW * row). Blocks sit
//    PADB elements apart. PADB and SW make the MFMA fragment reads (16 rows x
//    4 k for 4-byte elements, 16 rows x 1 k per pass for 16-byte ones) hit 64
//    distinct banks: f32 PADB 4, SW 0; c64 PADB 1, SW 4.
//  - M-major operands: one row of BM elements per k, lanes along M, pitch
//    BM + 16 for f32 (lanes of 4 k land on distinct bank quarters).
//  - 16-byte runs (DCfg::VEC; f32, f64, c32 whose scatters come in runs of
//    16 / ES consecutive elements): each lane fetches a whole run, K-major
//    parts as chunk (row, k run) at run * rows + row, M-major parts as chunk
//    (row run x, k) at k * rows / V + (x ^ XQ * (k % XR)).
// 8-byte elements outside run layouts arrive as two 4-byte DMAs into planes.
// Out-of-range rows and k use the buffer sentinels and arrive as zeros.
// Conjugation is applied to the imaginary parts as they are read for the MFMA
// (a sign flip of the high word), the f64 MFMA negate flag forms -im*im.
//
// Pipeline over two LDS stages, one barrier per K slab: MFMAs on slab kt,
// wait for this wave's DMA of slab kt + 1, barrier (slab kt + 1 visible,
// stage kt & 1 retired), then the DMA of slab kt + 2 into stage kt & 1.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Cx;
using dev::Mma;
using dev::Store;
using dev::v4f;

// One LDS-DMA instruction: every lane's ES bytes at byte offset `off` of the
// resource go to LDS at dst + 16 lane (ES 16) or dst + 4 lane (ES 4). The
// builtin is only visible to the device pass: the host pass rejects the
// 16-byte size for its own target and would then drop the kernel stubs.
template <int ES>
__device__ inline void dma_load(__amdgpu_buffer_rsrc_t r, void* dst, uint32_t off) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (ES == 16) __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 16, off, 0, 0, 0);
  else __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 4, off, 0, 0, 0);
#endif
}

// PLR_: read each k step's fragments while the previous step's MFMAs run
// (two fragment sets in registers); off where the accumulators leave no room.
// MINW_: waves per SIMD the register allocation must allow (1: no bound).
// VEC_ (f32, f64, c32): every lane's DMA moves 16 bytes, a run of V = 16 / ES
// elements consecutive in memory along the operand's major index (the
// planner's ContractArgs::ab_vec); a quarter of the DMA instructions of the
// 4-byte form for f32, an eighth for f64 and c32 (no word or re/im planes).
// KG_: K groups. The workgroup's waves form KG groups of NW waves; group g
// computes the whole BM x BN tile over slabs g, g + KG, g + 2 KG, ... (each
// group with its own LDS part of every stage), and the groups' accumulators
// are summed through LDS in group order before the write-back. A small tile
// then keeps every SIMD of its CU busy without a split-K workspace.
template <typename T, bool CPLX_, int BM_, int BN_, int BK_, int WM_, int WN_, bool AKM_, bool BKM_, int METH_,
          bool PLR_ = true, int MINW_ = 1, bool VEC_ = false, int KG_ = 1>
struct DCfg {
  typedef T Tv;
  static constexpr bool CPLX = CPLX_, AKM = AKM_, BKM = BKM_, PLR = PLR_;
  static constexpr int MINW = MINW_;
  static constexpr int METH = CPLX_ ? METH_ : CPLX_NATIVE;
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int KG = KG_;
  static constexpr int NW = (BM / WM) * (BN / WN);  // waves of one K group
  static constexpr int NT = NW * KG * 64;
  typedef typename Store<T, CPLX>::S S;
  static constexpr int ES = sizeof(S);  // element bytes in memory
  // 8-byte elements (no DMA size) arrive as two 4-byte DMAs into separate
  // planes: re and im for c32, the low and high words for f64; f32 and c64
  // elements as one DMA each.
  static constexpr bool VEC = VEC_;
  static_assert(!VEC || ES == 4 || ES == 8, "16-byte runs: f32, f64, c32");
  static constexpr bool PLANAR = ES == 8 && !VEC;
  static constexpr int PLANES = PLANAR ? 2 : 1;
  typedef typename std::conditional<PLANAR, float, S>::type W;  // LDS element
  static constexpr int EW = sizeof(W);
  static constexpr int DMAB = VEC ? 16 : EW;  // DMA size: 4 or 16
  // Run layout (VEC): V elements per 16-byte chunk. K-major parts hold chunk
  // (row r, k run q) at q * ROWS + r; M/N-major parts hold chunk (row run x,
  // k) at k * ROWS / V + (x ^ XS(k)), the XOR spreading the four k rows a
  // fragment read touches over distinct bank groups.
  static constexpr int V = VEC ? 16 / ES : 1;
  static constexpr int XQ = V == 4 ? 4 : 8, XR = V == 4 ? 4 : 2;
  static constexpr int RPI = 64 / BK;   // rows per K-major block
  static constexpr int PADB = EW == 16 ? 1 : 4;
  static constexpr int SW = EW == 16 ? 4 : 0;
  static constexpr int PADM = EW == 16 ? 0 : 16;
  static constexpr int A_ELEMS = VEC ? BM * BK : AKM ? (BM / RPI) * (64 + PADB) : BK * (BM + PADM);
  static constexpr int B_ELEMS = VEC ? BN * BK : BKM ? (BN / RPI) * (64 + PADB) : BK * (BN + PADM);
  static constexpr int PLANE = A_ELEMS + B_ELEMS;  // LDS elements of one plane
  static constexpr int STAGE = PLANES * PLANE;     // LDS elements of one K group's part of a stage
  static constexpr int SSTRIDE = KG * STAGE;       // LDS elements of one stage
  static constexpr int A_INSTR = BM * BK / (64 * V), B_INSTR = BN * BK / (64 * V);
  static constexpr int JA = A_INSTR / NW, JB = B_INSTR / NW;  // DMA instructions per wave per slab
  static constexpr int MI = WM / 16, NI = WN / 16;
  // 1m runs the real GEMM [Dre; Dim] = [[Ar, -Ai], [Ai, Ar]] [Br; Bi]: a
  // wave's WM complex rows are 2 WM real rows (re and im of a complex row
  // adjacent), so it holds MR = 2 MI row blocks of one accumulator.
  static constexpr int MR = METH == CPLX_1M ? 2 * MI : MI;
  // Accumulators: real and 1m 1; c64 native 2 (re, im; the f64 MFMA negates
  // -im*im for free); c32 native 3 (re*re, im*im subtracted at the end,
  // cross); 3m 3 (re*re, im*im, (re+im)*(re+im)).
  // c32 native keeps three accumulators: the f32 MFMA has no negate flag.
  static constexpr bool C3 = CPLX && sizeof(T) == 4 && METH == CPLX_NATIVE;
  static constexpr int ACCS = !CPLX || METH == CPLX_1M ? 1 : (C3 || METH == CPLX_3M ? 3 : 2);
  static constexpr int LDD = BN + 1;
  static constexpr size_t LDS_MAIN = (size_t)2 * SSTRIDE * EW;
  // The finished tile is staged whole when it fits in the main loop's LDS,
  // otherwise one wave row (WM rows) at a time.
  static constexpr size_t EPI_ROW_BYTES = (size_t)(CPLX ? 2 : 1) * LDD * sizeof(T);
  static constexpr int EPI_ROWS = (size_t)BM * EPI_ROW_BYTES <= LDS_MAIN ? BM : WM;
  static constexpr size_t LDS_EPI0 = (size_t)EPI_ROWS * EPI_ROW_BYTES;
  // f32 (real): a lane's four accumulator rows are consecutive, so an output
  // contiguous along M can be staged column-major, one wave column (WN
  // columns, pitch BM + 4) at a time, with 16-byte LDS stores and reads.
  static constexpr bool COLS = !CPLX && sizeof(T) == 4;
  static constexpr int LDM = BM + 4;
  static constexpr size_t LDS_COL = COLS ? (size_t)WN * LDM * 4 : 0;
  static constexpr size_t LDS_EPI = LDS_EPI0 > LDS_COL ? LDS_EPI0 : LDS_COL;
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  static_assert(DMAB == 4 || DMAB == 16, "LDS DMA moves 4 or 16 bytes per lane");
  // Run layout: whole 64-lane instructions; an M/N-major part needs a
  // multiple of 16 chunks per k so that x ^ XQ * (k % XR) stays in the row.
  static_assert(!VEC || ((BM * BK / V) % 64 == 0 && (BN * BK / V) % 64 == 0 && BK % V == 0 &&
                         (AKM || (BM / V) % 16 == 0) && (BKM || (BN / V) % 16 == 0)),
                "run layout: whole 64-lane instructions, room for the XOR");
  static_assert(KG == 1 || 64 % (BK * KG) == 0, "the K offset vectors are padded to a multiple of 64");
  static_assert(64 % BK == 0 && BM % RPI == 0 && BN % RPI == 0, "K-major blocks");
  static_assert(VEC || BM % 64 == 0 || AKM, "M-major rows are whole DMA instructions");
  static_assert(VEC || BN % 64 == 0 || BKM, "N-major rows are whole DMA instructions");
  static_assert(A_INSTR % NW == 0 && B_INSTR % NW == 0, "whole DMA instructions per wave");
  static_assert(BK % 8 == 0 && WM % 16 == 0 && WN % 16 == 0, "MFMA shape; an even number of k steps per slab");
  static_assert(LDS <= 160 * 1024, "LDS budget");

  // LDS slot of element (row r, k) of a stage's A (rows BM) or B (rows BN) part.
  template <bool KM, int ROWS>
  __device__ static inline int slot(int r, int k) {
    if constexpr (VEC && KM) return ((k / V) * ROWS + r) * V + k % V;
    if constexpr (VEC) return (k * (ROWS / V) + ((r / V) ^ (XQ * (k % XR)))) * V + r % V;
    if constexpr (KM) return (r / RPI) * (64 + PADB) + (r % RPI) * BK + (k ^ (SW * (r % RPI)));
    else return k * (ROWS + PADM) + r;
  }
};

// The barrier that hands the staged partial tile from K group g - 1 to group
// g: this wave's LDS stores complete before it. __syncthreads() alone
// compiles here to a bare s_barrier at the top of the groups' loop (the
// compiler counts on the LDS performing every wave's operations in one order
// and drops the wait), and the next group then sometimes read the f32 column
// staging without the previous group's 16-byte stores: wrong tiles in about 1
// call in 100 of the f32 K-group tile at M = 72, N = 140608, K = 72.
__device__ inline void handoff_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __syncthreads();
}

// One BM x BN output tile (tile row tm, column tn) of batch entry h, staged
// through smem (C::LDS bytes): over all of K, written back to D; or, with
// split K, over K range `split` (slabs [split * kslabs, + kslabs)), written
// unscaled to the split-K workspace.
// FAST: the vector epilogues (column-major staging for f32, 16-byte stores in
// the write-back); the grouped body, which inlines four gather orientations,
// keeps the element-wise write-back to hold its register count.
template <class C, bool SPLIT = false, bool FAST = true>
__device__ inline void dma_tile(const ContractArgs& p, int64_t h, int tm, int tn, char* smem, int split = 0) {
  typedef typename C::Tv T;
  typedef typename C::W W;
  typedef Mma<T> MM;
  typedef typename MM::acc Acc;
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, NW = C::NW, KG = C::KG;
  static_assert(!(SPLIT && KG > 1), "split K runs one K group");
  const int lane = threadIdx.x & 63;
  // K group kg (its own part of every stage) and the wave within it.
  const int wave_all = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kg = KG > 1 ? wave_all / NW : 0;
  const int wave = KG > 1 ? wave_all % NW : wave_all;
  W* const lds = (W*)smem + kg * C::STAGE;
  const int wm = (wave % C::WAVES_M) * C::WM;
  const int wn = (wave / C::WAVES_M) * C::WN;
  const int64_t M = p.M, N = p.N, K = p.K;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  typedef typename C::S S;
  const __amdgpu_buffer_rsrc_t rsA = dev::make_rsrc((const S*)p.A + p.ha[h]);
  const __amdgpu_buffer_rsrc_t rsB = dev::make_rsrc((const S*)p.B + p.hb[h]);

  // What each lane fetches: its row (fixed per instruction) and its k within
  // the slab (K-major: fixed per lane; M-major: fixed per instruction).
  uint32_t a_row[C::JA], b_row[C::JB];
  int a_kl[C::JA], b_kl[C::JB];
#pragma unroll
  for (int t = 0; t < C::JA; t++) {
    const int j = wave + NW * t;
    int r, kl;
    if constexpr (C::VEC && C::AKM) {
      const int q = j * 64 + lane;
      r = q % BM;
      kl = C::V * (q / BM);
    } else if constexpr (C::VEC) {
      const int q = j * 64 + lane, k = q / (BM / C::V);
      r = ((q % (BM / C::V)) ^ (C::XQ * (k % C::XR))) * C::V;
      kl = k;
    } else if constexpr (C::AKM) {
      r = j * C::RPI + lane / BK;
      kl = (lane % BK) ^ (C::SW * (lane / BK));
    } else {
      r = (j % (BM / 64)) * 64 + lane;
      kl = j / (BM / 64);
    }
    const int64_t m = m0 + r;
    a_row[t] = m < M ? (uint32_t)(p.am[m] * C::ES) : dev::ROW_OOB;
    a_kl[t] = kl;
  }
#pragma unroll
  for (int t = 0; t < C::JB; t++) {
    const int j = wave + NW * t;
    int r, kl;
    if constexpr (C::VEC && C::BKM) {
      const int q = j * 64 + lane;
      r = q % BN;
      kl = C::V * (q / BN);
    } else if constexpr (C::VEC) {
      const int q = j * 64 + lane, k = q / (BN / C::V);
      r = ((q % (BN / C::V)) ^ (C::XQ * (k % C::XR))) * C::V;
      kl = k;
    } else if constexpr (C::BKM) {
      r = j * C::RPI + lane / BK;
      kl = (lane % BK) ^ (C::SW * (lane / BK));
    } else {
      r = (j % (BN / 64)) * 64 + lane;
      kl = j / (BN / 64);
    }
    const int64_t n = n0 + r;
    b_row[t] = n < N ? (uint32_t)(p.bn[n] * C::ES) : dev::ROW_OOB;
    b_kl[t] = kl;
  }
  // LDS destination of instruction t of this wave within a stage.
  auto a_dst = [&](int t) {
    const int j = wave + NW * t;
    if constexpr (C::VEC) return j * 64 * C::V;
    else if constexpr (C::AKM) return j * (64 + C::PADB);
    else return (j / (BM / 64)) * (BM + C::PADM) + (j % (BM / 64)) * 64;
  };
  auto b_dst = [&](int t) {
    const int j = wave + NW * t;
    if constexpr (C::VEC) return C::A_ELEMS + j * 64 * C::V;
    else if constexpr (C::BKM) return C::A_ELEMS + j * (64 + C::PADB);
    else return C::A_ELEMS + (j / (BN / 64)) * (BN + C::PADM) + (j % (BN / 64)) * 64;
  };

  // K byte offsets of one slab from the padded 32-bit K scatters (the
  // sentinel past K), plain loads whose first use is the DMA issue.
  auto load_koff = [&](int64_t k0, uint32_t* ka, uint32_t* kb) {
#pragma unroll
    for (int t = 0; t < C::JA; t++) ka[t] = p.ak32[k0 + a_kl[t]];
#pragma unroll
    for (int t = 0; t < C::JB; t++) kb[t] = p.bk32[k0 + b_kl[t]];
  };
  // Plane c of an element holds its bytes [4 c, 4 c + 4) (re, then im for c32).
  auto issue = [&](int st, const uint32_t* ka, const uint32_t* kb) {
#pragma unroll
    for (int c = 0; c < C::PLANES; c++) {
      W* stage = lds + st * C::SSTRIDE + c * C::PLANE;
#pragma unroll
      for (int t = 0; t < C::JA; t++) dma_load<C::DMAB>(rsA, stage + a_dst(t), a_row[t] + ka[t] + 4 * c);
#pragma unroll
      for (int t = 0; t < C::JB; t++) dma_load<C::DMAB>(rsB, stage + b_dst(t), b_row[t] + kb[t] + 4 * c);
    }
  };

  Acc acc[C::ACCS][C::MR][C::NI];
#pragma unroll
  for (int s = 0; s < C::ACCS; s++)
#pragma unroll
    for (int i = 0; i < C::MR; i++)
#pragma unroll
      for (int j = 0; j < C::NI; j++) acc[s][i][j] = Acc{};

  const int lr = lane & 15, lk = lane >> 4;
  const uint32_t ca = p.conj_a ? 0x80000000u : 0u, cb = p.conj_b ? 0x80000000u : 0u;
  auto flip = [](T x, uint32_t sign) {
    if constexpr (sizeof(T) == 8) {
      return __builtin_bit_cast(T, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)sign << 32));
    } else {
      return __builtin_bit_cast(T, __builtin_bit_cast(uint32_t, x) ^ sign);
    }
  };
  // One k step as the values a lane feeds its MFMAs, read from LDS ahead of
  // the MFMAs that use them: 4 k for real (MI + NI values), native and 3m
  // ((re, im) pairs with conjugation applied); 2 complex k for 1m (one real
  // k step of 4 of the expanded A, MR values, and of B, NI values).
  constexpr int KSTEP = C::METH == CPLX_1M ? 2 : 4;
  constexpr int FA = !C::CPLX || C::METH == CPLX_1M ? C::MR : 2 * C::MI;
  constexpr int FB = !C::CPLX || C::METH == CPLX_1M ? C::NI : 2 * C::NI;
  struct Frag {
    T a[FA];
    T b[FB];
  };
  auto load_frag = [&](int st, int kk, Frag& f) {
    const W* As = lds + st * C::SSTRIDE;
    const W* Bs = As + C::A_ELEMS;
    if constexpr (!C::CPLX && C::PLANAR) {
      // f64 from its word planes.
      auto join = [](W lo, W hi) {
        return __builtin_bit_cast(T, ((uint64_t)__builtin_bit_cast(uint32_t, hi) << 32) | __builtin_bit_cast(uint32_t, lo));
      };
#pragma unroll
      for (int i = 0; i < C::MI; i++) {
        const int x = C::template slot<C::AKM, BM>(wm + 16 * i + lr, kk + lk);
        f.a[i] = join(As[x], As[C::PLANE + x]);
      }
#pragma unroll
      for (int j = 0; j < C::NI; j++) {
        const int x = C::template slot<C::BKM, BN>(wn + 16 * j + lr, kk + lk);
        f.b[j] = join(Bs[x], Bs[C::PLANE + x]);
      }
    } else if constexpr (!C::CPLX) {
#pragma unroll
      for (int i = 0; i < C::MI; i++) f.a[i] = As[C::template slot<C::AKM, BM>(wm + 16 * i + lr, kk + lk)];
#pragma unroll
      for (int j = 0; j < C::NI; j++) f.b[j] = Bs[C::template slot<C::BKM, BN>(wn + 16 * j + lr, kk + lk)];
    } else if constexpr (C::METH == CPLX_1M) {
      // Lane (lr, lk) holds expanded A at real row lr (complex row lr / 2,
      // part rp = lr & 1) and real k lk (complex k lk / 2, part kp = lk & 1):
      // the re part of A when rp == kp, else the im part, negated in the re
      // row (-Ai); and B's part kp at column lr.
      const int rp = lr & 1, kp = lk & 1, part = rp ^ kp;
      const uint32_t sa = (rp == 0 && kp == 1 ? 0x80000000u : 0u) ^ (part ? ca : 0u);
      const uint32_t sb = kp ? cb : 0u;
      const T* Af = (const T*)As;
      const T* Bf = (const T*)Bs;
      const int kc = kk + (lk >> 1);
#pragma unroll
      for (int i = 0; i < C::MR; i++) {
        const int x = C::template slot<C::AKM, BM>(wm + 8 * i + (lr >> 1), kc);
        f.a[i] = flip(C::PLANAR ? Af[part * C::PLANE + x] : Af[2 * x + part], sa);
      }
#pragma unroll
      for (int j = 0; j < C::NI; j++) {
        const int x = C::template slot<C::BKM, BN>(wn + 16 * j + lr, kc);
        f.b[j] = flip(C::PLANAR ? Bf[kp * C::PLANE + x] : Bf[2 * x + kp], sb);
      }
    } else {
#pragma unroll
      for (int i = 0; i < C::MI; i++) {
        const int x = C::template slot<C::AKM, BM>(wm + 16 * i + lr, kk + lk);
        if constexpr (C::PLANAR) {
          f.a[2 * i] = As[x];
          f.a[2 * i + 1] = flip(As[C::PLANE + x], ca);
        } else {
          const S v = As[x];
          f.a[2 * i] = v.re;
          f.a[2 * i + 1] = flip(v.im, ca);
        }
      }
#pragma unroll
      for (int j = 0; j < C::NI; j++) {
        const int x = C::template slot<C::BKM, BN>(wn + 16 * j + lr, kk + lk);
        if constexpr (C::PLANAR) {
          f.b[2 * j] = Bs[x];
          f.b[2 * j + 1] = flip(Bs[C::PLANE + x], cb);
        } else {
          const S v = Bs[x];
          f.b[2 * j] = v.re;
          f.b[2 * j + 1] = flip(v.im, cb);
        }
      }
    }
  };
  auto mma = [&](const Frag& f) {
    if constexpr (!C::CPLX) {
#pragma unroll
      for (int i = 0; i < C::MI; i++)
#pragma unroll
        for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(f.a[i], f.b[j], acc[0][i][j]);
    } else if constexpr (C::METH == CPLX_1M) {
#pragma unroll
      for (int i = 0; i < C::MR; i++)
#pragma unroll
        for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(f.a[i], f.b[j], acc[0][i][j]);
    } else if constexpr (C::METH == CPLX_3M) {
      // acc0 += ar br, acc1 += ai bi, acc2 += (ar + ai)(br + bi); the
      // epilogue forms re = acc0 - acc1, im = acc2 - acc0 - acc1.
      T as[C::MI], bs[C::NI];
#pragma unroll
      for (int i = 0; i < C::MI; i++) as[i] = f.a[2 * i] + f.a[2 * i + 1];
#pragma unroll
      for (int j = 0; j < C::NI; j++) bs[j] = f.b[2 * j] + f.b[2 * j + 1];
#pragma unroll
      for (int i = 0; i < C::MI; i++)
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
          acc[0][i][j] = MM::run(f.a[2 * i], f.b[2 * j], acc[0][i][j]);
          acc[1][i][j] = MM::run(f.a[2 * i + 1], f.b[2 * j + 1], acc[1][i][j]);
          acc[2][i][j] = MM::run(as[i], bs[j], acc[2][i][j]);
        }
    } else {
      // c64: acc0 += ar br - ai bi, acc1 += ar bi + ai br. c32: acc0 += ar br,
      // acc1 += ai bi (subtracted at the end), acc2 += ar bi + ai br.
      constexpr int XI = C::C3 ? 2 : 1;
#pragma unroll
      for (int i = 0; i < C::MI; i++)
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
          acc[0][i][j] = MM::run(f.a[2 * i], f.b[2 * j], acc[0][i][j]);
          acc[XI][i][j] = MM::run(f.a[2 * i], f.b[2 * j + 1], acc[XI][i][j]);
        }
#pragma unroll
      for (int i = 0; i < C::MI; i++)
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
          if constexpr (C::C3) acc[1][i][j] = MM::run(f.a[2 * i + 1], f.b[2 * j + 1], acc[1][i][j]);
          else acc[0][i][j] = MM::run_neg_a(f.a[2 * i + 1], f.b[2 * j + 1], acc[0][i][j]);
          acc[XI][i][j] = MM::run(f.a[2 * i + 1], f.b[2 * j], acc[XI][i][j]);
        }
    }
  };

  // Main loop: every k step's fragments are read from LDS while the previous
  // step's MFMAs run. The slab barrier sits before the last step's MFMAs (its
  // fragments are already in registers): it publishes slab kt + 1 (each
  // wave waited for its own DMA) and retires stage kt & 1 (every wave has
  // read it), so the DMA of slab kt + 2 and the first fragments of slab
  // kt + 1 are issued under those MFMAs.
  constexpr int KQ = BK / KSTEP;
  // K groups of one wave each read and fill only their own LDS parts: the
  // wave's own waits order its DMA and its reads, and no barrier is needed
  // until the groups meet in the epilogue.
  constexpr bool SOLO = KG > 1 && NW == 1;
  auto slab_barrier = [] {
    if constexpr (!SOLO) __syncthreads();
  };
  // p.kslabs counts slabs of 16 k (the planner's unit); a split's range is a
  // whole number of this kernel's slabs (the launcher checks). With K groups
  // an iteration covers KB = KG * BK k, group kg the BK of them at kg * BK.
  constexpr int KB = BK * KG;
  const int64_t nk_all = (K + KB - 1) / KB;
  const int64_t ks = SPLIT ? p.kslabs * 16 / BK : 0;
  const int64_t kt0 = SPLIT ? (int64_t)split * ks : 0;
  const int64_t nk_rest = nk_all - kt0;
  const int64_t nk = SPLIT ? (nk_rest < ks ? (nk_rest > 0 ? nk_rest : 0) : ks) : nk_all;
  const int64_t kb0 = kt0 * BK + kg * BK;  // this group's first k of the range
  uint32_t a_ko[C::JA], b_ko[C::JB];
  if (nk > 0) {
    uint32_t a1[C::JA], b1[C::JB];
    load_koff(kb0, a_ko, b_ko);
    if (nk > 1) load_koff(kb0 + KB, a1, b1);
    issue(0, a_ko, b_ko);
    if (nk > 1) issue(1, a1, b1);
  }
  __builtin_amdgcn_s_waitcnt(7 << 4);  // vmcnt(0) lgkmcnt(0)
  slab_barrier();
  Frag f[C::PLR ? 2 : 1];
  if constexpr (!C::PLR) {
    for (int64_t kt = 0; kt < nk; kt++) {
      const bool more = kt + 2 < nk;
      const int st = (int)(kt & 1);
      if (more) load_koff(kb0 + (kt + 2) * KB, a_ko, b_ko);
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        load_frag(st, KSTEP * q, f[0]);
        mma(f[0]);
      }
      __builtin_amdgcn_s_waitcnt(7 << 4);  // this wave's DMA of slab kt + 1 (and the offsets)
      slab_barrier();
      if (more) issue(st, a_ko, b_ko);
    }
  } else {
  if (nk > 0) load_frag(0, 0, f[0]);
  for (int64_t kt = 0; kt < nk; kt++) {
    const bool more = kt + 2 < nk;
    const int st = (int)(kt & 1);
    if (more) load_koff(kb0 + (kt + 2) * KB, a_ko, b_ko);  // lands during the MFMAs
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      if (q + 1 < KQ) {
        load_frag(st, KSTEP * (q + 1), f[(q + 1) & 1]);
      } else {
        __builtin_amdgcn_s_waitcnt(7 << 4);  // this wave's DMA of slab kt + 1, the offsets, its LDS reads
        slab_barrier();
        if (more) issue(st, a_ko, b_ko);
        if (kt + 1 < nk) load_frag(st ^ 1, 0, f[(q + 1) & 1]);
      }
      __builtin_amdgcn_sched_barrier(0);
      mma(f[q & 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  }

  if constexpr (C::COLS && !SPLIT && FAST) {
    // f32 output contiguous along M (affine offsets, no broadcast copies,
    // 16-byte aligned columns): stage one wave column of the tile at a time,
    // column-major, and write it back as 16-byte vectors (alpha and beta * C
    // in registers). K groups add in turn, in group order.
    float* __restrict__ Dg = (float*)p.D + p.hd[h];
    const float* __restrict__ Cg = (const float*)p.C + p.hc[h];
    const bool col_ok = p.d_mmajor && p.mn_affine && p.dr == nullptr && p.cd_stride[2] == 1 &&
                        p.cd_stride[3] % 4 == 0 && (uintptr_t)Dg % 16 == 0 &&
                        (!p.read_c || (p.cd_stride[0] == 1 && p.cd_stride[1] % 4 == 0 && (uintptr_t)Cg % 16 == 0));
    if (col_ok) {
      constexpr int LDM = C::LDM, WN = C::WN, NTH = C::NT;
      float* Dt = (float*)smem;
      const float alr = (float)p.alpha[0], ber = (float)p.beta[0];
      if constexpr (SOLO) __syncthreads();  // every group is done with its stages
#pragma unroll 1
      for (int c = 0; c < BN / WN; c++) {
        __syncthreads();  // the stages, or the previous column's staging, are free
#pragma unroll 1
        for (int g = 0; g < KG; g++) {
          if (g > 0) handoff_barrier();
          if (kg == g && wn == c * WN) {
#pragma unroll
            for (int i = 0; i < C::MR; i++)
#pragma unroll
              for (int j = 0; j < C::NI; j++) {
                v4f* dst = (v4f*)&Dt[(16 * j + lr) * LDM + wm + 16 * i + 4 * lk];
                const v4f v = __builtin_bit_cast(v4f, acc[0][i][j]);
                *dst = g > 0 ? *dst + v : v;
              }
          }
        }
        __syncthreads();
#pragma unroll 1
        for (int e = threadIdx.x; e < BM / 4 * WN; e += NTH) {
          const int ml = 4 * (e % (BM / 4)), nl = e / (BM / 4);
          const int64_t m = m0 + ml, n = n0 + c * WN + nl;
          if (m >= M || n >= N) continue;
          const v4f x = *(const v4f*)&Dt[nl * LDM + ml];
          const int64_t off = m + n * p.cd_stride[3], coff = m + n * p.cd_stride[1];
          if (M - m >= 4) {
            v4f y = alr * x;
            if (p.read_c) y += ber * *(const v4f*)(Cg + coff);
            *(v4f*)(Dg + off) = y;
          } else {
            for (int q = 0; q < M - m; q++) Dg[off + q] = alr * x[q] + (p.read_c ? ber * Cg[coff + q] : 0.0f);
          }
        }
      }
      return;
    }
  }

  // Stage the tile chunk by chunk (Dre/Dim [EPI_ROWS][LDD]) and write it back.
  // K groups store in turn, group 0 first and each later one adding its
  // partial sums, so the order of the sum is fixed.
  constexpr int ER = C::EPI_ROWS;
  T* Dre = (T*)smem;
  T* Dim = Dre + ER * C::LDD;
  auto put = [](T* dst, T v, bool add) { *dst = add ? *dst + v : v; };
  if constexpr (SOLO) __syncthreads();  // every group is done with its stages
#pragma unroll 1
  for (int c = 0; c < BM / ER; c++) {
    if (c > 0) __syncthreads();  // the previous chunk's write-back has read the staging rows
#pragma unroll 1
    for (int g = 0; g < KG; g++) {
      if (g > 0) handoff_barrier();  // group g - 1's values are in place
      const bool add = g > 0;
      if (kg == g && wm / ER == c) {
#pragma unroll
        for (int i = 0; i < C::MR; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
              const int col = wn + 16 * j + lr;
              if constexpr (C::METH == CPLX_1M) {
                const int rr = 16 * i + MM::row(lk, q), row = wm - c * ER + rr / 2;
                put(&(rr & 1 ? Dim : Dre)[row * C::LDD + col], acc[0][i][j][q], add);
                continue;
              }
              const int row = wm - c * ER + 16 * i + MM::row(lk, q);
              if constexpr (C::METH == CPLX_3M) {
                put(&Dre[row * C::LDD + col], acc[0][i][j][q] - acc[1][i][j][q], add);
                put(&Dim[row * C::LDD + col], acc[2][i][j][q] - acc[0][i][j][q] - acc[1][i][j][q], add);
              } else if constexpr (C::C3) {
                put(&Dre[row * C::LDD + col], acc[0][i][j][q] - acc[1][i][j][q], add);
                put(&Dim[row * C::LDD + col], acc[2][i][j][q], add);
              } else {
                put(&Dre[row * C::LDD + col], acc[0][i][j][q], add);
                if constexpr (C::CPLX) put(&Dim[row * C::LDD + col], acc[1][i][j][q], add);
              }
            }
      }
    }
    __syncthreads();
    if constexpr (SPLIT)
      dev::write_partial<T, C::CPLX, ER, BN, C::NT, C::LDD>(p, split, h, m0 + c * ER, n0, Dre, Dim);
    else
      dev::write_back<T, T, C::CPLX, ER, BN, C::NT, C::LDD, FAST>(p, h, m0 + c * ER, n0, Dre, Dim);
  }
}

// SPLIT: K split into p.ksplit ranges, the splits of a tile being ksplit
// consecutive workgroup ids.
template <class C, bool SPLIT>
__global__ __launch_bounds__(C::NT, C::MINW) void k_contract_dma(ContractArgs p) {
  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  const int S = SPLIT ? p.ksplit : 1;
  int64_t h;
  int tm, tn;
  dev::decode_tile((int64_t)blockIdx.x / S, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  dma_tile<C, SPLIT>(p, h, tm, tn, smem, (int)(blockIdx.x % S));
}

// The four gather orientations of one tile shape (A and B each K- or
// M/N-major), chosen per tile from the item's arguments.
template <typename T, bool CPLX, int BM, int BN, int BK, int WM, int WN, int METH, bool PLR, int MW>
struct DFamily {
  typedef DCfg<T, CPLX, BM, BN, BK, WM, WN, true, true, METH, PLR, MW> KK;
  typedef DCfg<T, CPLX, BM, BN, BK, WM, WN, true, false, METH, PLR, MW> KN;
  typedef DCfg<T, CPLX, BM, BN, BK, WM, WN, false, true, METH, PLR, MW> MK;
  typedef DCfg<T, CPLX, BM, BN, BK, WM, WN, false, false, METH, PLR, MW> MN;
  static constexpr size_t m2(size_t x, size_t y) { return x > y ? x : y; }
  static constexpr size_t LDS = m2(m2(KK::LDS, KN::LDS), m2(MK::LDS, MN::LDS));
  static constexpr int NT = KK::NT;
  static constexpr int MINW = KK::MINW;
};

// Persistent grouped kernel on the DMA tiles: workgroups stride over the
// flattened tile space of every item; each tile finds its item by binary
// search over tile_begin (every item has 32-bit gather offsets).
template <class F>
__global__ __launch_bounds__(F::NT, F::MINW) void k_grouped_dma(const GroupItem* __restrict__ items, int num_items,
                                                       int64_t total_tiles) {
  __shared__ __attribute__((aligned(16))) char smem[F::LDS];
  for (int64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    int lo = 0, hi = num_items - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    // The item is the workgroup's: its fields load once into scalar registers.
    lo = __builtin_amdgcn_readfirstlane(lo);
    const ContractArgs& p = items[lo].args;
    const int64_t local = t - items[lo].tile_begin;
    const int64_t per_h = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t h = local / per_h;
    const int r = (int)(local - h * per_h);
    const int tm = r % p.tiles_m, tn = r / p.tiles_m;
    if (p.a_kmajor) {
      if (p.b_kmajor) dma_tile<typename F::KK, false, false>(p, h, tm, tn, smem);
      else dma_tile<typename F::KN, false, false>(p, h, tm, tn, smem);
    } else {
      if (p.b_kmajor) dma_tile<typename F::MK, false, false>(p, h, tm, tn, smem);
      else dma_tile<typename F::MN, false, false>(p, h, tm, tn, smem);
    }
    __syncthreads();  // the next tile's DMA overwrites this tile's staging
  }
}

template <class C>
hipError_t launch_cfg(const ContractArgs& a, hipStream_t s) {
  if (a.ksplit > 1 && (a.kslabs * 16) % C::BK != 0) return hipErrorInvalidValue;  // split range not whole slabs
  ContractArgs p = a;
  p.tiles_m = (int)((p.M + C::BM - 1) / C::BM);
  p.tiles_n = (int)((p.N + C::BN - 1) / C::BN);
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H * (p.ksplit > 1 ? p.ksplit : 1);
  if (nwg == 0) return hipSuccess;
  if (nwg > max_grid(C::NT)) return hipErrorInvalidConfiguration;
  if constexpr (C::KG > 1) {
    if (p.ksplit > 1) return hipErrorInvalidValue;  // K groups run one K range
    hipLaunchKernelGGL((k_contract_dma<C, false>), dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
  } else {
    if (p.ksplit > 1) hipLaunchKernelGGL((k_contract_dma<C, true>), dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
    else hipLaunchKernelGGL((k_contract_dma<C, false>), dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
  }
  return hipGetLastError();
}

template <typename T, bool CPLX, int BM, int BN, int BK, int WM, int WN, int METH, bool PLR, int MINW, bool VEC = false>
hipError_t launch_majors(const ContractArgs& a, hipStream_t s) {
  if (a.a_kmajor) {
    if (a.b_kmajor) return launch_cfg<DCfg<T, CPLX, BM, BN, BK, WM, WN, true, true, METH, PLR, MINW, VEC>>(a, s);
    return launch_cfg<DCfg<T, CPLX, BM, BN, BK, WM, WN, true, false, METH, PLR, MINW, VEC>>(a, s);
  }
  if (a.b_kmajor) return launch_cfg<DCfg<T, CPLX, BM, BN, BK, WM, WN, false, true, METH, PLR, MINW, VEC>>(a, s);
  return launch_cfg<DCfg<T, CPLX, BM, BN, BK, WM, WN, false, false, METH, PLR, MINW, VEC>>(a, s);
}

// The wave layout of every DMA tile (BM x BN over waves of WM x WN, fragment
// prefetch PLR, register bound MINW waves per SIMD) per dtype and complex
// schedule, shared by single and grouped launches: act.go<...>() launches it.
// Each choice was measured against its alternatives on the benchmark shapes
// (same GPU, same run):
//  - c64 128x128: 16 waves of 32x32, prefetch for native (62.5 -> 64.3
//    TFLOP/s against 8 waves of 64x32), none for 3m (80.8 -> 86.1); 1m 8
//    waves of 32x64 with prefetch (60.7 against 58 for 64x32, 51 for 16
//    waves). c64 64x64: 8 waves of 32x16 (54.5 -> 58 against 4 of 32x32;
//    3m without prefetch: 63 against 54).
//  - c32 128x128: 16 waves of 32x32 (native 101 -> 109.5 and 3m 105 -> 119
//    without prefetch, 1m 89 -> 97 with it), slabs of 32 k (Single::go).
//  - f64 128x128 (grouped, M-major A, or any 16-byte run layout): 8 waves of
//    64x32 without prefetch, at most 128 VGPRs so two workgroups share a CU; f64 64x64: 8 waves of
//    32x16 (f64 24^6 31.2 -> 28.4 us against 4 of 32x32).
//  - f32 128x128: 8 waves of 64x32 without prefetch (122.7 against 121.1).
template <typename T, bool CPLX, int METH, class Act>
hipError_t by_tile(int tile_id, const Act& act) {
  if constexpr (CPLX && sizeof(T) == 8) {
    if (tile_id == TILE_128x128) {
      if constexpr (METH == CPLX_NATIVE) return act.template go<128, 128, 32, 32, true, 1>();
      else if constexpr (METH == CPLX_3M) return act.template go<128, 128, 32, 32, false, 1>();
      else return act.template go<128, 128, 32, 64, true, 1>();
    }
    return act.template go<64, 64, 32, 16, METH != CPLX_3M, 1>();
  } else if constexpr (CPLX) {
    if (tile_id == TILE_128x128) return act.template go<128, 128, 32, 32, METH == CPLX_1M, 1>();
    if (tile_id == TILE_128x64) return act.template go<128, 64, 64, 32, true, 1>();
    return act.template go<64, 64, 32, 32, true, 1>();
  } else if constexpr (sizeof(T) == 8) {
    if (tile_id == TILE_128x128) return act.template go<128, 128, 64, 32, false, 4>();
    if (tile_id == TILE_128x64) return act.template go<128, 64, 64, 32, true, 1>();
    return act.template go<64, 64, 32, 16, true, 1>();
  } else {
    if (tile_id == TILE_128x128) return act.template go<128, 128, 64, 32, false, 1>();
    if (tile_id == TILE_128x64) return act.template go<128, 64, 64, 32, true, 1>();
    return act.template go<64, 64, 32, 32, true, 1>();
  }
}

template <typename T, bool CPLX, int METH>
struct Single {
  const ContractArgs& a;
  hipStream_t s;
  template <int BM, int BN, int WM, int WN, bool PLR, int MINW>
  hipError_t go() const {
    // c32 128x128: slabs of 32 k (its 8-byte elements fill the LDS of the c64
    // tile's 16-k slabs): half the slab barriers of the f32 MFMA's shorter K
    // steps (complex benchmark contraction: native 101.4 -> 110.9 TFLOP/s, 3m
    // 111.3 -> 125.9). 1m only with A gathered along K (it spills with an
    // M-major A). A K split into an odd number of 16-k slabs keeps 16.
    constexpr bool k32 = CPLX && sizeof(T) == 4 && BM == 128 && BN == 128;
    if constexpr (k32) {
      if (a.ksplit <= 1 || a.kslabs % 2 == 0) {
        if constexpr (METH == CPLX_1M) {
          if (a.a_kmajor) {
            if (a.ab_vec) {
              if (a.b_kmajor) return launch_cfg<DCfg<T, CPLX, BM, BN, 32, WM, WN, true, true, METH, PLR, MINW, true>>(a, s);
              return launch_cfg<DCfg<T, CPLX, BM, BN, 32, WM, WN, true, false, METH, PLR, MINW, true>>(a, s);
            }
            if (a.b_kmajor) return launch_cfg<DCfg<T, CPLX, BM, BN, 32, WM, WN, true, true, METH, PLR, MINW>>(a, s);
            return launch_cfg<DCfg<T, CPLX, BM, BN, 32, WM, WN, true, false, METH, PLR, MINW>>(a, s);
          }
        } else {
          if (a.ab_vec) return launch_majors<T, CPLX, BM, BN, 32, WM, WN, METH, PLR, MINW, true>(a, s);
          return launch_majors<T, CPLX, BM, BN, 32, WM, WN, METH, PLR, MINW>(a, s);
        }
      }
    }
    if constexpr (!CPLX || sizeof(T) == 4)
      if (a.ab_vec) return launch_majors<T, CPLX, BM, BN, 16, WM, WN, METH, PLR, MINW, true>(a, s);
    return launch_majors<T, CPLX, BM, BN, 16, WM, WN, METH, PLR, MINW>(a, s);
  }
};

template <typename T, bool CPLX, int METH>
struct Grouped {
  const GroupItem* items;
  int n;
  int64_t total;
  int grid;
  hipStream_t s;
  // The grouped body holds the persistent loop's state and all four gather
  // orientations, so it runs its 128x128 tiles one workgroup per CU with no
  // register bound (f64: no spills, where the single launch's two workgroups
  // per CU spill outside the K loop). The c64 and c32 1m 128x128 bodies still
  // spill inside the K loop within a 16-wave workgroup's 128 registers, so
  // groups do not offer that tile for them and they are not built.
  template <int BM, int BN, int WM, int WN, bool PLR, int MINW>
  hipError_t go() const {
    constexpr bool big = BM == 128 && BN == 128;
    if constexpr (big && CPLX && (sizeof(T) == 8 || METH == CPLX_1M)) {
      return hipErrorInvalidValue;
    } else {
      return launch<BM, BN, WM, WN, PLR, big ? 1 : MINW>();
    }
  }
  template <int BM, int BN, int WM, int WN, bool PLR, int MINW>
  hipError_t launch() const {
    typedef DFamily<T, CPLX, BM, BN, 16, WM, WN, METH, PLR, MINW> F;
    static_assert(F::LDS <= 160 * 1024, "LDS budget");
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grouped_dma<F>, dim3((unsigned)grid), dim3(F::NT), 0, s, items, n, total);
    return hipGetLastError();
  }
};

template <typename T, bool CPLX>
hipError_t single_by_method(const ContractArgs& a, int method, int tile_id, hipStream_t s) {
  if constexpr (CPLX) {
    if (method == CPLX_3M) return by_tile<T, true, CPLX_3M>(tile_id, Single<T, true, CPLX_3M>{a, s});
    if (method == CPLX_1M) return by_tile<T, true, CPLX_1M>(tile_id, Single<T, true, CPLX_1M>{a, s});
  }
  return by_tile<T, CPLX, CPLX_NATIVE>(tile_id, Single<T, CPLX, CPLX_NATIVE>{a, s});
}

template <typename T, bool CPLX>
hipError_t grouped_by_method(const GroupItem* items, int n, int64_t total, int method, int tile_id, int grid,
                             hipStream_t s) {
  if constexpr (CPLX) {
    if (method == CPLX_3M)
      return by_tile<T, true, CPLX_3M>(tile_id, Grouped<T, true, CPLX_3M>{items, n, total, grid, s});
    if (method == CPLX_1M)
      return by_tile<T, true, CPLX_1M>(tile_id, Grouped<T, true, CPLX_1M>{items, n, total, grid, s});
  }
  return by_tile<T, CPLX, CPLX_NATIVE>(tile_id, Grouped<T, CPLX, CPLX_NATIVE>{items, n, total, grid, s});
}

// TILE_32x32K: a 32 x 32 tile on four K groups of one 32 x 32 wave each.
template <typename T, bool CPLX, int METH, bool VEC, bool AKM, bool BKM>
hipError_t launch_kgroup_cfg(const ContractArgs& a, hipStream_t s) {
  return launch_cfg<DCfg<T, CPLX, 32, 32, 16, 32, 32, AKM, BKM, METH, true, 1, VEC, 4>>(a, s);
}

template <typename T, bool CPLX, int METH>
hipError_t launch_kgroup_majors(const ContractArgs& a, hipStream_t s) {
  typedef typename Store<T, CPLX>::S S;
  if constexpr (sizeof(S) < 16) {
    if (a.ab_vec) {
      if (a.a_kmajor && a.b_kmajor) return launch_kgroup_cfg<T, CPLX, METH, true, true, true>(a, s);
      // 16 chunks of 2 elements per k of a 32-row M/N-major part (f64, c32).
      if constexpr (sizeof(S) == 8) {
        if (a.a_kmajor) return launch_kgroup_cfg<T, CPLX, METH, true, true, false>(a, s);
        if (a.b_kmajor) return launch_kgroup_cfg<T, CPLX, METH, true, false, true>(a, s);
        return launch_kgroup_cfg<T, CPLX, METH, true, false, false>(a, s);
      }
      return hipErrorInvalidValue;
    }
  }
  if (a.a_kmajor && a.b_kmajor) return launch_kgroup_cfg<T, CPLX, METH, false, true, true>(a, s);
  return hipErrorInvalidValue;
}

template <typename T, bool CPLX>
hipError_t kgroup_by_method(const ContractArgs& a, int method, hipStream_t s) {
  if constexpr (CPLX) {
    if (method == CPLX_3M) return launch_kgroup_majors<T, true, CPLX_3M>(a, s);
    if (method == CPLX_1M) return launch_kgroup_majors<T, true, CPLX_1M>(a, s);
  }
  return launch_kgroup_majors<T, CPLX, CPLX_NATIVE>(a, s);
}

}  // namespace

bool kgroup_supported(bool off32, bool ab_vec, bool a_kmajor, bool b_kmajor, DType t) {
  if (!off32) return false;
  if (ab_vec && t != DType::C64) return (a_kmajor && b_kmajor) || t == DType::F64 || t == DType::C32;
  return a_kmajor && b_kmajor;
}

hipError_t launch_contract_kgroup(const ContractArgs& a, DType t, int method, hipStream_t s) {
  if (a.ksplit > 1 || !kgroup_supported(a.off32, a.ab_vec, a.a_kmajor, a.b_kmajor, t)) return hipErrorInvalidValue;
  switch (t) {
    case DType::C64: return kgroup_by_method<double, true>(a, method, s);
    case DType::C32: return kgroup_by_method<float, true>(a, method, s);
    case DType::F64: return kgroup_by_method<double, false>(a, method, s);
    case DType::F32: return kgroup_by_method<float, false>(a, method, s);
  }
  return hipErrorInvalidValue;
}

// Every dtype and complex schedule with 32-bit gather offsets. On the complex
// benchmark the c64 64x64 DMA kernel ran 2.41 ms against 2.98 ms for the
// register-staged one on a GPU that holds its clock down under load (it moves
// no slab data through VGPRs, and so draws less power), and 2.44-2.50 ms
// against 2.35 ms on one that does not.
bool dma_supported(bool off32, DType t, int method, int tile_id) {
  if (!off32) return false;
  if (t == DType::C64) return tile_id == TILE_64x64 || tile_id == TILE_128x128;
  if (t == DType::C32) return tile_id == TILE_64x64 || tile_id == TILE_128x64 || tile_id == TILE_128x128;
  return tile_id == TILE_128x128 || tile_id == TILE_128x64 || tile_id == TILE_64x64;  // f32, f64
}

hipError_t launch_contract_dma(const ContractArgs& a, DType t, int method, int tile_id, hipStream_t s) {
  switch (t) {
    case DType::C64: return single_by_method<double, true>(a, method, tile_id, s);
    case DType::C32: return single_by_method<float, true>(a, method, tile_id, s);
    case DType::F64:
      // 128x128 with K-major A in 4-byte gathers: 16 waves of 32x32 with
      // fragment prefetch (f64 4096^3 GEMM 55.2 against 51.2 TFLOP/s for the
      // 8-wave layout). With 16-byte runs the 8-wave layout of the table
      // serves every major (K-major A 56.3 -> 62.8 TFLOP/s; M-major A 59.4
      // against 55.9 for 16 waves).
      if (tile_id == TILE_128x128 && a.a_kmajor && !a.ab_vec) {
        if (a.b_kmajor) return launch_cfg<DCfg<double, false, 128, 128, 16, 32, 32, true, true, CPLX_NATIVE>>(a, s);
        return launch_cfg<DCfg<double, false, 128, 128, 16, 32, 32, true, false, CPLX_NATIVE>>(a, s);
      }
      return single_by_method<double, false>(a, method, tile_id, s);
    case DType::F32: return single_by_method<float, false>(a, method, tile_id, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_grouped_dma(const GroupItem* items, int n, int64_t total, DType t, int method, int tile_id, int grid,
                              hipStream_t s) {
  switch (t) {
    case DType::C64: return grouped_by_method<double, true>(items, n, total, method, tile_id, grid, s);
    case DType::C32: return grouped_by_method<float, true>(items, n, total, method, tile_id, grid, s);
    case DType::F64: return grouped_by_method<double, false>(items, n, total, method, tile_id, grid, s);
    case DType::F32: return grouped_by_method<float, false>(items, n, total, method, tile_id, grid, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tp
