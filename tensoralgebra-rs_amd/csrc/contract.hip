// Storage-precision contraction kernels for gfx950 (f32/f64 real, c32/c64
// complex), the GPU form of the reference's packed five-loop driver:
//
//   crates/tprims-kernel/src/pack/pack.rs:57-265   packing (block-scatter gather,
//                                                  conj folded, planar / 3m planes)
//   crates/tprims-kernel/src/kernels/reference/    microkernel semantics
//     scalar.rs:31 real, :64 planar, :110 1m, :136 3m
//   crates/tprims-kernel/src/pack/writeback.rs:114-495  write-back
//   crates/tprims-contract/src/driver/tile.rs:142     K accumulation
//
// One workgroup computes one BM x BN tile of one batch entry h over the whole
// K range, or with split K over one of ksplit ranges (partial tiles to the
// workspace, reduced in split order afterwards), so the result does not
// depend on the grid. These kernels serve operands whose gathers need 64-bit
// offsets, the 32x32 tile and the grouped launches whose items do; the
// LDS-DMA kernels of dma.hip serve the rest. Each
// K step gathers a BM x BK slab of A and a BK x BN slab of B straight from HBM
// through the scatter vectors into LDS planes, folding conjugation and the 3m
// sum plane into the store. The lane order of a gather follows the operand's
// contiguous axis (K or M/N, chosen by the planner and fixed at compile time),
// and so does the plane layout in LDS ([m][k] for K-major gathers, [k][m]
// otherwise), which keeps both the stores and the MFMA fragment reads free of
// bank conflicts. Waves run 16x16x4 MFMAs (f64 or f32) from the planes.
//
// Software pipeline per K step over two LDS stages: the MFMAs of the current
// slab, then the store of the next slab (fetched during the previous step)
// into the other stage, the gathers of the slab after it (their K offsets
// were loaded one step earlier, so no load waits on another) and the offset
// loads of the one after that, then a single barrier. The stores and gathers
// do not touch the stage being multiplied, so they overlap the MFMAs. Out-of-range lanes
// load the operand's first element and discard it: the gathers have no
// branches. The finished tile is staged through LDS and written back along
// D's contiguous axis with alpha, beta, op_C and op_D fused.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Store;

using dev::Mma;

enum { M_REAL = -1, M_NATIVE = 0, M_1M = 1, M_3M = 2 };

template <typename T, int METHOD, int BM_, int BN_, int BK_, int WM_, int WN_, bool AKM_, bool BKM_, bool BUF_>
struct Cfg {
  typedef T Tv;
  static constexpr bool BUF = BUF_;  // 32-bit buffer offsets (ContractArgs::off32)
  static constexpr int METH = METHOD;
  static constexpr bool CPLX = METHOD != M_REAL;
  static constexpr bool AKM = AKM_, BKM = BKM_;
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int NT = (BM / WM) * (BN / WN) * 64;
  // The native complex schedule keeps re and im of an element together (one
  // 8- or 16-byte LDS element, one store and one read each); 1m and 3m use
  // separate planes of T.
  static constexpr bool ILV = METHOD == M_NATIVE;
  static constexpr int EW = ILV ? 2 : 1;  // T values per LDS element
  static constexpr int PLANES = METHOD == M_REAL || ILV ? 1 : (METHOD == M_3M ? 3 : 2);
  // Plane layouts (in elements): K-major gathers store [m][k], M-major
  // gathers [k][m] with pitch BM+16. The K-major pitch puts the rows of a
  // fragment read on distinct banks of the 64 x 4-byte LDS banks: odd (BK+1)
  // for 4- and 16-byte elements (16 rows of one k in one pass); BK+2 for
  // 8-byte elements (twice an odd number, so a pass's 32 reads, 16 rows x 2 k,
  // cover all 64 banks exactly once).
  static constexpr int KPAD = EW * sizeof(T) == 8 ? 2 : 1;
  static constexpr int LDAK = BK + KPAD, LDAM = BM + 16;
  static constexpr int LDBK = BK + KPAD, LDBN = BN + 16;
  static constexpr int A_PLANE = EW * (AKM ? BM * LDAK : BK * LDAM);  // in T
  static constexpr int B_PLANE = EW * (BKM ? BN * LDBK : BK * LDBN);
  static constexpr int LDD = BN + 1;  // row pitch of the staging tile
  static constexpr int EA = BM * BK / NT;
  static constexpr int EB = BN * BK / NT;
  static constexpr int KA = AKM ? 1 : EA;  // K offsets a thread tracks for A
  static constexpr int KB = BKM ? 1 : EB;
  static constexpr int RA = AKM ? EA : 1;  // row offsets a thread tracks for A
  static constexpr int RB = BKM ? EB : 1;
  static constexpr int MI = WM / 16, NI = WN / 16;
  // Accumulators: real 1; native 2 (re, im) for f64, whose MFMA negates an
  // operand for free, 3 (re*re, im*im, cross) for f32; 1m 1; 3m 3.
  static constexpr bool NEG_MFMA = sizeof(T) == 8;
  static constexpr int ACCS =
      METHOD == M_REAL || METHOD == M_1M ? 1 : (METHOD == M_NATIVE && NEG_MFMA ? 2 : 3);
  static constexpr int STAGE = PLANES * (A_PLANE + B_PLANE);  // T values of one LDS stage
  static constexpr size_t LDS_MAIN = (size_t)2 * STAGE * sizeof(T);
  // The finished tile is staged in row chunks: the whole tile when it fits in
  // the main loop's LDS, otherwise one wave row (WM rows) at a time, so the
  // staging never raises the workgroup's LDS footprint above the main loop's
  // by more than one wave row.
  static constexpr size_t EPI_ROW_BYTES = (size_t)(CPLX ? 2 : 1) * LDD * sizeof(T);
  static constexpr int EPI_ROWS = (size_t)BM * EPI_ROW_BYTES <= LDS_MAIN ? BM : WM;
  static constexpr size_t LDS_EPI = (size_t)EPI_ROWS * EPI_ROW_BYTES;
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  __device__ static inline int aidx(int m, int k) { return AKM ? m * LDAK + k : k * LDAM + m; }
  __device__ static inline int bidx(int n, int k) { return BKM ? n * LDBK + k : k * LDBN + n; }
  static_assert(NT % BK == 0 && BM % (NT / BK) == 0 && BN % (NT / BK) == 0, "K-major lane map");
  static_assert(NT % BM == 0 && NT % BN == 0 && BK % (NT / BM) == 0 && BK % (NT / BN) == 0, "MN-major lane map");
  static_assert(BK % 4 == 0 && WM % 16 == 0 && WN % 16 == 0, "MFMA shape");
};

// One tile of one contraction, the body shared by the single-contraction
// kernel and the grouped persistent kernel.
template <class C>
__device__ inline void contract_tile(const ContractArgs& p, int64_t h, int tm, int tn, char* smem, int split = 0) {
  typedef typename C::Tv T;
  typedef Store<T, C::CPLX> St;
  typedef typename St::S S;
  typedef Mma<T> MM;
  typedef typename MM::acc Acc;
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, NT = C::NT;
  constexpr int EA = C::EA, EB = C::EB;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave % C::WAVES_M) * C::WM;
  const int wn = (wave / C::WAVES_M) * C::WN;

  const S* __restrict__ A = (const S*)p.A + p.ha[h];
  const S* __restrict__ B = (const S*)p.B + p.hb[h];
  const int64_t M = p.M, N = p.N, K = p.K;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t* __restrict__ ak = p.ak;
  const int64_t* __restrict__ bk = p.bk;
  // Buffer resources over the batch entry (used by the BUF variant only).
  const __amdgpu_buffer_rsrc_t rsA = dev::make_rsrc(A);
  const __amdgpu_buffer_rsrc_t rsB = dev::make_rsrc(B);
  constexpr int64_t ES = sizeof(S);
  // Row and K offsets: elements (64-bit) or, under BUF, bytes (32-bit) with
  // out-of-range entries set to the sentinels.
  typedef typename std::conditional<C::BUF, uint32_t, int64_t>::type Off;

  // Two LDS stages, each [PLANES][A_PLANE] then [PLANES][B_PLANE].
  T* const lds = (T*)smem;

  // Gather coordinates. K-major: one k per thread, EA rows. M-major: one row,
  // EA k's. (Likewise for B.)
  const int a_k0 = C::AKM ? tid % BK : tid / BM;
  const int a_m0 = C::AKM ? tid / BK : tid % BM;
  constexpr int a_kstep = C::AKM ? 0 : NT / BM;
  constexpr int a_mstep = C::AKM ? NT / BK : 0;
  const int b_k0 = C::BKM ? tid % BK : tid / BN;
  const int b_n0 = C::BKM ? tid / BK : tid % BN;
  constexpr int b_kstep = C::BKM ? 0 : NT / BN;
  constexpr int b_nstep = C::BKM ? NT / BK : 0;

  Off a_row[C::RA], b_col[C::RB];
  uint32_t a_rok = 0, b_cok = 0;
#pragma unroll
  for (int r = 0; r < C::RA; r++) {
    const int64_t m = m0 + a_m0 + r * a_mstep;
    const bool ok = m < M;
    const int64_t v = p.am[ok ? m : 0];
    if constexpr (C::BUF) a_row[r] = ok ? (uint32_t)(v * ES) : dev::ROW_OOB;
    else a_row[r] = v;
    a_rok |= (uint32_t)ok << r;
  }
#pragma unroll
  for (int r = 0; r < C::RB; r++) {
    const int64_t n = n0 + b_n0 + r * b_nstep;
    const bool ok = n < N;
    const int64_t v = p.bn[ok ? n : 0];
    if constexpr (C::BUF) b_col[r] = ok ? (uint32_t)(v * ES) : dev::ROW_OOB;
    else b_col[r] = v;
    b_cok |= (uint32_t)ok << r;
  }

  // K offsets of the slab to gather next (loaded one step ahead).
  Off a_ko[C::KA], b_ko[C::KB];
  uint32_t a_kok = 0, b_kok = 0;
  const int64_t kmax = K - 1;
#define TP_LOAD_KOFF(k0)                                                       \
  {                                                                            \
    a_kok = 0;                                                                 \
    b_kok = 0;                                                                 \
    _Pragma("unroll") for (int q = 0; q < C::KA; q++) {                        \
      const int64_t k = (k0) + a_k0 + q * a_kstep;                             \
      const int64_t v = ak[k < K ? k : kmax];                                  \
      if constexpr (C::BUF) a_ko[q] = k < K ? (uint32_t)(v * ES) : dev::K_OOB; \
      else a_ko[q] = v;                                                        \
      a_kok |= (uint32_t)(k < K) << q;                                         \
    }                                                                          \
    _Pragma("unroll") for (int q = 0; q < C::KB; q++) {                        \
      const int64_t k = (k0) + b_k0 + q * b_kstep;                             \
      const int64_t v = bk[k < K ? k : kmax];                                  \
      if constexpr (C::BUF) b_ko[q] = k < K ? (uint32_t)(v * ES) : dev::K_OOB; \
      else b_ko[q] = v;                                                        \
      b_kok |= (uint32_t)(k < K) << q;                                         \
    }                                                                          \
  }

  S ra[EA], rb[EB];
  uint32_t a_live = 0, b_live = 0;  // which gathered elements are in range
#define TP_GATHER()                                                              \
  {                                                                              \
    a_live = 0;                                                                  \
    b_live = 0;                                                                  \
    _Pragma("unroll") for (int i = 0; i < EA; i++) {                             \
      const int r = C::AKM ? i : 0, q = C::AKM ? 0 : i;                          \
      const bool ok = ((a_rok >> r) & 1) && ((a_kok >> q) & 1);                 \
      if constexpr (C::BUF) ra[i] = dev::buf_load<S>(rsA, a_row[r] + a_ko[q]);  \
      else ra[i] = A[ok ? a_row[r] + a_ko[q] : 0];                               \
      a_live |= (uint32_t)ok << i;                                               \
    }                                                                            \
    _Pragma("unroll") for (int i = 0; i < EB; i++) {                             \
      const int r = C::BKM ? i : 0, q = C::BKM ? 0 : i;                          \
      const bool ok = ((b_cok >> r) & 1) && ((b_kok >> q) & 1);                  \
      if constexpr (C::BUF) rb[i] = dev::buf_load<S>(rsB, b_col[r] + b_ko[q]);  \
      else rb[i] = B[ok ? b_col[r] + b_ko[q] : 0];                               \
      b_live |= (uint32_t)ok << i;                                               \
    }                                                                            \
  }

  const bool ca = p.conj_a, cb = p.conj_b;
  auto stash = [&](int buf) {
    T* As = lds + buf * C::STAGE;
    T* Bs = As + C::PLANES * C::A_PLANE;
#pragma unroll
    for (int i = 0; i < EA; i++) {
      const bool ok = C::BUF || ((a_live >> i) & 1);  // BUF: out-of-range loads read zero
      const int e = C::aidx(a_m0 + i * a_mstep, a_k0 + i * a_kstep);
      const T re = ok ? St::re(ra[i]) : T(0);
      if constexpr (C::ILV) {
        T im = ok ? St::im(ra[i]) : T(0);
        if (ca) im = -im;
        ((dev::Cx<T>*)As)[e] = dev::Cx<T>{re, im};
        continue;
      }
      T* dst = As + e;
      dst[0] = re;
      if constexpr (C::CPLX) {
        T im = ok ? St::im(ra[i]) : T(0);
        if (ca) im = -im;
        dst[C::A_PLANE] = im;
        if constexpr (C::METH == M_3M) dst[2 * C::A_PLANE] = re + im;
      }
    }
#pragma unroll
    for (int i = 0; i < EB; i++) {
      const bool ok = C::BUF || ((b_live >> i) & 1);
      const int e = C::bidx(b_n0 + i * b_nstep, b_k0 + i * b_kstep);
      const T re = ok ? St::re(rb[i]) : T(0);
      if constexpr (C::ILV) {
        T im = ok ? St::im(rb[i]) : T(0);
        if (cb) im = -im;
        ((dev::Cx<T>*)Bs)[e] = dev::Cx<T>{re, im};
        continue;
      }
      T* dst = Bs + e;
      dst[0] = re;
      if constexpr (C::CPLX) {
        T im = ok ? St::im(rb[i]) : T(0);
        if (cb) im = -im;
        dst[C::B_PLANE] = im;
        if constexpr (C::METH == M_3M) dst[2 * C::B_PLANE] = re + im;
      }
    }
  };

  // Accumulators: [ACCS][MB][NI] (1m works on 8-row complex blocks).
  constexpr int MB = C::METH == M_1M ? 2 * C::MI : C::MI;
  Acc acc[C::ACCS][MB][C::NI];
#pragma unroll
  for (int s = 0; s < C::ACCS; s++)
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
      for (int j = 0; j < C::NI; j++) acc[s][i][j] = Acc{};

  const int lr = lane & 15, lk = lane >> 4;
  // This workgroup's K slabs: all of them, or split `split`'s share.
  const int64_t nk_all = (K + BK - 1) / BK;
  const int64_t kt0 = p.ksplit > 1 ? (int64_t)split * p.kslabs : 0;
  const int64_t nk_rest = nk_all - kt0;
  const int64_t nk = p.ksplit > 1 ? (nk_rest < p.kslabs ? (nk_rest > 0 ? nk_rest : 0) : p.kslabs) : nk_all;
  const int64_t kb0 = kt0 * BK;  // first k of the range
  TP_LOAD_KOFF(kb0);
  TP_GATHER();
  if (nk > 1) TP_LOAD_KOFF(kb0 + BK);
  stash(0);
  if (nk > 1) {
    TP_GATHER();
    if (nk > 2) TP_LOAD_KOFF(kb0 + 2 * BK);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; kt++) {
    // Slab kt is in stage kt & 1. Its MFMAs run while slab kt + 1 (gathered
    // during the previous step) is stored into the other stage and slab kt + 2
    // is gathered; one barrier per step then publishes slab kt + 1 and retires
    // stage kt & 1 for slab kt + 2.
    const T* As = lds + (kt & 1) * C::STAGE;
    const T* Bs = As + C::PLANES * C::A_PLANE;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      if constexpr (C::METH == M_REAL) {
        T a[C::MI], b[C::NI];
#pragma unroll
        for (int i = 0; i < C::MI; i++) a[i] = As[C::aidx(wm + 16 * i + lr, kk + lk)];
#pragma unroll
        for (int j = 0; j < C::NI; j++) b[j] = Bs[C::bidx(wn + 16 * j + lr, kk + lk)];
#pragma unroll
        for (int i = 0; i < C::MI; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(a[i], b[j], acc[0][i][j]);
      } else if constexpr (C::METH == M_NATIVE) {
        T ar[C::MI], ai[C::MI], br[C::NI], bi[C::NI];
        const dev::Cx<T>* Ac = (const dev::Cx<T>*)As;
        const dev::Cx<T>* Bc = (const dev::Cx<T>*)Bs;
#pragma unroll
        for (int i = 0; i < C::MI; i++) {
          const dev::Cx<T> v = Ac[C::aidx(wm + 16 * i + lr, kk + lk)];
          ar[i] = v.re;
          ai[i] = v.im;
        }
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
          const dev::Cx<T> v = Bc[C::bidx(wn + 16 * j + lr, kk + lk)];
          br[j] = v.re;
          bi[j] = v.im;
        }
        // f64: acc0 += ar br - ai bi, acc1 += ar bi + ai br. f32: acc0 += ar br,
        // acc1 += ai bi (subtracted at the end), acc2 += ar bi + ai br.
        constexpr int XI = C::NEG_MFMA ? 1 : 2;
#pragma unroll
        for (int i = 0; i < C::MI; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++) {
            acc[0][i][j] = MM::run(ar[i], br[j], acc[0][i][j]);
            acc[XI][i][j] = MM::run(ar[i], bi[j], acc[XI][i][j]);
          }
#pragma unroll
        for (int i = 0; i < C::MI; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++) {
            if constexpr (C::NEG_MFMA) acc[0][i][j] = MM::run_neg_a(ai[i], bi[j], acc[0][i][j]);
            else acc[1][i][j] = MM::run(ai[i], bi[j], acc[1][i][j]);
            acc[XI][i][j] = MM::run(ai[i], br[j], acc[XI][i][j]);
          }
      } else if constexpr (C::METH == M_3M) {
        T a[3][C::MI], b[3][C::NI];
#pragma unroll
        for (int s = 0; s < 3; s++) {
#pragma unroll
          for (int i = 0; i < C::MI; i++) a[s][i] = As[s * C::A_PLANE + C::aidx(wm + 16 * i + lr, kk + lk)];
#pragma unroll
          for (int j = 0; j < C::NI; j++) b[s][j] = Bs[s * C::B_PLANE + C::bidx(wn + 16 * j + lr, kk + lk)];
        }
#pragma unroll
        for (int s = 0; s < 3; s++)
#pragma unroll
          for (int i = 0; i < C::MI; i++)
#pragma unroll
            for (int j = 0; j < C::NI; j++) acc[s][i][j] = MM::run(a[s][i], b[s][j], acc[s][i][j]);
      } else {  // 1m: real product of the 1e-expanded A (2 rows x 2 k per element) and 1r-expanded B
        const int r = lr & 1;  // 0: real row, 1: imaginary row of the 1e block
        const int s = lk & 1;  // 0: real k, 1: imaginary k
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
          const int ck = kk + 2 * kc + (lk >> 1);
          T a[2 * C::MI], b[C::NI];
#pragma unroll
          for (int i = 0; i < 2 * C::MI; i++) {
            const int x = C::aidx(wm + 8 * i + (lr >> 1), ck);
            const T xr = As[x], xi = As[C::A_PLANE + x];
            a[i] = (r == s) ? xr : (r == 0 ? -xi : xi);
          }
#pragma unroll
          for (int j = 0; j < C::NI; j++) {
            const int x = C::bidx(wn + 16 * j + lr, ck);
            b[j] = s == 0 ? Bs[x] : Bs[C::B_PLANE + x];
          }
#pragma unroll
          for (int i = 0; i < 2 * C::MI; i++)
#pragma unroll
            for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(a[i], b[j], acc[0][i][j]);
        }
      }
    }
    // Unconditional (branch-free, so the compiler can interleave them with
    // the MFMAs): past the last slab the store rewrites the idle stage with
    // data nobody reads, and the offsets and gathers fall back to every
    // lane's safe address (K offsets clamp to K - 1, gathers to element 0).
    stash((int)((kt + 1) & 1));
    TP_GATHER();
    TP_LOAD_KOFF(kb0 + (kt + 3) * BK);
    __syncthreads();
  }
#undef TP_LOAD_KOFF
#undef TP_GATHER

  // Stage the tile chunk by chunk (Dre/Dim [EPI_ROWS][LDD]) and write it back.
  constexpr int ER = C::EPI_ROWS;
  T* Dre = (T*)smem;
  T* Dim = Dre + ER * C::LDD;
#pragma unroll 1
  for (int c = 0; c < BM / ER; c++) {
    if (c > 0) __syncthreads();  // the previous chunk's write-back has read the staging rows
    if (wm / ER == c) {
      const int wr = wm - c * ER;  // this wave's first row inside the chunk
#pragma unroll
  for (int i = 0; i < MB; i++)
#pragma unroll
    for (int j = 0; j < C::NI; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int col = wn + 16 * j + lr;
        const int rr = MM::row(lk, q);  // row inside the 16x16 block
        if constexpr (C::METH == M_1M) {
          // real row rr of the 1e product is complex row rr / 2, component rr % 2
          const int row = wr + 8 * i + (rr >> 1);
          if (rr & 1) Dim[row * C::LDD + col] = acc[0][i][j][q];
          else Dre[row * C::LDD + col] = acc[0][i][j][q];
        } else {
          const int row = wr + 16 * i + rr;
          if constexpr (C::METH == M_REAL) {
            Dre[row * C::LDD + col] = acc[0][i][j][q];
          } else if constexpr (C::METH == M_NATIVE && C::NEG_MFMA) {
            Dre[row * C::LDD + col] = acc[0][i][j][q];
            Dim[row * C::LDD + col] = acc[1][i][j][q];
          } else if constexpr (C::METH == M_NATIVE) {
            Dre[row * C::LDD + col] = acc[0][i][j][q] - acc[1][i][j][q];
            Dim[row * C::LDD + col] = acc[2][i][j][q];
          } else {
            const T m1 = acc[0][i][j][q], m2 = acc[1][i][j][q], m3 = acc[2][i][j][q];
            Dre[row * C::LDD + col] = m1 - m2;
            Dim[row * C::LDD + col] = m3 - m1 - m2;
          }
        }
      }
    }
    __syncthreads();
    if (p.ksplit > 1)
      dev::write_partial<T, C::CPLX, ER, BN, NT, C::LDD>(p, split, h, m0 + c * ER, n0, Dre, Dim);
    else
      // Element-wise write-back: the vector path costs these kernels registers
      // (occupancy of the grouped 32x32 bodies 2 -> 1 waves per SIMD).
      dev::write_back<T, T, C::CPLX, ER, BN, NT, C::LDD, false>(p, h, m0 + c * ER, n0, Dre, Dim);
  }
}

// One tile per workgroup; a grid capped below the launch limit strides over
// the remaining tiles.
template <class C>
__global__ __launch_bounds__(C::NT) void k_contract(ContractArgs p) {
  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  // With split K the splits of a tile are ksplit consecutive ids.
  const int S = p.ksplit > 1 ? p.ksplit : 1;
  for (int64_t t = blockIdx.x; t < nwg * S; t += gridDim.x) {
    if (t != (int64_t)blockIdx.x) __syncthreads();  // the previous tile's write-back has read LDS
    int64_t h;
    int tm, tn;
    dev::decode_tile(t / S, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
    contract_tile<C>(p, h, tm, tn, smem, (int)(t % S));
  }
}

template <typename T, int METH, int BM, int BN, int BK, int WM, int WN, bool BUF>
struct Majors {
  typedef Cfg<T, METH, BM, BN, BK, WM, WN, true, true, BUF> KK;
  typedef Cfg<T, METH, BM, BN, BK, WM, WN, true, false, BUF> KN;
  typedef Cfg<T, METH, BM, BN, BK, WM, WN, false, true, BUF> MK;
  typedef Cfg<T, METH, BM, BN, BK, WM, WN, false, false, BUF> MN;
};

// The eight gather variants of one tile: operand majors x offset width.
template <typename T, int METH, int BM, int BN, int BK, int WM, int WN>
struct Family {
  typedef Majors<T, METH, BM, BN, BK, WM, WN, false> W64;
  typedef Majors<T, METH, BM, BN, BK, WM, WN, true> W32;
  typedef typename W64::KK KK;
  typedef typename W64::KN KN;
  typedef typename W64::MK MK;
  typedef typename W64::MN MN;
  static constexpr int NT = KK::NT;
  static constexpr size_t LDS = KK::LDS > MN::LDS ? (KK::LDS > KN::LDS ? KK::LDS : KN::LDS)
                                                  : (MN::LDS > MK::LDS ? MN::LDS : MK::LDS);
};

template <class W>
__device__ inline void tile_by_major(const ContractArgs& p, int64_t h, int tm, int tn, char* smem) {
  if (p.a_kmajor) {
    if (p.b_kmajor) contract_tile<typename W::KK>(p, h, tm, tn, smem);
    else contract_tile<typename W::KN>(p, h, tm, tn, smem);
  } else {
    if (p.b_kmajor) contract_tile<typename W::MK>(p, h, tm, tn, smem);
    else contract_tile<typename W::MN>(p, h, tm, tn, smem);
  }
}

// Persistent grouped kernel: workgroups stride over the flattened tile space of
// every item; each tile finds its item by binary search over tile_begin and
// runs the gather variant that item's layout asks for.
template <class F>
__global__ __launch_bounds__(F::NT) void k_grouped(const GroupItem* __restrict__ items, int num_items,
                                                   int64_t total_tiles) {
  __shared__ __attribute__((aligned(16))) char smem[F::LDS];
  for (int64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    int lo = 0, hi = num_items - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (items[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    const ContractArgs& p = items[lo].args;
    const int64_t local = t - items[lo].tile_begin;
    const int64_t per_h = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t h = local / per_h;
    const int r = (int)(local - h * per_h);
    const int tm = r % p.tiles_m, tn = r / p.tiles_m;
    if (p.off32) tile_by_major<typename F::W32>(p, h, tm, tn, smem);
    else tile_by_major<typename F::W64>(p, h, tm, tn, smem);
    __syncthreads();
  }
}

template <class C>
hipError_t launch_cfg(const ContractArgs& a, hipStream_t s) {
  ContractArgs p = a;
  p.tiles_m = (int)((p.M + C::BM - 1) / C::BM);
  p.tiles_n = (int)((p.N + C::BN - 1) / C::BN);
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H * (p.ksplit > 1 ? p.ksplit : 1);
  if (nwg == 0) return hipSuccess;
  const int64_t grid = nwg < max_grid(C::NT) ? nwg : max_grid(C::NT);
  hipLaunchKernelGGL(k_contract<C>, dim3((unsigned)grid), dim3(C::NT), 0, s, p);
  return hipGetLastError();
}

template <class W>
hipError_t launch_majors(const ContractArgs& a, hipStream_t s) {
  if (a.a_kmajor) {
    if (a.b_kmajor) return launch_cfg<typename W::KK>(a, s);
    return launch_cfg<typename W::KN>(a, s);
  }
  if (a.b_kmajor) return launch_cfg<typename W::MK>(a, s);
  return launch_cfg<typename W::MN>(a, s);
}

template <class F>
hipError_t launch_family(const ContractArgs& a, hipStream_t s) {
  if constexpr (F::LDS > 160 * 1024) {
    return hipErrorInvalidValue;  // shape/method combination the planner never picks
  } else {
    return a.off32 ? launch_majors<typename F::W32>(a, s) : launch_majors<typename F::W64>(a, s);
  }
}

template <class F>
hipError_t launch_group(const GroupItem* items, int num_items, int64_t total, int grid, hipStream_t s) {
  if constexpr (F::LDS > 160 * 1024) {
    return hipErrorInvalidValue;  // shape/method combination the planner never picks
  } else {
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grouped<F>, dim3((unsigned)grid), dim3(F::NT), 0, s, items, num_items, total);
    return hipGetLastError();
  }
}

// Tile shapes (BM, BN, BK, WM, WN) per tile id.
#define TP_TILE_LIST(X)                 \
  X(TILE_64x64, 64, 64, 16, 32, 32)     \
  X(TILE_32x32, 32, 32, 16, 32, 32)     \
  X(TILE_128x64, 128, 64, 16, 64, 32)

template <typename T, int METH>
hipError_t dispatch_tile(const ContractArgs& a, int tile_id, hipStream_t s) {
  switch (tile_id) {
#define TP_CASE(id, bm, bn, bk, wm, wn) \
  case id: return launch_family<Family<T, METH, bm, bn, bk, wm, wn>>(a, s);
    TP_TILE_LIST(TP_CASE)
#undef TP_CASE
  }
  return hipErrorInvalidValue;
}

template <typename T, int METH>
hipError_t dispatch_group(const GroupItem* items, int n, int64_t total, int tile_id, int grid, hipStream_t s) {
  switch (tile_id) {
#define TP_CASE(id, bm, bn, bk, wm, wn) \
  case id: return launch_group<Family<T, METH, bm, bn, bk, wm, wn>>(items, n, total, grid, s);
    TP_TILE_LIST(TP_CASE)
#undef TP_CASE
  }
  return hipErrorInvalidValue;
}

}  // namespace

const TileCfg* tile_cfg(int id) {
  static const TileCfg tiles[] = {
#define TP_ROW(id, bm, bn, bk, wm, wn) {id, bm, bn, bk, (bm / wm) * (bn / wn) * 64, #id},
      TP_TILE_LIST(TP_ROW)
#undef TP_ROW
      {TILE_128x128, 128, 128, 16, 512, "TILE_128x128"},  // c64: 1024 threads
      {TILE_32x32K, 32, 32, 16, 256, "TILE_32x32K"},
      {TILE_32x32R, 32, 32, 16, 512, "TILE_32x32R"},
      {TILE_32x48R, 32, 48, 16, 512, "TILE_32x48R"},
      {STILE_128x128, 128, 128, 32, 256, "STILE_128x128"},
      {STILE_64x64, 64, 64, 32, 256, "STILE_64x64"},
      {STILE_256x128, 256, 128, 32, 512, "STILE_256x128"},
      {STILE_128x64, 128, 64, 32, 512, "STILE_128x64"},
      {STILE_256x256, 256, 256, 128, 512, "STILE_256x256"},
  };
  for (auto& t : tiles)
    if (t.id == id) return &t;
  return nullptr;
}

hipError_t launch_contract(const ContractArgs& a, DType t, int method, int tile_id, hipStream_t s) {
  if (dma_supported(a.off32, t, method, tile_id)) {
    // The DMA kernel runs one tile (or one K range of one) per workgroup,
    // within the launch limit.
    const TileCfg* c = tile_cfg(tile_id);
    const int64_t nwg =
        ((a.M + c->bm - 1) / c->bm) * ((a.N + c->bn - 1) / c->bn) * a.H * (a.ksplit > 1 ? a.ksplit : 1);
    if (nwg <= max_grid(c->threads)) return launch_contract_dma(a, t, method, tile_id, s);
  }
  // The K-group tiles run where their gathers fit (the register-direct ones
  // as the LDS-DMA K-group tile otherwise), else as the register-staged
  // 32x32 tile; the DMA-only 128x128 tile falls back to the register-staged
  // 64x64 kernel.
  if (tile_id == TILE_32x32R || tile_id == TILE_32x48R) {
    const bool wide = tile_id == TILE_32x48R && a.b_kmajor;
    if (a.ksplit <= 1 && direct_supported(a.off32, a.ab_vec, t) &&
        ((a.M + 31) / 32) * ((a.N + 31) / 32) * a.H <= max_grid(512))
      return launch_contract_direct(a, t, wide, s);
    tile_id = TILE_32x32K;
  }
  if (tile_id == TILE_32x32K) {
    if (a.ksplit <= 1 && kgroup_supported(a.off32, a.ab_vec, a.a_kmajor, a.b_kmajor, t) &&
        ((a.M + 31) / 32) * ((a.N + 31) / 32) * a.H <= max_grid(256))
      return launch_contract_kgroup(a, t, method, s);
    tile_id = TILE_32x32;
  }
  if (tile_id == TILE_128x128) tile_id = TILE_64x64;
  switch (t) {
    case DType::F64: return dispatch_tile<double, M_REAL>(a, tile_id, s);
    case DType::F32: return dispatch_tile<float, M_REAL>(a, tile_id, s);
    case DType::C64:
      if (method == CPLX_3M) return dispatch_tile<double, M_3M>(a, tile_id, s);
      if (method == CPLX_1M) return dispatch_tile<double, M_1M>(a, tile_id, s);
      return dispatch_tile<double, M_NATIVE>(a, tile_id, s);
    case DType::C32:
      if (method == CPLX_3M) return dispatch_tile<float, M_3M>(a, tile_id, s);
      if (method == CPLX_1M) return dispatch_tile<float, M_1M>(a, tile_id, s);
      return dispatch_tile<float, M_NATIVE>(a, tile_id, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_grouped(const GroupItem* items, int n, int64_t total, DType t, int method, int tile_id, int grid,
                          hipStream_t s) {
  switch (t) {
    case DType::F64: return dispatch_group<double, M_REAL>(items, n, total, tile_id, grid, s);
    case DType::F32: return dispatch_group<float, M_REAL>(items, n, total, tile_id, grid, s);
    case DType::C64:
      if (method == CPLX_3M) return dispatch_group<double, M_3M>(items, n, total, tile_id, grid, s);
      if (method == CPLX_1M) return dispatch_group<double, M_1M>(items, n, total, tile_id, grid, s);
      return dispatch_group<double, M_NATIVE>(items, n, total, tile_id, grid, s);
    case DType::C32:
      if (method == CPLX_3M) return dispatch_group<float, M_3M>(items, n, total, tile_id, grid, s);
      if (method == CPLX_1M) return dispatch_group<float, M_1M>(items, n, total, tile_id, grid, s);
      return dispatch_group<float, M_NATIVE>(items, n, total, tile_id, grid, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tp
