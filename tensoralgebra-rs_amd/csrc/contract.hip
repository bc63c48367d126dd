// Storage-precision contraction kernels for gfx950 (f32/f64 real, c32/c64
// complex), the GPU form of the reference's packed five-loop driver:
//
//   crates/tprims-kernel/src/pack/pack.rs:57-265   packing (block-scatter gather,
//                                                  conj folded, planar / 3m planes)
//   crates/tprims-kernel/src/kernels/reference/    microkernel semantics
//     scalar.rs:31 real, :64 planar, :110 1m, :136 3m
//   crates/tprims-kernel/src/pack/writeback.rs:114-495  write-back
//     op_D(alpha * AB + beta * op_C(C)); beta == 0 never reads C
//   crates/tprims-contract/src/driver/tile.rs:142     K accumulation
//
// One workgroup computes one BM x BN tile of one batch entry h over the whole
// K range (K is never split, so the result does not depend on the grid). Each
// K step gathers a BM x BK slab of A and a BK x BN slab of B straight from HBM
// through the scatter vectors into LDS planes (K-major or M/N-major lane
// order, whichever the planner found contiguous), folding conjugation and the
// 3m sum plane into the store. Waves run 16x16x4 MFMAs (f64 or f32) from the
// planes. The next slab is fetched into registers while the current one is
// multiplied. The tile is staged through LDS and written back along D's
// contiguous axis, with alpha, beta, op_C and op_D fused.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tp {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;
// Output placement of one 16x16 accumulator block: lane (lk = lane / 16,
// column lane % 16) holds four rows. The f64 instruction interleaves them
// (row lk + 4q); the f32 one keeps them together (row 4 lk + q). Both were
// measured on gfx950 with known inputs.
template <>
struct Mma<double> {
  typedef v4d acc;
  __device__ static inline acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int q) { return lk + 4 * q; }
};
template <>
struct Mma<float> {
  typedef v4f acc;
  __device__ static inline acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int q) { return 4 * lk + q; }
};

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T re, im;
};

enum { M_REAL = -1, M_NATIVE = 0, M_1M = 1, M_3M = 2 };

template <typename T, int METHOD, int BM_, int BN_, int BK_, int WM_, int WN_>
struct Cfg {
  typedef T Tv;
  static constexpr int METH = METHOD;
  static constexpr bool CPLX = METHOD != M_REAL;
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int NT = (BM / WM) * (BN / WN) * 64;
  static constexpr int PLANES = METHOD == M_REAL ? 1 : (METHOD == M_3M ? 3 : 2);
  static constexpr int LDA = BM + 16;  // row pitch of an A plane (elements)
  static constexpr int LDB = BN + 16;
  static constexpr int LDD = BN + 1;   // row pitch of the staging tile
  static constexpr int EA = BM * BK / NT;
  static constexpr int EB = BN * BK / NT;
  static constexpr int EPI = BM * BN / NT;
  static constexpr int MI = WM / 16, NI = WN / 16;
  static constexpr int ACCS = METHOD == M_REAL ? 1 : (METHOD == M_3M ? 3 : 2);
  static constexpr size_t LDS_MAIN = (size_t)PLANES * BK * (LDA + LDB) * sizeof(T);
  static constexpr size_t LDS_EPI = (size_t)(CPLX ? 2 : 1) * BM * LDD * sizeof(T);
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  static_assert(NT % BK == 0 && BM % (NT / BK) == 0 && BN % (NT / BK) == 0, "K-major lane map");
  static_assert(NT % BM == 0 && NT % BN == 0 && BK % (NT / BM) == 0 && BK % (NT / BN) == 0, "MN-major lane map");
  static_assert(BK % 4 == 0 && WM % 16 == 0 && WN % 16 == 0, "MFMA shape");
};

// Storage element of a dtype.
template <typename T, bool CPLX>
struct Store;
template <typename T>
struct Store<T, false> {
  typedef T S;
  __device__ static inline T re(S v) { return v; }
  __device__ static inline T im(S) { return T(0); }
};
template <typename T>
struct Store<T, true> {
  typedef Cx<T> S;
  __device__ static inline T re(S v) { return v.re; }
  __device__ static inline T im(S v) { return v.im; }
};

// Workgroup id -> (h, tile_m, tile_n). Consecutive hardware ids go round-robin
// over the 8 XCDs; remapping gives each XCD a contiguous run of logical tiles,
// and the logical order walks groups of 8 tile-rows so neighbours share A rows
// and B columns in that XCD's L2.
__device__ inline void decode_tile(int64_t bid, int64_t nwg, int tiles_m, int tiles_n, int64_t& h, int& tm,
                                   int& tn) {
  if ((nwg & 7) == 0) bid = (bid & 7) * (nwg >> 3) + (bid >> 3);
  const int64_t per_h = (int64_t)tiles_m * tiles_n;
  h = bid / per_h;
  int t = (int)(bid - h * per_h);
  const int GM = 8;
  int group = t / (GM * tiles_n);
  int first_m = group * GM;
  int gsize = min(tiles_m - first_m, GM);
  int r = t - group * GM * tiles_n;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

// One tile of one contraction, the body shared by the single-contraction
// kernel and the grouped persistent kernel.
template <class C>
__device__ inline void contract_tile(const ContractArgs& p, int64_t h, int tm, int tn, char* smem) {
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

  T* As = (T*)smem;                           // [PLANES][BK][LDA]
  T* Bs = As + C::PLANES * BK * C::LDA;       // [PLANES][BK][LDB]

  // Per-thread gather coordinates (local m/k of each element it loads).
  const bool akm = p.a_kmajor, bkm = p.b_kmajor;
  int a_ml[EA], a_kl[EA], b_nl[EB], b_kl[EB];
#pragma unroll
  for (int i = 0; i < EA; i++) {
    if (akm) {
      a_kl[i] = tid % BK;
      a_ml[i] = tid / BK + i * (NT / BK);
    } else {
      a_ml[i] = tid % BM;
      a_kl[i] = tid / BM + i * (NT / BM);
    }
  }
#pragma unroll
  for (int i = 0; i < EB; i++) {
    if (bkm) {
      b_kl[i] = tid % BK;
      b_nl[i] = tid / BK + i * (NT / BK);
    } else {
      b_nl[i] = tid % BN;
      b_kl[i] = tid / BN + i * (NT / BN);
    }
  }
  int64_t a_row[EA], b_col[EB];
  bool a_ok[EA], b_ok[EB];
#pragma unroll
  for (int i = 0; i < EA; i++) {
    int64_t m = m0 + a_ml[i];
    a_ok[i] = m < M;
    a_row[i] = a_ok[i] ? p.am[m] : 0;
  }
#pragma unroll
  for (int i = 0; i < EB; i++) {
    int64_t n = n0 + b_nl[i];
    b_ok[i] = n < N;
    b_col[i] = b_ok[i] ? p.bn[n] : 0;
  }

  S ra[EA], rb[EB];
  auto fetch = [&](int64_t k0) {
    const S zero = S{};
    if (akm) {
      int64_t k = k0 + a_kl[0];
      bool kok = k < K;
      int64_t ko = kok ? p.ak[k] : 0;
#pragma unroll
      for (int i = 0; i < EA; i++) ra[i] = (kok && a_ok[i]) ? A[a_row[i] + ko] : zero;
    } else {
#pragma unroll
      for (int i = 0; i < EA; i++) {
        int64_t k = k0 + a_kl[i];
        ra[i] = (k < K && a_ok[i]) ? A[a_row[i] + p.ak[k]] : zero;
      }
    }
    if (bkm) {
      int64_t k = k0 + b_kl[0];
      bool kok = k < K;
      int64_t ko = kok ? p.bk[k] : 0;
#pragma unroll
      for (int i = 0; i < EB; i++) rb[i] = (kok && b_ok[i]) ? B[b_col[i] + ko] : zero;
    } else {
#pragma unroll
      for (int i = 0; i < EB; i++) {
        int64_t k = k0 + b_kl[i];
        rb[i] = (k < K && b_ok[i]) ? B[b_col[i] + p.bk[k]] : zero;
      }
    }
  };

  const bool ca = p.conj_a, cb = p.conj_b;
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < EA; i++) {
      T* dst = As + a_kl[i] * C::LDA + a_ml[i];
      T re = St::re(ra[i]);
      dst[0] = re;
      if constexpr (C::CPLX) {
        T im = ca ? -St::im(ra[i]) : St::im(ra[i]);
        dst[BK * C::LDA] = im;
        if constexpr (C::METH == M_3M) dst[2 * BK * C::LDA] = re + im;
      }
    }
#pragma unroll
    for (int i = 0; i < EB; i++) {
      T* dst = Bs + b_kl[i] * C::LDB + b_nl[i];
      T re = St::re(rb[i]);
      dst[0] = re;
      if constexpr (C::CPLX) {
        T im = cb ? -St::im(rb[i]) : St::im(rb[i]);
        dst[BK * C::LDB] = im;
        if constexpr (C::METH == M_3M) dst[2 * BK * C::LDB] = re + im;
      }
    }
  };

  // Accumulators: [ACCS][MI or 2*MI for 1m][NI].
  constexpr int MB = C::METH == M_1M ? 2 * C::MI : C::MI;
  Acc acc[C::ACCS][MB][C::NI];
#pragma unroll
  for (int s = 0; s < C::ACCS; s++)
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
      for (int j = 0; j < C::NI; j++) acc[s][i][j] = Acc{};

  const int lr = lane & 15, lk = lane >> 4;
  const int64_t nk = (K + BK - 1) / BK;
  fetch(0);
  for (int64_t kt = 0; kt < nk; kt++) {
    stash();
    __syncthreads();
    if (kt + 1 < nk) fetch((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      if constexpr (C::METH == M_REAL) {
        T a[C::MI], b[C::NI];
#pragma unroll
        for (int i = 0; i < C::MI; i++) a[i] = As[(kk + lk) * C::LDA + wm + 16 * i + lr];
#pragma unroll
        for (int j = 0; j < C::NI; j++) b[j] = Bs[(kk + lk) * C::LDB + wn + 16 * j + lr];
#pragma unroll
        for (int i = 0; i < C::MI; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(a[i], b[j], acc[0][i][j]);
      } else if constexpr (C::METH == M_NATIVE) {
        T ar[C::MI], ai[C::MI], br[C::NI], bi[C::NI];
#pragma unroll
        for (int i = 0; i < C::MI; i++) {
          const T* src = As + (kk + lk) * C::LDA + wm + 16 * i + lr;
          ar[i] = src[0];
          ai[i] = src[BK * C::LDA];
        }
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
          const T* src = Bs + (kk + lk) * C::LDB + wn + 16 * j + lr;
          br[j] = src[0];
          bi[j] = src[BK * C::LDB];
        }
#pragma unroll
        for (int i = 0; i < C::MI; i++)
#pragma unroll
          for (int j = 0; j < C::NI; j++) {
            acc[0][i][j] = MM::run(ar[i], br[j], acc[0][i][j]);
            acc[0][i][j] = MM::run(-ai[i], bi[j], acc[0][i][j]);
            acc[1][i][j] = MM::run(ar[i], bi[j], acc[1][i][j]);
            acc[1][i][j] = MM::run(ai[i], br[j], acc[1][i][j]);
          }
      } else if constexpr (C::METH == M_3M) {
        T a[3][C::MI], b[3][C::NI];
#pragma unroll
        for (int s = 0; s < 3; s++) {
#pragma unroll
          for (int i = 0; i < C::MI; i++) a[s][i] = As[s * BK * C::LDA + (kk + lk) * C::LDA + wm + 16 * i + lr];
#pragma unroll
          for (int j = 0; j < C::NI; j++) b[s][j] = Bs[s * BK * C::LDB + (kk + lk) * C::LDB + wn + 16 * j + lr];
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
            const T* src = As + ck * C::LDA + wm + 8 * i + (lr >> 1);
            T xr = src[0], xi = src[BK * C::LDA];
            a[i] = (r == s) ? xr : (r == 0 ? -xi : xi);
          }
#pragma unroll
          for (int j = 0; j < C::NI; j++) {
            const T* src = Bs + ck * C::LDB + wn + 16 * j + lr;
            b[j] = s == 0 ? src[0] : src[BK * C::LDB];
          }
#pragma unroll
          for (int i = 0; i < 2 * C::MI; i++)
#pragma unroll
            for (int j = 0; j < C::NI; j++) acc[0][i][j] = MM::run(a[i], b[j], acc[0][i][j]);
        }
      }
    }
    __syncthreads();
  }

  // Stage the tile: Dre/Dim [BM][LDD].
  T* Dre = (T*)smem;
  T* Dim = Dre + BM * C::LDD;
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
          const int row = wm + 8 * i + (rr >> 1);
          if (rr & 1) Dim[row * C::LDD + col] = acc[0][i][j][q];
          else Dre[row * C::LDD + col] = acc[0][i][j][q];
        } else {
          const int row = wm + 16 * i + rr;
          if constexpr (C::METH == M_REAL) {
            Dre[row * C::LDD + col] = acc[0][i][j][q];
          } else if constexpr (C::METH == M_NATIVE) {
            Dre[row * C::LDD + col] = acc[0][i][j][q];
            Dim[row * C::LDD + col] = acc[1][i][j][q];
          } else {
            T m1 = acc[0][i][j][q], m2 = acc[1][i][j][q], m3 = acc[2][i][j][q];
            Dre[row * C::LDD + col] = m1 - m2;
            Dim[row * C::LDD + col] = m3 - m1 - m2;
          }
        }
      }
  __syncthreads();

  // Write-back along D's contiguous axis.
  S* __restrict__ D = (S*)p.D + p.hd[h];
  const S* __restrict__ Cp = (const S*)p.C + p.hc[h];
  const T alr = (T)p.alpha[0], ali = (T)p.alpha[1];
  const T ber = (T)p.beta[0], bei = (T)p.beta[1];
  const bool rc = p.read_c, cc = p.conj_c, cd = p.conj_d;
  const int64_t R = p.R;
#pragma unroll 4
  for (int t = 0; t < C::EPI; t++) {
    const int e = tid + t * NT;
    int ml, nl;
    if (p.d_mmajor) {
      ml = e % BM;
      nl = e / BM;
    } else {
      ml = e / BN;
      nl = e % BN;
    }
    const int64_t m = m0 + ml, n = n0 + nl;
    if (m >= M || n >= N) continue;
    T xr = Dre[ml * C::LDD + nl];
    T xi = C::CPLX ? Dim[ml * C::LDD + nl] : T(0);
    T vr = alr * xr - ali * xi;
    T vi = alr * xi + ali * xr;
    if (rc) {
      S cv = Cp[p.cm[m] + p.cn[n]];
      T cr = St::re(cv), ci = St::im(cv);
      if (cc) ci = -ci;
      vr += ber * cr - bei * ci;
      vi += ber * ci + bei * cr;
    }
    if (cd) vi = -vi;
    S out;
    if constexpr (C::CPLX) {
      out.re = vr;
      out.im = vi;
    } else {
      out = vr;
    }
    const int64_t off = p.dm[m] + p.dn[n];
    if (p.dr == nullptr) {
      D[off] = out;
    } else {
      for (int64_t r = 0; r < R; r++) D[off + p.dr[r]] = out;
    }
  }
}

template <class C>
__global__ __launch_bounds__(C::NT) void k_contract(ContractArgs p) {
  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  int64_t h;
  int tm, tn;
  decode_tile(blockIdx.x, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  contract_tile<C>(p, h, tm, tn, smem);
}

// Persistent grouped kernel: workgroups stride over the flattened tile space of
// every item; each tile finds its item by binary search over tile_begin.
template <class C>
__global__ __launch_bounds__(C::NT) void k_grouped(const GroupItem* __restrict__ items, int num_items,
                                                   int64_t total_tiles) {
  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
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
    contract_tile<C>(p, h, tm, tn, smem);
    __syncthreads();
  }
}

template <class C>
hipError_t launch_one(const ContractArgs& a, hipStream_t s) {
  ContractArgs p = a;
  p.tiles_m = (int)((p.M + C::BM - 1) / C::BM);
  p.tiles_n = (int)((p.N + C::BN - 1) / C::BN);
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  if (nwg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_contract<C>, dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
  return hipGetLastError();
}

template <class C>
hipError_t launch_group(const GroupItem* items, int num_items, int64_t total, int grid, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grouped<C>, dim3((unsigned)grid), dim3(C::NT), 0, s, items, num_items, total);
  return hipGetLastError();
}

// Tile shapes (BM, BN, BK, WM, WN) per tile id.
#define TP_TILE_LIST(X)          \
  X(TILE_64x64, 64, 64, 16, 32, 32) \
  X(TILE_32x32, 32, 32, 16, 32, 32) \
  X(TILE_128x64, 128, 64, 16, 64, 32) \
  X(TILE_64x32, 64, 32, 16, 32, 32)

template <typename T, int METH>
hipError_t dispatch_tile(const ContractArgs& a, int tile_id, hipStream_t s) {
  switch (tile_id) {
#define TP_CASE(id, bm, bn, bk, wm, wn) \
  case id: return launch_one<Cfg<T, METH, bm, bn, bk, wm, wn>>(a, s);
    TP_TILE_LIST(TP_CASE)
#undef TP_CASE
  }
  return hipErrorInvalidValue;
}

template <typename T, int METH>
hipError_t dispatch_group(const GroupItem* items, int n, int64_t total, int tile_id, int grid, hipStream_t s) {
  switch (tile_id) {
#define TP_CASE(id, bm, bn, bk, wm, wn) \
  case id: return launch_group<Cfg<T, METH, bm, bn, bk, wm, wn>>(items, n, total, grid, s);
    TP_TILE_LIST(TP_CASE)
#undef TP_CASE
  }
  return hipErrorInvalidValue;
}

template <typename T, bool CPLX>
__global__ void k_output_pass(ContractArgs p) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int64_t total = p.M * p.N * p.H;
  const T ber = (T)p.beta[0], bei = (T)p.beta[1];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t m, n, h;
    if (p.d_mmajor) {
      m = e % p.M;
      int64_t r = e / p.M;
      n = r % p.N;
      h = r / p.N;
    } else {
      n = e % p.N;
      int64_t r = e / p.N;
      m = r % p.M;
      h = r / p.M;
    }
    T vr = 0, vi = 0;
    if (p.read_c) {
      S cv = ((const S*)p.C)[p.hc[h] + p.cm[m] + p.cn[n]];
      T cr = St::re(cv), ci = St::im(cv);
      if (p.conj_c) ci = -ci;
      vr = ber * cr - bei * ci;
      vi = ber * ci + bei * cr;
    }
    if (p.conj_d) vi = -vi;
    S out;
    if constexpr (CPLX) {
      out.re = vr;
      out.im = vi;
    } else {
      out = vr;
    }
    S* D = (S*)p.D;
    const int64_t off = p.hd[h] + p.dm[m] + p.dn[n];
    if (p.dr == nullptr) D[off] = out;
    else
      for (int64_t r = 0; r < p.R; r++) D[off + p.dr[r]] = out;
  }
}

struct ZeroLayout {
  int rank;
  int64_t dims[8];
  int64_t strides[8];
  int64_t total;
};

template <typename S>
__global__ void k_zero_layout(S* base, ZeroLayout z) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < z.total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = e, off = 0;
    for (int d = 0; d < z.rank; d++) {
      int64_t i = rem % z.dims[d];
      rem /= z.dims[d];
      off += i * z.strides[d];
    }
    base[off] = S{};
  }
}

}  // namespace

const TileCfg* tile_cfg(int id) {
  static const TileCfg tiles[] = {
#define TP_ROW(id, bm, bn, bk, wm, wn) {id, bm, bn, bk, (bm / wm) * (bn / wn) * 64, #id},
      TP_TILE_LIST(TP_ROW)
#undef TP_ROW
      {STILE_128x128, 128, 128, 32, 256, "STILE_128x128"},
      {STILE_64x64, 64, 64, 32, 256, "STILE_64x64"},
  };
  for (auto& t : tiles)
    if (t.id == id) return &t;
  return nullptr;
}

hipError_t launch_contract(const ContractArgs& a, DType t, int method, int tile_id, hipStream_t s) {
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

hipError_t launch_output_pass(const ContractArgs& a, DType t, hipStream_t s) {
  const int64_t total = a.M * a.N * a.H;
  if (total == 0) return hipSuccess;
  const int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case DType::F64: hipLaunchKernelGGL((k_output_pass<double, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::F32: hipLaunchKernelGGL((k_output_pass<float, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C64: hipLaunchKernelGGL((k_output_pass<double, true>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C32: hipLaunchKernelGGL((k_output_pass<float, true>), dim3(blocks), dim3(threads), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_zero_layout(void* base, DType t, int rank, const int64_t* dims, const int64_t* strides,
                              hipStream_t s) {
  if (rank > 8) return hipErrorInvalidValue;
  ZeroLayout z;
  z.rank = rank;
  z.total = 1;
  for (int i = 0; i < rank; i++) {
    z.dims[i] = dims[i];
    z.strides[i] = strides[i];
    z.total *= dims[i];
  }
  if (z.total == 0) return hipSuccess;
  const int threads = 256;
  int64_t blocks = (z.total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case DType::F32: hipLaunchKernelGGL(k_zero_layout<float>, dim3(blocks), dim3(threads), 0, s, (float*)base, z); break;
    case DType::F64:
    case DType::C32: hipLaunchKernelGGL(k_zero_layout<double>, dim3(blocks), dim3(threads), 0, s, (double*)base, z); break;
    case DType::C64:
      hipLaunchKernelGGL(k_zero_layout<Cx<double>>, dim3(blocks), dim3(threads), 0, s, (Cx<double>*)base, z);
      break;
  }
  return hipGetLastError();
}

}  // namespace tp
