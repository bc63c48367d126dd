// Split-BF16 contraction kernels for f32/c32 storage on gfx950's bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation). No counterpart exists in
// the reference (crates/tprims-capi/src/product.rs:132 refuses any precision
// other than the storage one); the gather, conjugation and write-back
// semantics are those of the storage kernels.
//
// Every f32 value x entering LDS is split on the fly, in registers, into
// bf16 terms:  hi = bf16(x), r = x - hi (exact in f32), mid = bf16(r),
// lo = bf16(r - mid). Nothing split is ever written to HBM.
//   2 terms (hi, lo = bf16(x - hi)):  a*b ~ ah*bh + ah*bl + al*bh           (3 MFMAs)
//   3 terms (hi, mid, lo):            + am*bm + ah*bl + al*bh on top        (6 MFMAs)
// The dropped products are below 2^-16 (2 terms) or 2^-24 (3 terms) of |a||b|.
// A complex product takes four real ones (Re = rr - ii, Im = ri + ir), with
// the imaginary-imaginary sum kept in its own accumulator and subtracted at
// the end.
//
// One workgroup computes a BM x BN tile over all of K in steps of 32 (one MFMA
// depth). Each thread gathers runs of 4 consecutive k of one row (lanes along
// K or along M/N, following the operand's contiguous axis), splits them and
// stores each term plane row-major in k, so a lane's MFMA fragment (8 k of one
// row) is one 16-byte LDS read. K offsets are loaded one step ahead of the
// gathers, and the gathers of the next step overlap the MFMAs of this one.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Cx;
using dev::Store;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

template <bool CPLX_, int TERMS_, int BM_, int BN_, int WM_, int WN_, bool AKM_, bool BKM_>
struct SCfg {
  static constexpr bool CPLX = CPLX_, AKM = AKM_, BKM = BKM_;
  static constexpr int TERMS = TERMS_;
  static constexpr int BM = BM_, BN = BN_, BK = 32, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int NT = (BM / WM) * (BN / WN) * 64;
  static constexpr int KV = 4;             // consecutive k per gathered run
  static constexpr int KG = BK / KV;       // runs per row and slab
  static constexpr int COMPS = CPLX ? 2 : 1;
  static constexpr int LDK = BK + 8;       // bf16 pitch of a plane row (80 bytes)
  static constexpr int A_PLANE = BM * LDK, B_PLANE = BN * LDK;
  static constexpr int RUNS_A = BM * KG / NT, RUNS_B = BN * KG / NT;
  static constexpr int KA = AKM ? KV : RUNS_A * KV;  // K offsets tracked per thread
  static constexpr int KB = BKM ? KV : RUNS_B * KV;
  static constexpr int RA = AKM ? RUNS_A : 1;        // row offsets tracked per thread
  static constexpr int RB = BKM ? RUNS_B : 1;
  static constexpr int MI = WM / 16, NI = WN / 16;
  static constexpr int LDD = BN + 1;
  static constexpr size_t LDS_MAIN = (size_t)COMPS * TERMS * (A_PLANE + B_PLANE) * 2;
  static constexpr size_t LDS_EPI = (size_t)COMPS * BM * LDD * 4;
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  static_assert(NT % KG == 0 && BM % (NT / KG) == 0 && BN % (NT / KG) == 0, "K-major run map");
  static_assert(NT % BM == 0 && NT % BN == 0 && KG % (NT / BM) == 0 && KG % (NT / BN) == 0, "MN-major run map");
};

// The bf16 terms of x (TERMS of them, leading first).
template <int TERMS>
__device__ inline void split(float x, __bf16* t) {
  const __bf16 h = (__bf16)x;
  const float r = x - (float)h;
  t[0] = h;
  if constexpr (TERMS == 2) {
    t[1] = (__bf16)r;
  } else {
    const __bf16 m = (__bf16)r;
    t[1] = m;
    t[2] = (__bf16)(r - (float)m);
  }
}

template <class C>
__device__ inline void split_tile(const ContractArgs& p, int64_t h, int tm, int tn, char* smem) {
  typedef Store<float, C::CPLX> St;
  typedef typename St::S S;
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, NT = C::NT, KV = C::KV, KG = C::KG;
  constexpr int RUNS_A = C::RUNS_A, RUNS_B = C::RUNS_B;

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

  __bf16* As = (__bf16*)smem;                                // [COMPS][TERMS][BM][LDK]
  __bf16* Bs = As + C::COMPS * C::TERMS * C::A_PLANE;        // [COMPS][TERMS][BN][LDK]

  // Run coordinates: K-major: one k-group, RUNS rows; M-major: one row, RUNS k-groups.
  const int a_g0 = C::AKM ? tid % KG : tid / BM;
  const int a_m0 = C::AKM ? tid / KG : tid % BM;
  constexpr int a_gstep = C::AKM ? 0 : NT / BM;
  constexpr int a_mstep = C::AKM ? NT / KG : 0;
  const int b_g0 = C::BKM ? tid % KG : tid / BN;
  const int b_n0 = C::BKM ? tid / KG : tid % BN;
  constexpr int b_gstep = C::BKM ? 0 : NT / BN;
  constexpr int b_nstep = C::BKM ? NT / KG : 0;

  int64_t a_row[C::RA], b_col[C::RB];
  uint32_t a_rok = 0, b_cok = 0;
#pragma unroll
  for (int r = 0; r < C::RA; r++) {
    const int64_t m = m0 + a_m0 + r * a_mstep;
    const bool ok = m < M;
    a_row[r] = p.am[ok ? m : 0];
    a_rok |= (uint32_t)ok << r;
  }
#pragma unroll
  for (int r = 0; r < C::RB; r++) {
    const int64_t n = n0 + b_n0 + r * b_nstep;
    const bool ok = n < N;
    b_col[r] = p.bn[ok ? n : 0];
    b_cok |= (uint32_t)ok << r;
  }

  int64_t a_ko[C::KA], b_ko[C::KB];
  uint32_t a_kok = 0, b_kok = 0;
  const int64_t kmax = K - 1;
  auto load_koff = [&](int64_t k0) {
    a_kok = 0;
    b_kok = 0;
#pragma unroll
    for (int q = 0; q < C::KA; q++) {
      const int g = a_g0 + (q / KV) * a_gstep;
      const int64_t k = k0 + g * KV + q % KV;
      a_ko[q] = ak[k < K ? k : kmax];
      a_kok |= (uint32_t)(k < K) << q;
    }
#pragma unroll
    for (int q = 0; q < C::KB; q++) {
      const int g = b_g0 + (q / KV) * b_gstep;
      const int64_t k = k0 + g * KV + q % KV;
      b_ko[q] = bk[k < K ? k : kmax];
      b_kok |= (uint32_t)(k < K) << q;
    }
  };

  S ra[RUNS_A][KV], rb[RUNS_B][KV];
  uint32_t a_live = 0, b_live = 0;
  auto gather = [&]() {
    a_live = 0;
    b_live = 0;
#pragma unroll
    for (int r = 0; r < RUNS_A; r++)
#pragma unroll
      for (int v = 0; v < KV; v++) {
        const int ri = C::AKM ? r : 0, q = C::AKM ? v : r * KV + v;
        const bool ok = ((a_rok >> ri) & 1) && ((a_kok >> q) & 1);
        ra[r][v] = A[ok ? a_row[ri] + a_ko[q] : 0];
        a_live |= (uint32_t)ok << (r * KV + v);
      }
#pragma unroll
    for (int r = 0; r < RUNS_B; r++)
#pragma unroll
      for (int v = 0; v < KV; v++) {
        const int ri = C::BKM ? r : 0, q = C::BKM ? v : r * KV + v;
        const bool ok = ((b_cok >> ri) & 1) && ((b_kok >> q) & 1);
        rb[r][v] = B[ok ? b_col[ri] + b_ko[q] : 0];
        b_live |= (uint32_t)ok << (r * KV + v);
      }
  };

  const bool ca = p.conj_a, cb = p.conj_b;
  auto stash = [&]() {
#pragma unroll
    for (int r = 0; r < RUNS_A; r++) {
      const int m = a_m0 + r * a_mstep, g = a_g0 + r * a_gstep;
      v4bf t[C::COMPS][C::TERMS];
#pragma unroll
      for (int v = 0; v < KV; v++) {
        const bool ok = (a_live >> (r * KV + v)) & 1;
        __bf16 s[C::TERMS];
        split<C::TERMS>(ok ? St::re(ra[r][v]) : 0.0f, s);
#pragma unroll
        for (int x = 0; x < C::TERMS; x++) t[0][x][v] = s[x];
        if constexpr (C::CPLX) {
          float im = ok ? St::im(ra[r][v]) : 0.0f;
          split<C::TERMS>(ca ? -im : im, s);
#pragma unroll
          for (int x = 0; x < C::TERMS; x++) t[1][x][v] = s[x];
        }
      }
#pragma unroll
      for (int c = 0; c < C::COMPS; c++)
#pragma unroll
        for (int x = 0; x < C::TERMS; x++)
          *(v4bf*)&As[(c * C::TERMS + x) * C::A_PLANE + m * C::LDK + g * KV] = t[c][x];
    }
#pragma unroll
    for (int r = 0; r < RUNS_B; r++) {
      const int n = b_n0 + r * b_nstep, g = b_g0 + r * b_gstep;
      v4bf t[C::COMPS][C::TERMS];
#pragma unroll
      for (int v = 0; v < KV; v++) {
        const bool ok = (b_live >> (r * KV + v)) & 1;
        __bf16 s[C::TERMS];
        split<C::TERMS>(ok ? St::re(rb[r][v]) : 0.0f, s);
#pragma unroll
        for (int x = 0; x < C::TERMS; x++) t[0][x][v] = s[x];
        if constexpr (C::CPLX) {
          float im = ok ? St::im(rb[r][v]) : 0.0f;
          split<C::TERMS>(cb ? -im : im, s);
#pragma unroll
          for (int x = 0; x < C::TERMS; x++) t[1][x][v] = s[x];
        }
      }
#pragma unroll
      for (int c = 0; c < C::COMPS; c++)
#pragma unroll
        for (int x = 0; x < C::TERMS; x++)
          *(v4bf*)&Bs[(c * C::TERMS + x) * C::B_PLANE + n * C::LDK + g * KV] = t[c][x];
    }
  };

  // Term pairs multiplied: (a term, b term).
  constexpr int NP = C::TERMS == 2 ? 3 : 6;
  constexpr int PA[6] = {0, 0, 1, 1, 0, 2};
  constexpr int PB[6] = {0, 1, 0, 1, 2, 0};
  // Accumulators: real [1]; complex [3] = rr, ii, ri + ir.
  constexpr int ACCS = C::CPLX ? 3 : 1;
  v4f acc[ACCS][C::MI][C::NI];
#pragma unroll
  for (int s = 0; s < ACCS; s++)
#pragma unroll
    for (int i = 0; i < C::MI; i++)
#pragma unroll
      for (int j = 0; j < C::NI; j++) acc[s][i][j] = v4f{0, 0, 0, 0};

  const int lr = lane & 15, lk = lane >> 4;
  const int64_t nk = (K + BK - 1) / BK;
  load_koff(0);
  gather();
  if (nk > 1) load_koff(BK);
  for (int64_t kt = 0; kt < nk; kt++) {
    stash();
    __syncthreads();
    if (kt + 1 < nk) {
      gather();
      if (kt + 2 < nk) load_koff((kt + 2) * BK);
    }
#pragma unroll
    for (int c = 0; c < C::COMPS; c++) {
#pragma unroll
      for (int d = 0; d < C::COMPS; d++) {
        // (A component c) x (B component d) -> accumulator slot
        const int slot = !C::CPLX ? 0 : (c == 0 && d == 0 ? 0 : (c == 1 && d == 1 ? 1 : 2));
        v8bf a[C::TERMS][C::MI], b[C::TERMS][C::NI];
#pragma unroll
        for (int x = 0; x < C::TERMS; x++) {
#pragma unroll
          for (int i = 0; i < C::MI; i++)
            a[x][i] = *(const v8bf*)&As[(c * C::TERMS + x) * C::A_PLANE + (wm + 16 * i + lr) * C::LDK + 8 * lk];
#pragma unroll
          for (int j = 0; j < C::NI; j++)
            b[x][j] = *(const v8bf*)&Bs[(d * C::TERMS + x) * C::B_PLANE + (wn + 16 * j + lr) * C::LDK + 8 * lk];
        }
#pragma unroll
        for (int pp = 0; pp < NP; pp++)
#pragma unroll
          for (int i = 0; i < C::MI; i++)
#pragma unroll
            for (int j = 0; j < C::NI; j++)
              acc[slot][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[PA[pp]][i], b[PB[pp]][j], acc[slot][i][j],
                                                                        0, 0, 0);
      }
    }
    __syncthreads();
  }

  float* Dre = (float*)smem;
  float* Dim = Dre + BM * C::LDD;
#pragma unroll
  for (int i = 0; i < C::MI; i++)
#pragma unroll
    for (int j = 0; j < C::NI; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = wm + 16 * i + 4 * lk + q, col = wn + 16 * j + lr;
        if constexpr (C::CPLX) {
          Dre[row * C::LDD + col] = acc[0][i][j][q] - acc[1][i][j][q];
          Dim[row * C::LDD + col] = acc[2][i][j][q];
        } else {
          Dre[row * C::LDD + col] = acc[0][i][j][q];
        }
      }
  __syncthreads();
  dev::write_back<float, float, C::CPLX, BM, BN, NT, C::LDD>(p, h, m0, n0, Dre, Dim);
}

template <class C>
__global__ __launch_bounds__(C::NT) void k_split(ContractArgs p) {
  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  int64_t h;
  int tm, tn;
  dev::decode_tile(blockIdx.x, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  split_tile<C>(p, h, tm, tn, smem);
}

template <class C>
hipError_t launch_cfg(const ContractArgs& a, hipStream_t s) {
  ContractArgs p = a;
  p.tiles_m = (int)((p.M + C::BM - 1) / C::BM);
  p.tiles_n = (int)((p.N + C::BN - 1) / C::BN);
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  if (nwg == 0) return hipSuccess;
  if (nwg > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_split<C>, dim3((unsigned)nwg), dim3(C::NT), 0, s, p);
  return hipGetLastError();
}

template <bool CPLX, int TERMS, int BM, int BN, int WM, int WN>
hipError_t launch_majors(const ContractArgs& a, hipStream_t s) {
  if (a.a_kmajor) {
    if (a.b_kmajor) return launch_cfg<SCfg<CPLX, TERMS, BM, BN, WM, WN, true, true>>(a, s);
    return launch_cfg<SCfg<CPLX, TERMS, BM, BN, WM, WN, true, false>>(a, s);
  }
  if (a.b_kmajor) return launch_cfg<SCfg<CPLX, TERMS, BM, BN, WM, WN, false, true>>(a, s);
  return launch_cfg<SCfg<CPLX, TERMS, BM, BN, WM, WN, false, false>>(a, s);
}

template <bool CPLX, int TERMS>
hipError_t launch_tile(const ContractArgs& a, int tile_id, hipStream_t s) {
  if constexpr (!CPLX) {
    if (tile_id == STILE_128x128) return launch_majors<CPLX, TERMS, 128, 128, 64, 64>(a, s);
  }
  if (tile_id == STILE_64x64) return launch_majors<CPLX, TERMS, 64, 64, 32, 32>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_contract_split(const ContractArgs& a, DType t, int nterms, int tile_id, hipStream_t s) {
  if (t == DType::F32) return nterms == 2 ? launch_tile<false, 2>(a, tile_id, s) : launch_tile<false, 3>(a, tile_id, s);
  if (t == DType::C32) return nterms == 2 ? launch_tile<true, 2>(a, tile_id, s) : launch_tile<true, 3>(a, tile_id, s);
  return hipErrorInvalidValue;
}

}  // namespace tp
