// Device-side pieces shared by the contraction kernels: storage element
// types, the workgroup-to-tile map and the fused write-back
// D = op_D(alpha * tile + beta * op_C(C)) (crates/tprims-kernel/src/pack/writeback.rs:114-300).
#pragma once
#include <hip/hip_runtime.h>

#include "kargs.h"

namespace tp {
namespace dev {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T re, im;
};

// Raw buffer loads: 32-bit byte offsets from a base through a buffer resource
// of BUF_RANGE bytes. An offset at or past the range reads zero, so gathers
// mark an out-of-range row or k with a sentinel instead of selecting: every
// valid offset is below BUF_RANGE (checked on the host), and any sum holding
// one or both sentinels lands in [BUF_RANGE, 2^32).
constexpr uint32_t BUF_RANGE = 1u << 30;
constexpr uint32_t ROW_OOB = 0x80000000u;
constexpr uint32_t K_OOB = 0x40000000u;
// Word 3 of a gfx9 buffer resource for untyped dword access.
constexpr int BUF_FLAGS = 0x00020000;

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)BUF_RANGE, BUF_FLAGS);
}

template <typename S>
__device__ inline S buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(S) == 4) return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  else if constexpr (sizeof(S) == 8) return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  else return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Output placement of one 16x16 accumulator block: lane (lk = lane / 16,
// column lane % 16) holds four rows. The f64 instruction interleaves them
// (row lk + 4q); the f32 one keeps them together (row 4 lk + q). Both were
// measured on gfx950 with known inputs.
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef v4d acc;
  __device__ static inline acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // c - a b: the f64 instructions read their last field as negate flags
  // (bit 0 negates A), so the negation costs nothing.
  __device__ static inline acc run_neg_a(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
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

// Storage element of a dtype.
template <typename T, bool CPLX>
struct Store;
template <typename T>
struct Store<T, false> {
  typedef T S;
  __device__ static inline T re(S v) { return v; }
  __device__ static inline T im(S) { return T(0); }
  __device__ static inline S make(T r, T) { return r; }
};
template <typename T>
struct Store<T, true> {
  typedef Cx<T> S;
  __device__ static inline T re(S v) { return v.re; }
  __device__ static inline T im(S v) { return v.im; }
  __device__ static inline S make(T r, T i) { return S{r, i}; }
};

// Workgroup id -> (h, tile_m, tile_n). Consecutive hardware ids go round-robin
// over the 8 XCDs; remapping gives each XCD a contiguous run of logical tiles
// (XCD x of id b = 8 i + x takes run x: the first nwg % 8 runs are one tile
// longer), and the logical order walks groups of 8 tile-rows so neighbours
// share A rows and B columns in that XCD's L2.
__device__ inline void decode_tile(int64_t bid, int64_t nwg, int tiles_m, int tiles_n, int64_t& h, int& tm, int& tn) {
  {
    const int64_t q = nwg >> 3, r = nwg & 7, x = bid & 7, i = bid >> 3;
    bid = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
  }
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

// Writes a staged BM x BN partial tile of split `split` (Dre/Dim, row pitch
// LDD, in LDS) unscaled to the split-K workspace, rows past M and columns past
// N skipped. The workspace follows D's contiguous axis so that the reduction
// reads and writes along the same axis: work[split][h][n][m] (lanes along m)
// when D is m-major, work[split][h][m][n] (lanes along n) otherwise.
template <typename T, bool CPLX, int BM, int BN, int NT, int LDD>
__device__ inline void write_partial(const ContractArgs& p, int split, int64_t h, int64_t m0, int64_t n0, const T* Dre,
                                     const T* Dim) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  S* W = (S*)p.work + (((int64_t)split * p.H + h) * p.M) * p.N;
  constexpr int EPI = BM * BN / NT;
#pragma unroll 4
  for (int t = 0; t < EPI; t++) {
    const int e = threadIdx.x + t * NT;
    const int ml = p.d_mmajor ? e % BM : e / BN, nl = p.d_mmajor ? e / BM : e % BN;
    const int64_t m = m0 + ml, n = n0 + nl;
    const int64_t w = p.d_mmajor ? n * p.M + m : m * p.N + n;
    if (m < p.M && n < p.N) W[w] = St::make(Dre[ml * LDD + nl], CPLX ? Dim[ml * LDD + nl] : T(0));
  }
}

// D[h, m, n] = op_D(alpha * s + beta * op_C(C)) for one element given its sum
// s = (sr, si), with the broadcast copies of an adjoint (C is D there and each
// copy reads its own previous value). Offsets of h come from the scatter
// vectors or, for one folded batch axis, from h * stride.
template <typename T, bool CPLX>
__device__ inline void store_element(const ContractArgs& p, int64_t h, int64_t m, int64_t n, T sr, T si) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const T alr = (T)p.alpha[0], ali = (T)p.alpha[1];
  const T ber = (T)p.beta[0], bei = (T)p.beta[1];
  const int64_t oc = p.read_c ? (p.h_affine ? h * p.h_stride[2] : p.hc[h]) : 0;
  const int64_t od = p.h_affine ? h * p.h_stride[3] : p.hd[h];
  const T vr = alr * sr - ali * si, vi = alr * si + ali * sr;
  S* D = (S*)p.D + od + (p.mn_affine ? m * p.cd_stride[2] + n * p.cd_stride[3] : p.dm[m] + p.dn[n]);
  const int64_t coff =
      !p.read_c ? 0 : oc + (p.mn_affine ? m * p.cd_stride[0] + n * p.cd_stride[1] : p.cm[m] + p.cn[n]);
  const int64_t copies = p.dr == nullptr ? 1 : p.R;
  for (int64_t r = 0; r < copies; r++) {
    const int64_t roff = p.dr == nullptr ? 0 : p.dr[r];
    T wr = vr, wi = vi;
    if (p.read_c) {
      const S cv = ((const S*)p.C)[coff + roff];
      const T cr = St::re(cv), ci = p.conj_c ? -St::im(cv) : St::im(cv);
      wr += ber * cr - bei * ci;
      wi += ber * ci + bei * cr;
    }
    if (p.conj_d) wi = -wi;
    D[roff] = St::make(wr, wi);
  }
}

// Writes a staged BM x BN tile (Dre/Dim, row pitch LDD, in LDS) to D along
// D's contiguous axis: alpha scaling, beta * op_C(C) only when beta != 0 (C is
// never read otherwise), op_D last, and the broadcast copies of an adjoint.
// VEC: allow the 16-byte vector path below (it costs registers).
template <typename T, typename TS, bool CPLX, int BM, int BN, int NT, int LDD, bool VEC = true>
__device__ inline void write_back(const ContractArgs& p, int64_t h, int64_t m0, int64_t n0, const TS* Dre,
                                  const TS* Dim) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int tid = threadIdx.x;
  S* __restrict__ D = (S*)p.D + p.hd[h];
  const S* __restrict__ Cp = (const S*)p.C + p.hc[h];
  const T alr = (T)p.alpha[0], ali = (T)p.alpha[1];
  const T ber = (T)p.beta[0], bei = (T)p.beta[1];
  const bool rc = p.read_c, cc = p.conj_c, cd = p.conj_d;
  const int64_t M = p.M, N = p.N, R = p.R;
  // Real dtypes whose D (and C, when read) is contiguous along the staged
  // axis, with affine offsets, no broadcast copies and 16-byte aligned runs:
  // V consecutive elements per lane as one 16-byte load / store, with
  // (alpha, beta) applied in registers. A ragged end of the contiguous axis
  // falls back to single elements.
  constexpr int V = 16 / (int)sizeof(T);
  if constexpr (VEC && !CPLX && BM % V == 0 && BN % V == 0 && BM * BN % (V * NT) == 0) {
    const bool mm = p.d_mmajor;
    constexpr int RM = BM / V, RN = BN / V;  // vectors per row / column of the tile
    if (p.mn_affine && p.dr == nullptr) {
      const int64_t ds_c = mm ? p.cd_stride[2] : p.cd_stride[3];  // contiguous axis
      const int64_t ds_o = mm ? p.cd_stride[3] : p.cd_stride[2];  // other axis
      const int64_t cs_c = mm ? p.cd_stride[0] : p.cd_stride[1];
      const int64_t cs_o = mm ? p.cd_stride[1] : p.cd_stride[0];
      const bool ok = ds_c == 1 && ds_o % V == 0 && (uintptr_t)D % 16 == 0 &&
                      (!rc || (cs_c == 1 && cs_o % V == 0 && (uintptr_t)Cp % 16 == 0));
      if (ok) {
        typedef T vec __attribute__((ext_vector_type(16 / sizeof(T))));
        constexpr int EPV = BM * BN / (V * NT);
        // Not unrolled: unrolling raised the register count of the f32
        // 128x128 contraction kernel from 68 to 128 (half the occupancy).
#pragma unroll 1
        for (int t = 0; t < EPV; t++) {
          const int e = tid + t * NT;
          const int ml = mm ? V * (e % RM) : e / RN, nl = mm ? e / RM : V * (e % RN);
          const int64_t m = m0 + ml, n = n0 + nl;
          if (m >= M || n >= N) continue;
          const int64_t lim = mm ? M - m : N - n;  // elements left along the contiguous axis
          T x[V];
#pragma unroll
          for (int q = 0; q < V; q++) x[q] = (T)(mm ? Dre[(ml + q) * LDD + nl] : Dre[ml * LDD + nl + q]);
          const int64_t off = m * p.cd_stride[2] + n * p.cd_stride[3];
          const int64_t coff = m * p.cd_stride[0] + n * p.cd_stride[1];
          if (lim >= V) {
            vec y;
#pragma unroll
            for (int q = 0; q < V; q++) y[q] = alr * x[q];
            if (rc) {
              const vec c = *(const vec*)(Cp + coff);
#pragma unroll
              for (int q = 0; q < V; q++) y[q] += ber * c[q];
            }
            *(vec*)(D + off) = y;
          } else {
            // A constant trip count keeps x in registers (a variable index
            // would put it in scratch memory).
#pragma unroll
            for (int q = 0; q < V; q++) {
              if (q >= lim) break;
              T y = alr * x[q];
              if (rc) y += ber * Cp[coff + q];
              D[off + q] = y;
            }
          }
        }
        return;
      }
    }
  }
  constexpr int EPI = BM * BN / NT;
#pragma unroll 4
  for (int t = 0; t < EPI; t++) {
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
    T xr = (T)Dre[ml * LDD + nl];
    T xi = CPLX ? (T)Dim[ml * LDD + nl] : T(0);
    const T pr = alr * xr - ali * xi;
    const T pi = alr * xi + ali * xr;
    // One folded axis per role: affine offsets, no scatter-vector loads (a
    // latency chain per element otherwise; the f32 transform step's 16-slab
    // tiles spent 40 % of their time here).
    const int64_t off = p.mn_affine ? m * p.cd_stride[2] + n * p.cd_stride[3] : p.dm[m] + p.dn[n];
    const int64_t coff = !rc ? 0 : p.mn_affine ? m * p.cd_stride[0] + n * p.cd_stride[1] : p.cm[m] + p.cn[n];
    // A broadcast (adjoint) writes every copy; its C, when read, is D itself,
    // so each copy accumulates its own previous value.
    const int64_t copies = p.dr == nullptr ? 1 : R;
    for (int64_t r = 0; r < copies; r++) {
      const int64_t roff = p.dr == nullptr ? 0 : p.dr[r];
      T vr = pr, vi = pi;
      if (rc) {
        S cv = Cp[coff + roff];
        T cr = St::re(cv), ci = St::im(cv);
        if (cc) ci = -ci;
        vr += ber * cr - bei * ci;
        vi += ber * ci + bei * cr;
      }
      if (cd) vi = -vi;
      D[off + roff] = St::make(vr, vi);
    }
  }
}

}  // namespace dev
}  // namespace tp
