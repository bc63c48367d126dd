// FP8 precision for f32/c32 storage: operands quantised to OCP e4m3 with one
// power-of-two scale per 128-element K panel of every row, multiplied on
// gfx950's scaled fp8 MFMA (v_mfma_scale_f32_16x16x128_f8f6f4, fp32
// accumulation), which applies both operands' scales in the instruction. No
// counterpart exists in the reference (crates/tprims-capi/src/product.rs:132
// refuses any precision other than the storage one); gather, conjugation and
// write-back semantics are those of the storage kernels.
//
// Scale of a panel (row r, K block kb of 128): e = the smallest integer with
// amax * 2^-e <= 448 (the largest e4m3 value), stored as the E8M0 exponent
// 127 + e; elements are converted as x * 2^-e with round-to-nearest. A complex
// operand has separate re and im panels, each with its own scales.
//
// Instruction layout (checked on gfx950 with known inputs): lane l supplies
// row l % 16 and 32 bytes of fp8; A and B lanes pair byte for byte, and the
// lane's scale register (its low byte) scales one 32-k quarter of the 128-k
// step. With one scale per 128-k panel every quarter of a row carries the
// same value, so any assignment of a panel's 128 k to (quarter, byte) works
// when A and B share it. Every packer uses kslot(q, b) = 16 (b / 4) + 4 q +
// b % 4 (byte b of quarter q = lane / 16): the four lanes of a row then read
// 64 contiguous bytes of a K-major row per 16-byte load.
//
// Two stages, as in the split-BF16 path:
//  1. k_pack_fp8: one wave per (batch h, 128-k block, 16-row block): each lane
//     gathers its 32 values, the four lanes of a row reduce the panel's amax
//     with cross-lane shuffles, and the lane stores its 32 bytes of the 2 KiB
//     fragment as two 16-byte halves, at 16 lane and 1024 + 16 lane, so that
//     each 16-byte LDS read of the GEMM hits 64 distinct banks per 16 lanes
//     (at 32 lane, lanes 0 and 8 of every 16 shared banks: 1.05e7 bank-conflict
//     cycles per 4096^3 GEMM), plus (lanes of the first quarter) the row's
//     int32 scale word.
//  2. k_gemm_fp8: per 128-k step the workgroup copies its row blocks'
//     fragments and scale words into LDS by DMA; each wave reads 32 bytes per
//     fragment and runs one scaled MFMA per 16x16 block and real product.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Store;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

constexpr int F8_FRAG = 2048;  // bytes per fragment (16 rows x 128 k)
constexpr int F8_KBLK = 128;
constexpr float F8_MAX = 448.0f;

// The k within a 128-k panel that byte b (0..31) of quarter q (0..3) holds.
__device__ inline int kslot(int q, int b) { return 16 * (b >> 2) + 4 * q + (b & 3); }

struct F8PackArgs {
  const void* X;
  const int64_t* rv;  // row scatter (length R)
  const int64_t* kv;  // K scatter (length K)
  const int64_t* hv;  // batch scatter (length H)
  int64_t R, K, H;
  int64_t RB;  // row blocks (of 16) per (h, K block), padded to whole tiles and to 4
  int64_t KB;  // K blocks of 128
  int32_t conj;
  int32_t vec;     // ContractArgs::a_kvec / b_kvec of this operand
  int32_t kmajor;  // ContractArgs::a_kmajor / b_kmajor: contiguous along K
  int32_t rvec;    // rows in 16-byte runs of 4 consecutive f32 (M- or N-major, 16-byte aligned)
  uint8_t* out;
  int32_t* scales;  // [h][kb][RB * 16 rows][comp]
};

// Exponent e of the panel scale 2^e for a panel maximum amax (0 for an all-zero
// panel): the smallest e with amax * 2^-e <= 448.
__device__ inline int panel_exponent(float amax) {
  if (!(amax > 0.0f)) return 0;
  int E;
  (void)frexpf(amax / F8_MAX, &E);  // amax / 448 in [2^(E-1), 2^E)
  int e = E - 1;
  if (ldexpf(amax, -e) > F8_MAX) e = E;
  return e < -126 ? -126 : (e > 127 ? 127 : e);
}

template <bool CPLX, bool VEC>
__global__ __launch_bounds__(256) void k_pack_fp8(F8PackArgs p) {
  typedef Store<float, CPLX> St;
  typedef typename St::S S;
  constexpr int COMPS = CPLX ? 2 : 1;
  const int lane = threadIdx.x & 63;
  const int64_t rgroups = p.RB / 4;
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int64_t rb = rg * 4 + (threadIdx.x >> 6);
  const int64_t r = rb * 16 + (lane & 15);
  const int lq = lane >> 4;
  const int64_t k0 = kb * F8_KBLK;
  const bool rok = r < p.R;
  // Row 0 and k 0 have offset 0: X (and X + 0..3 under VEC) is a valid address.
  const S* __restrict__ X = (const S*)p.X + p.hv[h] + (rok ? p.rv[r] : 0);
  float re[32], im[32];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    // Bytes 4 t .. 4 t + 3: k = k0 + 16 t + 4 lq + 0..3.
    const int64_t kt = k0 + kslot(lq, 4 * t);
    if constexpr (VEC) {
      const bool ok = rok && kt < p.K;  // K % 8 == 0: the 4 k in or out together
      const v4f* src = (const v4f*)(X + (!ok ? 0 : p.vec == 2 ? kt : p.kv[kt]));
      if constexpr (!CPLX) {
        const v4f x = src[0];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          re[4 * t + j] = ok ? x[j] : 0.0f;
          im[4 * t + j] = 0.0f;
        }
      } else {
        const v4f x0 = src[0], x1 = src[1];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const v4f& x = j < 2 ? x0 : x1;
          re[4 * t + j] = ok ? x[2 * (j & 1)] : 0.0f;
          im[4 * t + j] = ok ? x[2 * (j & 1) + 1] : 0.0f;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t k = kt + j;
        const bool ok = rok && k < p.K;
        const S v = X[ok ? p.kv[k] : 0];
        re[4 * t + j] = ok ? St::re(v) : 0.0f;
        im[4 * t + j] = ok ? St::im(v) : 0.0f;
      }
    }
  }
  uint8_t* out = p.out + ((h * p.KB + kb) * p.RB + rb) * COMPS * F8_FRAG + lane * 16;
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    const bool neg = c == 1 && p.conj;
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < 32; j++) amax = fmaxf(amax, fabsf(c == 0 ? re[j] : im[j]));
    amax = fmaxf(amax, __shfl_xor(amax, 16));
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    const int e = panel_exponent(amax);
    const float inv = ldexpf(neg ? -1.0f : 1.0f, -e);
    int w[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
      const float* v = (c == 0 ? re : im) + 4 * d;
      int x = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * inv, v[1] * inv, 0, false);
      w[d] = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * inv, v[3] * inv, x, true);
    }
    *(v4u*)(out + c * F8_FRAG) = v4u{(unsigned)w[0], (unsigned)w[1], (unsigned)w[2], (unsigned)w[3]};
    *(v4u*)(out + c * F8_FRAG + 1024) = v4u{(unsigned)w[4], (unsigned)w[5], (unsigned)w[6], (unsigned)w[7]};
    if (lane < 16) p.scales[((h * p.KB + kb) * p.RB * 16 + r) * COMPS + c] = 127 + e;
  }
}

// The same panels for K-major operands without 16-byte K vectors: the
// workgroup walks its 64-row x 128-k block in four 32-k chunks, each read
// with lanes along K (one run of nearby addresses per load instruction) into
// LDS (row pitch 33 floats); the lanes whose 32 k lie in the chunk take their
// row from there.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_pack_fp8_lds(F8PackArgs p) {
  typedef Store<float, CPLX> St;
  typedef typename St::S S;
  constexpr int COMPS = CPLX ? 2 : 1;
  constexpr int LDW = 33;
  __shared__ float tre[64 * LDW];
  __shared__ float tim[CPLX ? 64 * LDW : 1];
  const int lane = threadIdx.x & 63;
  const int64_t rgroups = p.RB / 4;
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int w = threadIdx.x >> 6, lk = lane >> 4;
  const int64_t rb = rg * 4 + w;
  const int64_t r = rb * 16 + (lane & 15);
  const int rl = w * 16 + (lane & 15);
  // Row 0 and k 0 have offset 0: out-of-range lanes read X[0] and store zero.
  const S* __restrict__ X = (const S*)p.X + p.hv[h];
  float re[32], im[32];
#pragma unroll
  for (int ch = 0; ch < 4; ch++) {
    if (ch > 0) __syncthreads();  // the previous chunk has been read
    const int kl = threadIdx.x & 31;
    const int64_t k = kb * F8_KBLK + 32 * ch + kl;
    const bool kok = k < p.K;
    const int64_t ko = kok ? p.kv[k] : 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int rr = (threadIdx.x >> 5) + 8 * i;
      const int64_t rg_r = rg * 64 + rr;
      const bool ok = kok && rg_r < p.R;
      const S v = X[ok ? p.rv[rg_r] + ko : 0];
      tre[rr * LDW + kl] = ok ? St::re(v) : 0.0f;
      if constexpr (CPLX) tim[rr * LDW + kl] = ok ? St::im(v) : 0.0f;
    }
    __syncthreads();
    // This chunk holds bytes 8 ch .. 8 ch + 7 of every lane: k = 32 ch + 16 tt
    // + 4 lk + j within the panel.
#pragma unroll
    for (int tt = 0; tt < 2; tt++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int b = 4 * (2 * ch + tt) + j, kl = 16 * tt + 4 * lk + j;
        re[b] = tre[rl * LDW + kl];
        im[b] = CPLX ? tim[rl * LDW + kl] : 0.0f;
      }
  }
  uint8_t* out = p.out + ((h * p.KB + kb) * p.RB + rb) * COMPS * F8_FRAG + lane * 16;
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    const bool neg = c == 1 && p.conj;
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < 32; j++) amax = fmaxf(amax, fabsf(c == 0 ? re[j] : im[j]));
    amax = fmaxf(amax, __shfl_xor(amax, 16));
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    const int e = panel_exponent(amax);
    const float inv = ldexpf(neg ? -1.0f : 1.0f, -e);
    int wd[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
      const float* v = (c == 0 ? re : im) + 4 * d;
      int x = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * inv, v[1] * inv, 0, false);
      wd[d] = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * inv, v[3] * inv, x, true);
    }
    *(v4u*)(out + c * F8_FRAG) = v4u{(unsigned)wd[0], (unsigned)wd[1], (unsigned)wd[2], (unsigned)wd[3]};
    *(v4u*)(out + c * F8_FRAG + 1024) = v4u{(unsigned)wd[4], (unsigned)wd[5], (unsigned)wd[6], (unsigned)wd[7]};
    if (lane < 16) p.scales[((h * p.KB + kb) * p.RB * 16 + r) * COMPS + c] = 127 + e;
  }
}

// The same panels for operands gathered along their rows (M- or N-major):
// the workgroup's four waves take the same 64 rows, one lane per row, and
// one 32-k quarter of the 128-k block each, so every load instruction reads
// 64 rows at one k (consecutive addresses where the row scatter is); the
// waves combine the rows' panel maxima through LDS, and each lane stores its
// quarter's 32 bytes into the fragment slot of (row, quarter).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_pack_fp8_rows(F8PackArgs p) {
  typedef Store<float, CPLX> St;
  typedef typename St::S S;
  constexpr int COMPS = CPLX ? 2 : 1;
  __shared__ float red[COMPS][4][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;  // q: this wave's 32-k quarter
  const int64_t rgroups = p.RB / 4;
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int64_t r = rg * 64 + lane;
  const bool rok = r < p.R;
  const int64_t k0 = kb * F8_KBLK;
  // Row 0 and k 0 have offset 0: out-of-range lanes read X[0] and store zero.
  const S* __restrict__ X = (const S*)p.X + p.hv[h] + (rok ? p.rv[r] : 0);
  float re[32], im[32];
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int64_t k = k0 + kslot(q, j);
    const bool ok = rok && k < p.K;
    const S v = X[k < p.K ? p.kv[k] : 0];
    re[j] = ok ? St::re(v) : 0.0f;
    im[j] = ok ? St::im(v) : 0.0f;
  }
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < 32; j++) amax = fmaxf(amax, fabsf(c == 0 ? re[j] : im[j]));
    red[c][q][lane] = amax;
  }
  __syncthreads();
  // Fragment of row block r / 16; this lane's slot is (quarter q, row r % 16).
  uint8_t* out = p.out + ((h * p.KB + kb) * p.RB + rg * 4 + lane / 16) * COMPS * F8_FRAG + (q * 16 + lane % 16) * 16;
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    const float amax = fmaxf(fmaxf(red[c][0][lane], red[c][1][lane]), fmaxf(red[c][2][lane], red[c][3][lane]));
    const int e = panel_exponent(amax);
    const bool neg = c == 1 && p.conj;
    const float inv = ldexpf(neg ? -1.0f : 1.0f, -e);
    int w[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
      const float* v = (c == 0 ? re : im) + 4 * d;
      int x = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * inv, v[1] * inv, 0, false);
      w[d] = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * inv, v[3] * inv, x, true);
    }
    *(v4u*)(out + c * F8_FRAG) = v4u{(unsigned)w[0], (unsigned)w[1], (unsigned)w[2], (unsigned)w[3]};
    *(v4u*)(out + c * F8_FRAG + 1024) = v4u{(unsigned)w[4], (unsigned)w[5], (unsigned)w[6], (unsigned)w[7]};
    if (q == 0) p.scales[((h * p.KB + kb) * p.RB * 16 + r) * COMPS + c] = 127 + e;
  }
}

// Real operands whose rows come in 16-byte runs (4 consecutive f32 at
// 16-byte aligned offsets): each lane loads 4 rows at a time, so a load
// instruction reads 1 KiB (256 rows at one k); the workgroup's eight waves
// take the same 256 rows and 16 consecutive k each (K offsets by scalar
// loads: the range is uniform across the wave) and combine the rows' panel
// maxima through LDS. The workgroup's output is 16 consecutive fragments (32
// KiB): each lane puts its four rows' bytes into their fragment slots in LDS
// (4 bytes per quarter, as kslot places the wave's k; fragment pitch 2064
// bytes, so the 16 fragments a store instruction touches start on different
// banks), then the workgroup writes the block out as whole 8 KiB rows.
// Fragments past the padded panel (RB row blocks) are not written.
constexpr int F8_LPITCH = F8_FRAG + 16;

__global__ __launch_bounds__(512) void k_pack_fp8_rows4(F8PackArgs p) {
  __shared__ float red[8][256];
  __shared__ __attribute__((aligned(16))) uint8_t frag[16 * F8_LPITCH];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t rgroups = (p.RB + 15) / 16;
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int64_t r0 = rg * 256 + 4 * lane;  // the lane's first row
  const bool rok = r0 < p.R;               // R is a multiple of 4: all four rows or none
  const int64_t k0 = kb * F8_KBLK + 16 * w;  // wave w loads k0 .. k0 + 15
  const float* __restrict__ X = (const float*)p.X + p.hv[h] + (rok ? p.rv[r0] : 0);
  v4f v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int64_t k = k0 + j;
    const bool ok = rok && k < p.K;
    const v4f x = *(const v4f*)(X + (k < p.K ? p.kv[k] : 0));
    v[j] = ok ? x : v4f{0, 0, 0, 0};
  }
  v4f amax = v4f{0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 16; j++)
#pragma unroll
    for (int i = 0; i < 4; i++) amax[i] = fmaxf(amax[i], fabsf(v[j][i]));
#pragma unroll
  for (int i = 0; i < 4; i++) red[w][4 * lane + i] = amax[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int rl = 4 * lane + i;  // row within the group of 256
    float m = red[0][rl];
#pragma unroll
    for (int g = 1; g < 8; g++) m = fmaxf(m, red[g][rl]);
    const int e = panel_exponent(m);
    const float inv = ldexpf(1.0f, -e);
    // The wave's 16 k are byte word w of every quarter c (k 16 w + 4 c + 0..3
    // is byte 4 w + 0..3 of quarter c): one 4-byte word per quarter, in half
    // w / 4 of the slot, at 4 (w % 4).
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const v4f* x = v + 4 * c;
      const int y = __builtin_amdgcn_cvt_pk_fp8_f32(x[0][i] * inv, x[1][i] * inv, 0, false);
      const int wd = __builtin_amdgcn_cvt_pk_fp8_f32(x[2][i] * inv, x[3][i] * inv, y, true);
      *(int*)(frag + (rl / 16) * F8_LPITCH + (w >> 2) * 1024 + (c * 16 + rl % 16) * 16 + 4 * (w & 3)) = wd;
    }
    const int64_t r = rg * 256 + rl;
    if (w == 0 && r < p.RB * 16) p.scales[(h * p.KB + kb) * p.RB * 16 + r] = 127 + e;
  }
  __syncthreads();
  // The block: fragments rg * 16 .. + 15 of (h, kb), 128 16-byte pieces each.
  const int64_t nfr = p.RB - rg * 16 < 16 ? p.RB - rg * 16 : 16;
  uint8_t* out = p.out + ((h * p.KB + kb) * p.RB + rg * 16) * F8_FRAG;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int idx = t * 512 + threadIdx.x;  // 16-byte piece of the 32 KiB block
    const int fr = idx >> 7, pc = idx & 127;
    if (fr < nfr) *(v4u*)(out + (int64_t)fr * F8_FRAG + pc * 16) = *(const v4u*)(frag + fr * F8_LPITCH + pc * 16);
  }
}

template <bool CPLX_, int BM_, int BN_, int WM_, int WN_>
struct F8Cfg {
  static constexpr bool CPLX = CPLX_;
  static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int NW = (BM / WM) * (BN / WN);
  static constexpr int NT = NW * 64;
  static constexpr int COMPS = CPLX ? 2 : 1;
  static constexpr int A_FRAGS = BM / 16 * COMPS, B_FRAGS = BN / 16 * COMPS;
  // One stage: fragments (2 KiB each, copied as two 1 KiB DMA pieces), then
  // the scale words of A's and B's rows (int32 [row][comp], 256 B per piece).
  static constexpr int PIECES = 2 * (A_FRAGS + B_FRAGS);
  static constexpr int A_SC = BM * COMPS * 4, B_SC = BN * COMPS * 4;  // bytes
  static constexpr int SC_PIECES = (A_SC + B_SC) / 256;
  static constexpr size_t FRAG_BYTES = (size_t)(A_FRAGS + B_FRAGS) * F8_FRAG;
  static constexpr size_t STAGE_BYTES = FRAG_BYTES + A_SC + B_SC;
  static constexpr int MI = WM / 16, NI = WN / 16;
  static constexpr int LDD = BN + 1;
  static constexpr size_t LDS_MAIN = 2 * STAGE_BYTES;
  // The staged epilogue holds the whole tile when it fits the main loop's
  // LDS, else one wave row (WM rows) at a time.
  static constexpr int EPI_ROWS = (size_t)COMPS * BM * LDD * 4 <= LDS_MAIN ? BM : WM;
  static constexpr size_t LDS_EPI = (size_t)COMPS * EPI_ROWS * LDD * 4;
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  static_assert(PIECES % NW == 0, "whole fragment pieces per wave");
  static_assert(A_SC % 256 == 0 && B_SC % 256 == 0, "scale pieces");
  static_assert(LDS <= 160 * 1024, "LDS budget");
};

struct F8GemmArgs {
  ContractArgs c;
  const uint8_t* Ap;
  const uint8_t* Bp;
  const int32_t* As;  // scales
  const int32_t* Bs;
  int64_t RBa, RBb, KB;
};

template <class G>
__global__ __launch_bounds__(G::NT) void k_gemm_fp8(F8GemmArgs g) {
  const ContractArgs& p = g.c;
  __shared__ __attribute__((aligned(16))) char smem[G::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  int64_t h;
  int tm, tn;
  dev::decode_tile(blockIdx.x, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  constexpr int BM = G::BM, BN = G::BN, COMPS = G::COMPS;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = (wave % G::WAVES_M) * G::WM;
  const int wn = (wave / G::WAVES_M) * G::WN;

  // Fragments of K block kb: A's BM/16 row blocks (all comps) contiguous in the
  // panel, then B's; the scale words of the tile's rows likewise.
  const uint8_t* Ag = g.Ap + (h * g.KB * g.RBa + (int64_t)tm * (BM / 16)) * COMPS * F8_FRAG;
  const uint8_t* Bg = g.Bp + (h * g.KB * g.RBb + (int64_t)tn * (BN / 16)) * COMPS * F8_FRAG;
  const int32_t* Asg = g.As + (h * g.KB * g.RBa * 16 + (int64_t)tm * BM) * COMPS;
  const int32_t* Bsg = g.Bs + (h * g.KB * g.RBb * 16 + (int64_t)tn * BN) * COMPS;
  const int64_t a_step = g.RBa * COMPS * F8_FRAG, b_step = g.RBb * COMPS * F8_FRAG;
  const int64_t as_step = g.RBa * 16 * COMPS, bs_step = g.RBb * 16 * COMPS;
  constexpr int PW = G::PIECES / G::NW;
  auto issue = [&](int64_t kb, int st) {
    char* stage = smem + (size_t)st * G::STAGE_BYTES;
#pragma unroll
    for (int i = 0; i < PW; i++) {
      const int f = wave + i * G::NW;  // 1 KiB piece index
      const uint8_t* src = f < 2 * G::A_FRAGS ? Ag + kb * a_step + (int64_t)f * 1024
                                              : Bg + kb * b_step + (int64_t)(f - 2 * G::A_FRAGS) * 1024;
      __builtin_amdgcn_global_load_lds((const void*)(src + lane * 16),
                                       (__attribute__((address_space(3))) void*)(stage + (size_t)f * 1024), 16, 0, 0);
    }
    for (int f = wave; f < G::SC_PIECES; f += G::NW) {  // 256 B scale pieces
      const int32_t* src = f < G::A_SC / 256 ? Asg + kb * as_step + f * 64 : Bsg + kb * bs_step + (f - G::A_SC / 256) * 64;
      __builtin_amdgcn_global_load_lds((const void*)(src + lane),
                                       (__attribute__((address_space(3))) void*)(stage + G::FRAG_BYTES + (size_t)f * 256),
                                       4, 0, 0);
    }
  };

  constexpr int ACCS = G::CPLX ? 3 : 1;
  v4f acc[ACCS][G::MI][G::NI];
#pragma unroll
  for (int s = 0; s < ACCS; s++)
#pragma unroll
    for (int i = 0; i < G::MI; i++)
#pragma unroll
      for (int j = 0; j < G::NI; j++) acc[s][i][j] = v4f{0, 0, 0, 0};

  const int64_t nk = g.KB;
  issue(0, 0);
  if (nk > 1) issue(1, 1);
  __builtin_amdgcn_s_waitcnt(7 << 4);  // vmcnt(0) lgkmcnt(0)
  __syncthreads();
  const int lr = lane & 15;
  for (int64_t kb = 0; kb < nk; kb++) {
    const char* stage = smem + (size_t)(kb & 1) * G::STAGE_BYTES;
    const int32_t* sc = (const int32_t*)(stage + G::FRAG_BYTES);
    v8i a[COMPS][G::MI], b[COMPS][G::NI];
    int sa[COMPS][G::MI], sb[COMPS][G::NI];
#pragma unroll
    for (int c = 0; c < COMPS; c++) {
#pragma unroll
      for (int i = 0; i < G::MI; i++) {
        const char* q = stage + (size_t)((((wm >> 4) + i) * COMPS + c) * F8_FRAG) + lane * 16;
        const v4u lo = *(const v4u*)q, hi = *(const v4u*)(q + 1024);
        a[c][i] = v8i{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
        sa[c][i] = sc[(wm + 16 * i + lr) * COMPS + c];
      }
#pragma unroll
      for (int j = 0; j < G::NI; j++) {
        const char* q = stage + (size_t)((G::A_FRAGS + ((wn >> 4) + j) * COMPS + c) * F8_FRAG) + lane * 16;
        const v4u lo = *(const v4u*)q, hi = *(const v4u*)(q + 1024);
        b[c][j] = v8i{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
        sb[c][j] = sc[BM * COMPS + (wn + 16 * j + lr) * COMPS + c];
      }
    }
#pragma unroll
    for (int c = 0; c < COMPS; c++)
#pragma unroll
      for (int d = 0; d < COMPS; d++) {
        const int slot = !G::CPLX ? 0 : (c == 0 && d == 0 ? 0 : (c == 1 && d == 1 ? 1 : 2));
#pragma unroll
        for (int i = 0; i < G::MI; i++)
#pragma unroll
          for (int j = 0; j < G::NI; j++)
            acc[slot][i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[c][i], b[d][j], acc[slot][i][j], 0, 0, 0,
                                                                               sa[c][i], 0, sb[d][j]);
      }
    // Own pieces of block kb + 1 landed and reads of block kb done; after the
    // barrier block kb + 1 is visible and stage kb & 1 takes block kb + 2.
    __builtin_amdgcn_s_waitcnt(7 << 4);
    __syncthreads();
    if (kb + 2 < nk) issue(kb + 2, (int)(kb & 1));
  }

  const int lk = lane >> 4;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  if constexpr (!G::CPLX) {
    // D contiguous along M with affine offsets, no broadcast copies and
    // 16-byte aligned columns: a lane's four consecutive accumulator rows go
    // straight to D as one 16-byte store (alpha and beta * C in registers).
    float* __restrict__ Dg = (float*)p.D + p.hd[h];
    const float* __restrict__ Cg = (const float*)p.C + p.hc[h];
    const bool direct = p.d_mmajor && p.mn_affine && p.dr == nullptr && p.cd_stride[2] == 1 &&
                        p.cd_stride[3] % 4 == 0 && (uintptr_t)Dg % 16 == 0 &&
                        (!p.read_c || (p.cd_stride[0] == 1 && p.cd_stride[1] % 4 == 0 && (uintptr_t)Cg % 16 == 0));
    if (direct) {
      // Staged column-major through LDS (pitch BM + 4: each 16-lane group's
      // 16-byte stores hit distinct banks), CH columns at a time, then written
      // as whole columns of the tile: 64 lanes cover 1 KiB of one column per
      // store (16-byte stores straight from the accumulators covered 64-byte
      // pieces of 16 columns: f32 transform step GEMM 175.5 against 171.7 us).
      constexpr int LDM = BM + 4;
      constexpr int CH = (size_t)BN * LDM * 4 <= G::LDS ? BN : BN / 2;
      static_assert((size_t)CH * LDM * 4 <= G::LDS && BN % CH == 0 && G::WN <= CH && CH % G::WN == 0, "staging");
      static_assert(BM * CH % (4 * G::NT) == 0, "whole vectors per thread");
      float* Dt = (float*)smem;
      const float alr = (float)p.alpha[0], ber = (float)p.beta[0];
#pragma unroll 1
      for (int c = 0; c < BN / CH; c++) {
        if (c > 0) __syncthreads();  // the previous columns have been written out
        if (wn / CH == c) {
#pragma unroll
          for (int i = 0; i < G::MI; i++)
#pragma unroll
            for (int j = 0; j < G::NI; j++)
              *(v4f*)&Dt[(wn - c * CH + 16 * j + lr) * LDM + wm + 16 * i + 4 * lk] = acc[0][i][j];
        }
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < BM * CH / (4 * G::NT); t++) {
          const int e = threadIdx.x + t * G::NT;
          const int ml = 4 * (e % (BM / 4)), nl = e / (BM / 4);
          const int64_t m = m0 + ml, n = n0 + c * CH + nl;
          if (m >= p.M || n >= p.N) continue;
          const v4f x = *(const v4f*)&Dt[nl * LDM + ml];
          const int64_t off = m + n * p.cd_stride[3], coff = m + n * p.cd_stride[1];
          if (p.M - m >= 4) {
            v4f y = alr * x;
            if (p.read_c) y += ber * *(const v4f*)(Cg + coff);
            *(v4f*)(Dg + off) = y;
          } else {
            for (int q = 0; q < p.M - m; q++) Dg[off + q] = alr * x[q] + (p.read_c ? ber * Cg[coff + q] : 0.0f);
          }
        }
      }
      return;
    }
  }
  constexpr int ER = G::EPI_ROWS;
  float* Dre = (float*)smem;
  float* Dim = Dre + ER * G::LDD;
#pragma unroll 1
  for (int c = 0; c < BM / ER; c++) {
    if (c > 0) __syncthreads();  // the previous chunk's write-back has read the staging rows
    if (wm / ER == c) {
#pragma unroll
      for (int i = 0; i < G::MI; i++)
#pragma unroll
        for (int j = 0; j < G::NI; j++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int row = wm - c * ER + 16 * i + 4 * lk + q, col = wn + 16 * j + lr;
            if constexpr (G::CPLX) {
              Dre[row * G::LDD + col] = acc[0][i][j][q] - acc[1][i][j][q];
              Dim[row * G::LDD + col] = acc[2][i][j][q];
            } else {
              Dre[row * G::LDD + col] = acc[0][i][j][q];
            }
          }
    }
    __syncthreads();
    dev::write_back<float, float, G::CPLX, ER, BN, G::NT, G::LDD>(p, h, m0 + c * ER, n0, Dre, Dim);
  }
}

struct Shape {
  int bm, bn;
};

Shape f8_shape(bool cplx) { return cplx ? Shape{128, 64} : Shape{256, 256}; }

int64_t pad_rb(int64_t rows, int bm) {
  const int64_t rb = (rows + bm - 1) / bm * (bm / 16);
  return (rb + 3) / 4 * 4;
}

template <bool CPLX>
hipError_t run(const ContractArgs& a, void* panels, hipStream_t s) {
  constexpr int COMPS = CPLX ? 2 : 1;
  const Shape sh = f8_shape(CPLX);
  const int64_t KB = (a.K + F8_KBLK - 1) / F8_KBLK;
  const int64_t RBa = pad_rb(a.M, sh.bm), RBb = pad_rb(a.N, sh.bn);
  uint8_t* Ap = (uint8_t*)panels;
  uint8_t* Bp = Ap + a.H * KB * RBa * COMPS * F8_FRAG;
  int32_t* As = (int32_t*)(Bp + a.H * KB * RBb * COMPS * F8_FRAG);
  int32_t* Bs = As + a.H * KB * RBa * 16 * COMPS;
  // Rows in 16-byte runs: A alone (a_vec16), B with A (ab_vec); real only.
  const int a_rvec = !CPLX && !a.a_kmajor && a.a_vec16;
  const int b_rvec = !CPLX && !a.b_kmajor && a.ab_vec;
  F8PackArgs pa{a.A, a.am, a.ak, a.ha, a.M, a.K, a.H, RBa, KB, a.conj_a, a.a_kvec, a.a_kmajor, a_rvec, Ap, As};
  F8PackArgs pb{a.B, a.bn, a.bk, a.hb, a.N, a.K, a.H, RBb, KB, a.conj_b, a.b_kvec, a.b_kmajor, b_rvec, Bp, Bs};
  for (const F8PackArgs* q : {&pa, &pb}) {
    const int64_t grid = q->H * q->KB * (q->RB / 4);
    if (grid > max_grid(256)) return hipErrorInvalidConfiguration;
    if (q->vec) hipLaunchKernelGGL((k_pack_fp8<CPLX, true>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    else if (q->rvec) {
      const int64_t g4 = q->H * q->KB * ((q->RB + 15) / 16);
      hipLaunchKernelGGL(k_pack_fp8_rows4, dim3((unsigned)g4), dim3(512), 0, s, *q);
    }
    else if (q->kmajor) hipLaunchKernelGGL((k_pack_fp8_lds<CPLX>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    else hipLaunchKernelGGL((k_pack_fp8_rows<CPLX>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  F8GemmArgs g{a, Ap, Bp, As, Bs, RBa, RBb, KB};
  g.c.tiles_m = (int)((a.M + sh.bm - 1) / sh.bm);
  g.c.tiles_n = (int)((a.N + sh.bn - 1) / sh.bn);
  const int64_t nwg = (int64_t)g.c.tiles_m * g.c.tiles_n * a.H;
  if (nwg > max_grid(512)) return hipErrorInvalidConfiguration;
  if constexpr (CPLX) {
    typedef F8Cfg<true, 128, 64, 32, 32> G;
    hipLaunchKernelGGL(k_gemm_fp8<G>, dim3((unsigned)nwg), dim3(G::NT), 0, s, g);
  } else {
    // 256 x 256 on eight waves of 128 x 64 (two per SIMD): fewer LDS
    // fragment reads per MFMA than 64 x 64 wave tiles.
    typedef F8Cfg<false, 256, 256, 128, 64> G;
    hipLaunchKernelGGL(k_gemm_fp8<G>, dim3((unsigned)nwg), dim3(G::NT), 0, s, g);
  }
  return hipGetLastError();
}

}  // namespace

size_t fp8_panel_bytes(const ContractArgs& a, DType t) {
  const bool cplx = t == DType::C32;
  const Shape sh = f8_shape(cplx);
  const int64_t KB = (a.K + F8_KBLK - 1) / F8_KBLK;
  const int64_t rows = pad_rb(a.M, sh.bm) + pad_rb(a.N, sh.bn);  // row blocks
  const int64_t comps = cplx ? 2 : 1;
  return (size_t)(a.H * KB * rows * comps * (F8_FRAG + 16 * 4));
}

hipError_t launch_contract_fp8(const ContractArgs& a, DType t, void* panels, hipStream_t s) {
  if (t == DType::F32) return run<false>(a, panels, s);
  if (t == DType::C32) return run<true>(a, panels, s);
  return hipErrorInvalidValue;
}

}  // namespace tp
