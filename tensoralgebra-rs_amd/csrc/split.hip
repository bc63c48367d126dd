// Split-BF16 contraction for f32/c32 storage on gfx950's bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation). No counterpart exists in
// the reference (crates/tprims-capi/src/product.rs:132 refuses any precision
// other than the storage one); gather, conjugation and write-back semantics are
// those of the storage kernels.
//
// Every f32 value x is split into bf16 terms:
//   hi = bf16(x), r = x - hi (exact in f32), mid = bf16(r), lo = bf16(r - mid).
//   2 terms (hi, lo = bf16(x - hi)):  a*b ~ ah*bh + ah*bl + al*bh           (3 MFMAs)
//   3 terms (hi, mid, lo):            + am*bm + ah*bl + al*bh on top        (6 MFMAs)
// The dropped products are below 2^-16 (2 terms) or 2^-24 (3 terms) of |a||b|.
// A complex product takes four real ones (Re = rr - ii, Im = ri + ir), the
// imaginary-imaginary sum in its own accumulator, subtracted at the end.
//
// Two stages, like the reference's pack-then-multiply driver
// (crates/tprims-kernel/src/pack/pack.rs:152 packs through the scatter
// vectors, with conjugation folded, before the microkernel runs):
//  1. k_pack gathers each operand once through its scatter vectors (each
//     lane the 8 values it later feeds the MFMA), splits it in registers and
//     writes bf16 panels in MFMA fragment order: for every (batch h, 32-wide K block, 16-row
//     block, component, term) one 1 KiB fragment whose 16 bytes at lane*16
//     are exactly what lane `lane` feeds the instruction (row lane%16,
//     k 8*(lane/16) .. +7). Splitting once per element (not once per output
//     tile) is what takes the split arithmetic off the MFMA loop.
//  2. k_gemm streams the panels: per 32-wide K step a workgroup copies its
//     row blocks (contiguous in the panel) into LDS with 16-byte loads, every
//     fragment read is one conflict-free 16-byte LDS read, and the waves run
//     3 or 6 bf16 MFMAs per real product. The next step's panels are loaded
//     into registers while this step's MFMAs run.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Store;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

constexpr int FRAG = 512;  // bf16 per fragment (16 rows x 32 k)
constexpr int KBLK = 32;

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

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

// The same split for two values at once, each term as the packed dword of
// the pair (x0 in the low half), which is what the MFMA operand holds: one
// v_cvt_pk_bf16_f32 per term and pair, the terms' f32 values recovered by
// shifting or masking the packed halves.
template <int TERMS>
__device__ inline void split_pair(float x0, float x1, uint32_t* t) {
  const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){x0, x1}, v2bf));
  const float r0 = x0 - __builtin_bit_cast(float, h << 16);
  const float r1 = x1 - __builtin_bit_cast(float, h & 0xffff0000u);
  const uint32_t m = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){r0, r1}, v2bf));
  t[0] = h;
  t[1] = m;
  if constexpr (TERMS == 3) {
    const float s0 = r0 - __builtin_bit_cast(float, m << 16);
    const float s1 = r1 - __builtin_bit_cast(float, m & 0xffff0000u);
    t[2] = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){s0, s1}, v2bf));
  }
}

struct PackArgs {
  const void* X;
  const int64_t* rv;  // row scatter (length R)
  const int64_t* kv;  // K scatter (length K)
  const int64_t* hv;  // batch scatter (length H)
  int64_t R, K, H;
  int64_t RB;  // row blocks (of 16) per (h, K block) in the panel, padded
  int64_t KB;  // K blocks of 32
  int32_t conj;
  int32_t vec;     // ContractArgs::a_kvec / b_kvec of this operand
  int32_t kmajor;  // ContractArgs::a_kmajor / b_kmajor: contiguous along K
  __bf16* out;
};

// Each wave packs one 16-row x 32-k block of one batch entry (4 row blocks per
// workgroup): lane l gathers exactly the 8 values it will feed the MFMA (row
// l % 16, k 8 (l / 16) .. + 7), splits them in registers and stores its 16
// bytes of every component/term fragment, so each store instruction writes one
// whole 1 KiB fragment. With `VEC` the 8 values are two (f32) or four (c32)
// 16-byte loads per lane; otherwise one load each through the scatter vectors (rows that are
// contiguous in memory then make each load instruction 4 runs of 16 lanes).
template <bool CPLX, int TERMS, bool VEC>
__global__ __launch_bounds__(256) void k_pack(PackArgs p) {
  typedef Store<float, CPLX> St;
  typedef typename St::S S;
  constexpr int COMPS = CPLX ? 2 : 1;
  constexpr int PER = COMPS * TERMS;  // fragments per row block
  const int lane = threadIdx.x & 63;
  const int64_t rgroups = p.RB / 4;  // RB is a multiple of 4
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int64_t rb = rg * 4 + (threadIdx.x >> 6);
  const int64_t r = rb * 16 + (lane & 15);
  const int64_t k0 = kb * KBLK + 8 * (lane >> 4);
  const bool rok = r < p.R;
  // Row 0 and k 0 have offset 0, so X (and X + 0..7 under VEC) is always a
  // valid address of this batch entry; out-of-range values are read from
  // there and replaced by zero.
  const S* __restrict__ X = (const S*)p.X + p.hv[h] + (rok ? p.rv[r] : 0);
  float re[8], im[8];
  if constexpr (VEC) {
    const bool ok = rok && k0 < p.K;  // K % 8 == 0: all 8 k in range or none
    const v4f* src = (const v4f*)(X + (!ok ? 0 : p.vec == 2 ? k0 : p.kv[k0]));
    if constexpr (!CPLX) {
      const v4f x0 = src[0], x1 = src[1];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        re[j] = ok ? x0[j] : 0.0f;
        re[4 + j] = ok ? x1[j] : 0.0f;
        im[j] = im[4 + j] = 0.0f;
      }
    } else {
      v4f x[4];
#pragma unroll
      for (int c = 0; c < 4; c++) x[c] = src[c];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        re[j] = ok ? x[j >> 1][2 * (j & 1)] : 0.0f;
        im[j] = ok ? x[j >> 1][2 * (j & 1) + 1] : 0.0f;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int64_t k = k0 + j;
      const bool ok = rok && k < p.K;
      const S v = X[ok ? p.kv[k] : 0];
      re[j] = ok ? St::re(v) : 0.0f;
      im[j] = ok ? St::im(v) : 0.0f;
    }
  }
  __bf16* out = p.out + ((h * p.KB + kb) * p.RB + rb) * PER * FRAG + lane * 8;
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    v8bf t[TERMS];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float v = c == 0 ? re[j] : (p.conj ? -im[j] : im[j]);
      __bf16 s[TERMS];
      split<TERMS>(v, s);
#pragma unroll
      for (int x = 0; x < TERMS; x++) t[x][j] = s[x];
    }
#pragma unroll
    for (int x = 0; x < TERMS; x++) *(v8bf*)(out + (c * TERMS + x) * FRAG) = t[x];
  }
}

// The same panels for operands without 16-byte K vectors: a workgroup first
// reads its 64-row x 32-k block with lanes along the operand's contiguous
// axis (K when KMAJ, the rows otherwise), so that every load instruction is
// one run of nearby addresses where the layout has one, and stages it in LDS
// (row pitch 33 floats); each lane then takes its row and 8 k from there.
template <bool CPLX, int TERMS, bool KMAJ>
__global__ __launch_bounds__(256) void k_pack_lds(PackArgs p) {
  typedef Store<float, CPLX> St;
  typedef typename St::S S;
  constexpr int COMPS = CPLX ? 2 : 1;
  constexpr int PER = COMPS * TERMS;
  constexpr int LDW = KBLK + 1;
  __shared__ float tre[64 * LDW];
  __shared__ float tim[CPLX ? 64 * LDW : 1];
  const int64_t rgroups = p.RB / 4;
  int64_t bid = blockIdx.x;
  const int64_t rg = bid % rgroups;
  bid /= rgroups;
  const int64_t kb = bid % p.KB;
  const int64_t h = bid / p.KB;
  const int64_t r0 = rg * 64, kg0 = kb * KBLK;
  // Row 0 and k 0 have offset 0: out-of-range lanes read X[0] and store zero.
  const S* __restrict__ X = (const S*)p.X + p.hv[h];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int rl, kl;
    if constexpr (KMAJ) {
      kl = threadIdx.x & 31;
      rl = (threadIdx.x >> 5) + 8 * i;
    } else {
      rl = threadIdx.x & 63;
      kl = (threadIdx.x >> 6) + 4 * i;
    }
    const int64_t r = r0 + rl, k = kg0 + kl;
    const bool ok = r < p.R && k < p.K;
    const S v = X[ok ? p.rv[r] + p.kv[k] : 0];
    tre[rl * LDW + kl] = ok ? St::re(v) : 0.0f;
    if constexpr (CPLX) tim[rl * LDW + kl] = ok ? St::im(v) : 0.0f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rl = w * 16 + (lane & 15), kq = 8 * (lane >> 4);
  const int64_t rb = rg * 4 + w;
  __bf16* out = p.out + ((h * p.KB + kb) * p.RB + rb) * PER * FRAG + lane * 8;
#pragma unroll
  for (int c = 0; c < COMPS; c++) {
    const float* t = c == 0 ? tre : tim;
    v8bf o[TERMS];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float v = t[rl * LDW + kq + j];
      if (c == 1 && p.conj) v = -v;
      __bf16 s[TERMS];
      split<TERMS>(v, s);
#pragma unroll
      for (int x = 0; x < TERMS; x++) o[x][j] = s[x];
    }
#pragma unroll
    for (int x = 0; x < TERMS; x++) *(v8bf*)(out + (c * TERMS + x) * FRAG) = o[x];
  }
}

// AF_: A is not packed; it is an f32 operand gathered along M in 16-byte
// runs (ContractArgs::a_vec16 with A M-major). Each K step DMAs A's BM x 32
// f32 block straight into LDS (buffer_load ... lds, 16 bytes per lane), and
// every lane splits the 8 values it feeds the MFMA as it reads them. The A
// block is [32 k][BM m] with the 16-byte chunks of row k XOR-permuted by
// 4 ((k / 8) % 4): a fragment read (16 rows x 8 k per lane group, the four
// lane groups 8 k apart) then hits 64 distinct banks.
template <bool CPLX_, int TERMS_, int BM_, int BN_, int WM_, int WN_, int STAGES_, bool AF_ = false>
struct GCfg {
  static constexpr bool CPLX = CPLX_, AF = AF_;
  static constexpr int TERMS = TERMS_, STAGES = STAGES_;
  static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM;
  static constexpr int NW = (BM / WM) * (BN / WN);
  static constexpr int NT = NW * 64;
  static constexpr int COMPS = CPLX ? 2 : 1;
  static constexpr int PER = COMPS * TERMS;
  static constexpr int A_FRAGS = AF ? 0 : BM / 16 * PER, B_FRAGS = BN / 16 * PER;
  static constexpr int FRAGS = A_FRAGS + B_FRAGS;  // 1 KiB fragments per K step
  static constexpr int PW = FRAGS / NW;            // fragments each wave copies per K step
  // The raw A block (AF): 16-byte chunks, 64 per DMA instruction.
  static constexpr int A_CHUNKS_ROW = BM / 4;
  static constexpr int A_INSTR = AF ? KBLK * A_CHUNKS_ROW / 64 : 0;
  static constexpr int AW = A_INSTR / NW;          // A instructions each wave issues per K step
  static constexpr int IW = PW + AW;               // every copy instruction of a wave per K step
  static constexpr size_t A_RAW_BYTES = AF ? (size_t)KBLK * BM * 4 : 0;
  static constexpr int MI = WM / 16, NI = WN / 16;
  // Every wave spans the tile's columns (waves stacked along M) on a split A:
  // each A value of the tile is split by one wave only.
  static constexpr bool WIDE = AF && WN == BN;
  static constexpr int MINB = WIDE ? 512 / NT : 1;  // workgroups per CU the registers are bounded for
  static constexpr int LDD = BN + 1;
  static constexpr size_t STAGE_BYTES = A_RAW_BYTES + (size_t)FRAGS * FRAG * 2;
  static constexpr size_t LDS_MAIN = STAGES * STAGE_BYTES;
  // Epilogue staging: row-major re/im planes, or (real) the column-major
  // tile of the 16-byte write-back of outputs contiguous along M.
  static constexpr size_t LDS_ROWS = (size_t)COMPS * BM * LDD * 4, LDS_COLS = CPLX ? 0 : (size_t)BN * (BM + 4) * 4;
  static constexpr size_t LDS_EPI = LDS_ROWS > LDS_COLS ? LDS_ROWS : LDS_COLS;
  static constexpr size_t LDS = LDS_MAIN > LDS_EPI ? LDS_MAIN : LDS_EPI;
  static_assert(FRAGS % NW == 0, "whole fragments per wave");
  static_assert(!AF || (!CPLX && A_INSTR % NW == 0 && A_CHUNKS_ROW % 16 == 0),
                "raw A: f32, whole instructions per wave, room for the chunk XOR");
  static_assert(LDS <= 160 * 1024, "LDS budget");
};

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

struct GemmArgs {
  ContractArgs c;
  const __bf16* Ap;  // unused when A is split in the kernel
  const __bf16* Bp;
  int64_t RBa, RBb, KB;
};

// LDS float index of raw A element (row m, k) within a stage's A block.
template <int BM>
__device__ inline int araw_slot(int m, int k) {
  return k * BM + ((((m >> 2) ^ (4 * ((k >> 3) & 3)))) << 2) + (m & 3);
}

template <class G>
__global__ __launch_bounds__(G::NT, G::MINB) void k_gemm(GemmArgs g) {
  const ContractArgs& p = g.c;
  __shared__ __attribute__((aligned(16))) char smem[G::LDS];
  const int64_t nwg = (int64_t)p.tiles_m * p.tiles_n * p.H;
  int64_t h;
  int tm, tn;
  dev::decode_tile(blockIdx.x, nwg, p.tiles_m, p.tiles_n, h, tm, tn);
  constexpr int BM = G::BM, BN = G::BN, NT = G::NT, PER = G::PER;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave % G::WAVES_M) * G::WM;
  const int wn = (wave / G::WAVES_M) * G::WN;

  // Panel fragments of this tile for K block kb: A's BM/16 row blocks then B's,
  // each group contiguous in its panel. A stage in LDS holds them in the same
  // order: [A_FRAGS][FRAG] then [B_FRAGS][FRAG].
  const __bf16* Ag = g.Ap + (h * g.KB * g.RBa + (int64_t)tm * (BM / 16)) * PER * FRAG;
  const __bf16* Bg = g.Bp + (h * g.KB * g.RBb + (int64_t)tn * (BN / 16)) * PER * FRAG;
  const int64_t a_step = g.RBa * PER * FRAG;  // bf16 per K block
  const int64_t b_step = g.RBb * PER * FRAG;
  // Raw A (AF): instruction t of this wave moves flat chunks
  // (wave + NW t) * 64 + lane of the [32 k][BM / 4 chunk] block; the chunk at
  // position x of row k holds rows 4 (x ^ 4 ((k / 8) % 4)) .. + 3.
  [[maybe_unused]] const __amdgpu_buffer_rsrc_t rsA = dev::make_rsrc(G::AF ? (const void*)((const float*)p.A + p.ha[h]) : p.A);
  uint32_t a_row[G::AF ? G::AW : 1];
  int a_kl[G::AF ? G::AW : 1];
  if constexpr (G::AF) {
#pragma unroll
    for (int t = 0; t < G::AW; t++) {
      const int f = (wave + G::NW * t) * 64 + lane;
      const int k = f / G::A_CHUNKS_ROW, x = f % G::A_CHUNKS_ROW;
      const int64_t m = (int64_t)tm * BM + 4 * (x ^ (4 * ((k >> 3) & 3)));
      a_row[t] = m < p.M ? (uint32_t)(p.am[m] * 4) : dev::ROW_OOB;
      a_kl[t] = k;
    }
  }
  // Copies K block kb into LDS stage `st`, straight from global memory: one
  // instruction per fragment, lane l moving bytes [16 l, 16 l + 16); with a
  // raw A its block first, then B's fragments.
  auto issue = [&](int64_t kb, int st) {
    char* stage = smem + (size_t)st * G::STAGE_BYTES;
    if constexpr (G::AF) {
#pragma unroll
      for (int t = 0; t < G::AW; t++) {
        [[maybe_unused]] const uint32_t off = a_row[t] + p.ak32[kb * KBLK + a_kl[t]];
#if defined(__HIP_DEVICE_COMPILE__)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rsA, (__attribute__((address_space(3))) void*)(stage + (size_t)(wave + G::NW * t) * 1024), 16, off, 0, 0,
            0);
#endif
      }
      stage += G::A_RAW_BYTES;
    }
#pragma unroll
    for (int i = 0; i < G::PW; i++) {
      const int f = wave + i * G::NW;  // wave-uniform fragment index
      const __bf16* src = f < G::A_FRAGS ? Ag + kb * a_step + (int64_t)f * FRAG
                                         : Bg + kb * b_step + (int64_t)(f - G::A_FRAGS) * FRAG;
      __builtin_amdgcn_global_load_lds((const void*)(src + lane * 8),
                                       (__attribute__((address_space(3))) void*)(stage + (size_t)f * FRAG * 2), 16,
                                       0, 0);
    }
  };

  constexpr int NP = G::TERMS == 2 ? 3 : 6;
  constexpr int PA[6] = {0, 0, 1, 1, 0, 2};
  constexpr int PB[6] = {0, 1, 0, 1, 2, 0};
  constexpr int ACCS = G::CPLX ? 3 : 1;
  v4f acc[ACCS][G::MI][G::NI];
#pragma unroll
  for (int s = 0; s < ACCS; s++)
#pragma unroll
    for (int i = 0; i < G::MI; i++)
#pragma unroll
      for (int j = 0; j < G::NI; j++) acc[s][i][j] = v4f{0, 0, 0, 0};

  // Fragment (row block blk, component c, term x) of a stage; B's follow A's.
  auto frag = [&](const char* stage, int blk, int c, int x) -> v8bf {
    return *(const v8bf*)((const __bf16*)stage + ((size_t)(blk * PER + c * G::TERMS + x) * FRAG + lane * 8));
  };
  v8bf a[G::COMPS][G::TERMS][G::MI], b[G::COMPS][G::TERMS][G::NI], nb[G::COMPS][G::TERMS][G::NI];
  // Raw A (AF): the 8 f32 values lane (row lane % 16, k group lane / 16) feeds
  // the MFMA for row block i of a stage, and their split into a[0][*][i].
  const int ar = lane & 15, ag = lane >> 4;
  auto raw_read = [&](const char* stage, int i, float* v) {
    const float* As = (const float*)stage;
    const int m = wm + 16 * i + ar;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = As[araw_slot<BM>(m, 8 * ag + j)];
  };
  auto raw_split = [&](const float* v, int i, auto& dst) {
    v4u w[G::TERMS];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint32_t t[3];
      split_pair<G::TERMS>(v[2 * q], v[2 * q + 1], t);
#pragma unroll
      for (int x = 0; x < G::TERMS; x++) w[x][q] = t[x];
    }
#pragma unroll
    for (int x = 0; x < G::TERMS; x++) dst[0][x][i] = __builtin_bit_cast(v8bf, w[x]);
  };
  // B's fragments of a stage start after the A block (raw) or A's fragments.
  constexpr size_t B_OFF = G::AF ? G::A_RAW_BYTES : (size_t)G::A_FRAGS * FRAG * 2;

  // Two stages; K block kb lives in stage kb & 1. Registers are one block
  // ahead of the MFMAs: while block kb is multiplied, block kb + 1 is read
  // from LDS (B into nb up front, A's row block i as soon as the MFMAs of
  // block kb no longer need it), so LDS reads overlap the matrix cores.
  // At the top of iteration kb every wave has its own copies of block kb + 1
  // landed (vmcnt) and its reads of block kb done (lgkmcnt); after the
  // barrier block kb + 1 is visible to all and nobody reads stage kb & 1
  // again, which then receives block kb + 2.
  const int64_t nk = g.KB;
  issue(0, 0);
  if (nk > 1) {
    issue(1, 1);
    wait_vmcnt<G::IW>();
  } else {
    wait_vmcnt<0>();
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < G::COMPS; c++)
#pragma unroll
    for (int x = 0; x < G::TERMS; x++) {
      if constexpr (!G::AF) {
#pragma unroll
        for (int i = 0; i < G::MI; i++) a[c][x][i] = frag(smem, (wm >> 4) + i, c, x);
      }
#pragma unroll
      for (int j = 0; j < G::NI; j++) b[c][x][j] = frag(smem + B_OFF, (wn >> 4) + j, c, x);
    }
  float rv[2][8];
  if constexpr (G::AF) {
#pragma unroll
    for (int i = 0; i < G::MI; i++) {
      raw_read(smem, i, rv[0]);
      raw_split(rv[0], i, a);
    }
  }
  if constexpr (G::WIDE) {
    // Column block outer: column block j's B fragments of block kb + 1
    // replace block kb's once its MFMAs are issued (no second B register
    // set). A's block kb + 1 is read up front and split one value pair per
    // column block into `an`. Within a column block the LDS reads and the
    // split's VALU work are placed between the MFMAs (sched_group_barrier),
    // so a wave's own VALU issues while its MFMAs run: f32 transform step,
    // 3 terms, 164 -> 172 TFLOP/s against the same loop with them grouped
    // after the MFMAs. Same stage protocol as the loop below.
    static_assert(G::MI * 4 == G::NI, "one value pair of A per column block");
    v8bf an[1][G::TERMS][G::MI];
    for (int64_t kb = 0; kb < nk; kb++) {
      __builtin_amdgcn_s_waitcnt(7 << 4);
      __syncthreads();
      if (kb + 2 < nk) issue(kb + 2, (int)(kb & 1));
      const char* An = smem + (size_t)((kb + 1) & 1) * G::STAGE_BYTES;
      const char* Bn = An + B_OFF;
      float rw[G::MI][8];
#pragma unroll
      for (int i = 0; i < G::MI; i++) raw_read(An, i, rw[i]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < G::NI; j++) {
#pragma unroll
        for (int i = 0; i < G::MI; i++)
#pragma unroll
          for (int pp = 0; pp < NP; pp++)
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][PA[pp]][i], b[0][PB[pp]][j], acc[0][i][j],
                                                                    0, 0, 0);
        // Column block j - 1's B fragments of block kb + 1 (its MFMAs are all
        // issued) and one split pair go between column block j's MFMAs.
        if (j > 0) {
#pragma unroll
          for (int x = 0; x < G::TERMS; x++) b[0][x][j - 1] = frag(Bn, (wn >> 4) + j - 1, 0, x);
        }
        {
          const int i = j >> 2, q = j & 3;
          uint32_t t[3];
          split_pair<G::TERMS>(rw[i][2 * q], rw[i][2 * q + 1], t);
#pragma unroll
          for (int x = 0; x < G::TERMS; x++) {
            v4u w = __builtin_bit_cast(v4u, an[0][x][i]);
            w[q] = t[x];
            an[0][x][i] = __builtin_bit_cast(v8bf, w);
          }
        }
#pragma unroll
        for (int t = 0; t < G::TERMS; t++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
        }
#pragma unroll
        for (int t = 0; t < G::MI * NP - G::TERMS; t++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int x = 0; x < G::TERMS; x++) b[0][x][G::NI - 1] = frag(Bn, (wn >> 4) + G::NI - 1, 0, x);
#pragma unroll
      for (int x = 0; x < G::TERMS; x++)
#pragma unroll
        for (int i = 0; i < G::MI; i++) a[0][x][i] = an[0][x][i];
    }
  } else
  for (int64_t kb = 0; kb < nk; kb++) {
    // vmcnt(0) and lgkmcnt(0): own copies landed, own LDS reads complete.
    __builtin_amdgcn_s_waitcnt(7 << 4);
    __syncthreads();
    if (kb + 2 < nk) issue(kb + 2, (int)(kb & 1));
    // Block kb + 1 (stale data in the last iteration, never used).
    const char* An = smem + (size_t)((kb + 1) & 1) * G::STAGE_BYTES;
    const char* Bn = An + B_OFF;
#pragma unroll
    for (int c = 0; c < G::COMPS; c++)
#pragma unroll
      for (int x = 0; x < G::TERMS; x++)
#pragma unroll
        for (int j = 0; j < G::NI; j++) nb[c][x][j] = frag(Bn, (wn >> 4) + j, c, x);
    // Scheduling fences keep the reads where they are written (the compiler
    // would otherwise sink them below the MFMAs, serialising LDS and MFMA).
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < G::MI; i++) {
#pragma unroll
      for (int c = 0; c < G::COMPS; c++)
#pragma unroll
        for (int d = 0; d < G::COMPS; d++) {
          const int slot = !G::CPLX ? 0 : (c == 0 && d == 0 ? 0 : (c == 1 && d == 1 ? 1 : 2));
#pragma unroll
          for (int pp = 0; pp < NP; pp++)
#pragma unroll
            for (int j = 0; j < G::NI; j++)
              acc[slot][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c][PA[pp]][i], b[d][PB[pp]][j],
                                                                        acc[slot][i][j], 0, 0, 0);
        }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (G::AF) {
        // Row block i of block kb + 1: its values are read now and split
        // after the next row block's MFMAs (the LDS latency hides under them).
        if (i > 0) raw_split(rv[(i - 1) & 1], i - 1, a);
        raw_read(An, i, rv[i & 1]);
      } else {
#pragma unroll
        for (int c = 0; c < G::COMPS; c++)
#pragma unroll
          for (int x = 0; x < G::TERMS; x++) a[c][x][i] = frag(An, (wm >> 4) + i, c, x);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (G::AF) raw_split(rv[(G::MI - 1) & 1], G::MI - 1, a);
#pragma unroll
    for (int c = 0; c < G::COMPS; c++)
#pragma unroll
      for (int x = 0; x < G::TERMS; x++)
#pragma unroll
        for (int j = 0; j < G::NI; j++) b[c][x][j] = nb[c][x][j];
  }
  __syncthreads();  // every wave is done with the stages before they become the staging tile

  const int lr = lane & 15, lk = lane >> 4;
  if constexpr (!G::CPLX && (size_t)BN * (BM + 4) * 4 <= G::LDS && BM * BN % (4 * NT) == 0) {
    // D contiguous along M with affine offsets, no broadcast copies, 16-byte
    // aligned columns: the tile is staged column by column (pitch BM + 4, so
    // each 16-lane group's 16-byte stores hit distinct banks), a lane's four
    // consecutive rows of an accumulator block as one 16-byte store, and
    // written back as 16-byte vectors.
    float* __restrict__ Dg = (float*)p.D + p.hd[h];
    const float* __restrict__ Cg = (const float*)p.C + p.hc[h];
    const bool ok = p.d_mmajor && p.mn_affine && p.dr == nullptr && p.cd_stride[2] == 1 && p.cd_stride[3] % 4 == 0 &&
                    (uintptr_t)Dg % 16 == 0 &&
                    (!p.read_c || (p.cd_stride[0] == 1 && p.cd_stride[1] % 4 == 0 && (uintptr_t)Cg % 16 == 0));
    if (ok) {
      constexpr int LDM = BM + 4;
      float* Dt = (float*)smem;
#pragma unroll
      for (int i = 0; i < G::MI; i++)
#pragma unroll
        for (int j = 0; j < G::NI; j++) *(v4f*)&Dt[(wn + 16 * j + lr) * LDM + wm + 16 * i + 4 * lk] = acc[0][i][j];
      __syncthreads();
      const float alr = (float)p.alpha[0], ber = (float)p.beta[0];
      const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
#pragma unroll 4
      for (int t = 0; t < BM * BN / (4 * NT); t++) {
        const int e = tid + t * NT;
        const int ml = 4 * (e % (BM / 4)), nl = e / (BM / 4);
        const int64_t m = m0 + ml, n = n0 + nl;
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
      return;
    }
  }
  float* Dre = (float*)smem;
  float* Dim = Dre + BM * G::LDD;
#pragma unroll
  for (int i = 0; i < G::MI; i++)
#pragma unroll
    for (int j = 0; j < G::NI; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = wm + 16 * i + 4 * lk + q, col = wn + 16 * j + lr;
        if constexpr (G::CPLX) {
          Dre[row * G::LDD + col] = acc[0][i][j][q] - acc[1][i][j][q];
          Dim[row * G::LDD + col] = acc[2][i][j][q];
        } else {
          Dre[row * G::LDD + col] = acc[0][i][j][q];
        }
      }
  __syncthreads();
  dev::write_back<float, float, G::CPLX, BM, BN, NT, G::LDD>(p, h, (int64_t)tm * BM, (int64_t)tn * BN, Dre, Dim);
}

struct Shape {
  int bm, bn;
};

Shape shape_of(int tile_id, bool cplx) {
  if (tile_id == STILE_256x128 && !cplx) return {256, 128};
  if (tile_id == STILE_128x128 && !cplx) return {128, 128};
  if (tile_id == STILE_128x64 && cplx) return {128, 64};
  return {64, 64};
}

template <class G>
void launch_gemm(const GemmArgs& g, int64_t nwg, hipStream_t s) {
  hipLaunchKernelGGL(k_gemm<G>, dim3((unsigned)nwg), dim3(G::NT), 0, s, g);
}

bool split_in_gemm(const ContractArgs& a, bool cplx, int tile_id) {
  return a.off32 && a.ak32 != nullptr &&
         split_a_in_gemm(cplx, a.a_kmajor, a.a_vec16, a.N, tile_id);
}

template <bool CPLX, int TERMS>
hipError_t run(const ContractArgs& a, int tile_id, void* panels, hipStream_t s) {
  constexpr int PER = (CPLX ? 2 : 1) * TERMS;
  const bool fuse_a = split_in_gemm(a, CPLX, tile_id);
  const Shape sh = shape_of(tile_id, CPLX);
  const int64_t KB = (a.K + KBLK - 1) / KBLK;
  const int64_t tiles_m = (a.M + sh.bm - 1) / sh.bm, tiles_n = (a.N + sh.bn - 1) / sh.bn;
  // Row blocks per panel slice: whole tiles, and a multiple of 4 for the packer.
  auto pad = [](int64_t rows, int bm) {
    int64_t rb = (rows + bm - 1) / bm * (bm / 16);
    return (rb + 3) / 4 * 4;
  };
  const int64_t RBa = pad(a.M, sh.bm), RBb = pad(a.N, sh.bn);
  __bf16* Ap = (__bf16*)panels;
  __bf16* Bp = fuse_a ? Ap : Ap + a.H * KB * RBa * PER * FRAG;
  PackArgs pa{a.A, a.am, a.ak, a.ha, a.M, a.K, a.H, RBa, KB, a.conj_a, a.a_kvec, a.a_kmajor, Ap};
  PackArgs pb{a.B, a.bn, a.bk, a.hb, a.N, a.K, a.H, RBb, KB, a.conj_b, a.b_kvec, a.b_kmajor, Bp};
  for (const PackArgs* q : {&pa, &pb}) {
    if (q == &pa && fuse_a) continue;
    const int64_t grid = q->H * q->KB * (q->RB / 4);
    if (grid > max_grid(256)) return hipErrorInvalidConfiguration;
    if (q->vec) hipLaunchKernelGGL((k_pack<CPLX, TERMS, true>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    // K-major without 16-byte vectors: lanes along K through LDS (c32
    // benchmark contraction, 3 terms: 110 -> 129 TFLOP/s). M-major operands
    // keep the direct gather, whose 16-row runs measured faster there (f32
    // transform step: 98 against 93).
    else if (q->kmajor) hipLaunchKernelGGL((k_pack_lds<CPLX, TERMS, true>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    else hipLaunchKernelGGL((k_pack<CPLX, TERMS, false>), dim3((unsigned)grid), dim3(256), 0, s, *q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  GemmArgs g{a, Ap, Bp, RBa, RBb, KB};
  g.c.tiles_m = (int)tiles_m;
  g.c.tiles_n = (int)tiles_n;
  const int64_t nwg = tiles_m * tiles_n * a.H;
  if (nwg > max_grid(512)) return hipErrorInvalidConfiguration;
  if constexpr (!CPLX) {
    if (fuse_a) {
      if (sh.bm == 256) launch_gemm<GCfg<CPLX, TERMS, 256, 128, 64, 64, 2, true>>(g, nwg, s);
      else if (sh.bm == 128) launch_gemm<GCfg<CPLX, TERMS, 128, 128, 32, 128, 2, true>>(g, nwg, s);
      else launch_gemm<GCfg<CPLX, TERMS, 64, 64, 32, 32, 2, true>>(g, nwg, s);
      return hipGetLastError();
    }
    if (sh.bm == 256) launch_gemm<GCfg<CPLX, TERMS, 256, 128, 64, 64, 2>>(g, nwg, s);
    else if (sh.bm == 128) launch_gemm<GCfg<CPLX, TERMS, 128, 128, 64, 64, 2>>(g, nwg, s);
    else launch_gemm<GCfg<CPLX, TERMS, 64, 64, 32, 32, 2>>(g, nwg, s);
  } else {
    if (sh.bm == 128) launch_gemm<GCfg<CPLX, TERMS, 128, 64, 32, 32, 2>>(g, nwg, s);
    else launch_gemm<GCfg<CPLX, TERMS, 64, 64, 32, 32, 2>>(g, nwg, s);
  }
  return hipGetLastError();
}

}  // namespace

// An f32 A gathered along M in 16-byte runs is split inside the GEMM (no A
// panels): the transform step's 537 MB A was otherwise read, written as 805 MB
// of panels and read again. Each of a tile column's workgroups splits its own
// copy, so wide products (many tile columns) keep the panels, which split
// every element once (transform step, N = 512: 100.9 -> 115 TFLOP/s; the
// 4096^3 GEMM, 32 tile columns: 170 with panels against 150).
bool split_a_in_gemm(bool cplx, bool a_kmajor, bool a_vec16, int64_t N, int tile_id) {
  const int64_t bn = shape_of(tile_id, cplx).bn;
  return !cplx && !a_kmajor && a_vec16 && (N + bn - 1) / bn <= 8;
}

size_t split_panel_bytes(const ContractArgs& a, DType t, int nterms, int tile_id) {
  const bool cplx = t == DType::C32;
  const Shape sh = shape_of(tile_id, cplx);
  const int64_t KB = (a.K + KBLK - 1) / KBLK;
  auto pad = [](int64_t rows, int bm) {
    int64_t rb = (rows + bm - 1) / bm * (bm / 16);
    return (rb + 3) / 4 * 4;
  };
  const int64_t per = (cplx ? 2 : 1) * nterms;
  const int64_t rows = (split_in_gemm(a, cplx, tile_id) ? 0 : pad(a.M, sh.bm)) + pad(a.N, sh.bn);
  return (size_t)a.H * KB * rows * per * FRAG * 2;
}

hipError_t launch_contract_split(const ContractArgs& a, DType t, int nterms, int tile_id, void* panels,
                                 hipStream_t s) {
  if (t == DType::F32) return nterms == 2 ? run<false, 2>(a, tile_id, panels, s) : run<false, 3>(a, tile_id, panels, s);
  if (t == DType::C32) return nterms == 2 ? run<true, 2>(a, tile_id, panels, s) : run<true, 3>(a, tile_id, panels, s);
  return hipErrorInvalidValue;
}

}  // namespace tp
