// Element-parallel kernels: the output pass D = op_D(beta * op_C(C)) used for
// an empty contraction or alpha == 0 (crates/tprims-kernel/src/pack/writeback.rs:495
// scale_only; crates/tprims-contract/src/strategy/elementwise.rs:210), the
// elementwise (all-batch) strategy for tiny M x N and K with many batch entries
// (strategy/elementwise.rs:194, the Hadamard-type products), and the zero fill
// of a layout used by the adjoints. All are HBM-bound element passes.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace tp {
namespace {

using dev::Store;

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
    S* D = (S*)p.D;
    const int64_t off = p.hd[h] + (p.mn_affine ? m * p.cd_stride[2] + n * p.cd_stride[3] : p.dm[m] + p.dn[n]);
    const int64_t coff =
        !p.read_c ? 0 : p.hc[h] + (p.mn_affine ? m * p.cd_stride[0] + n * p.cd_stride[1] : p.cm[m] + p.cn[n]);
    // Broadcast copies (adjoints) read their own previous value: C is D there.
    const int64_t copies = p.dr == nullptr ? 1 : p.R;
    for (int64_t r = 0; r < copies; r++) {
      const int64_t roff = p.dr == nullptr ? 0 : p.dr[r];
      T vr = 0, vi = 0;
      if (p.read_c) {
        S cv = ((const S*)p.C)[coff + roff];
        T cr = St::re(cv), ci = St::im(cv);
        if (p.conj_c) ci = -ci;
        vr = ber * cr - bei * ci;
        vi = ber * ci + bei * cr;
      }
      if (p.conj_d) vi = -vi;
      D[off + roff] = St::make(vr, vi);
    }
  }
}

// One thread per output element (h, m, n), K summed serially in the storage
// type, then the output-pass write-back. Elements run along D's contiguous
// role: the batch index when M = N = 1, else M or N as the write-back does.
template <typename T, bool CPLX>
__global__ void k_elementwise(ContractArgs p) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int64_t M = p.M, N = p.N, H = p.H, K = p.K;
  const int64_t total = M * N * H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t m, n, h;
    if (M * N == 1) {
      m = n = 0;
      h = e;
    } else if (p.d_mmajor) {
      m = e % M;
      const int64_t r = e / M;
      n = r % N;
      h = r / N;
    } else {
      n = e % N;
      const int64_t r = e / N;
      m = r % M;
      h = r / M;
    }
    const int64_t oa = p.h_affine ? h * p.h_stride[0] : p.ha[h];
    const int64_t ob = p.h_affine ? h * p.h_stride[1] : p.hb[h];
    const S* A = (const S*)p.A + oa + p.am[m];
    const S* B = (const S*)p.B + ob + p.bn[n];
    T sr = 0, si = 0;
    for (int64_t k = 0; k < K; k++) {
      const S av = A[p.ak[k]], bv = B[p.bk[k]];
      const T xr = St::re(av), yr = St::re(bv);
      if constexpr (CPLX) {
        const T xi = p.conj_a ? -St::im(av) : St::im(av);
        const T yi = p.conj_b ? -St::im(bv) : St::im(bv);
        sr += xr * yr - xi * yi;
        si += xr * yi + xi * yr;
      } else {
        sr += xr * yr;
      }
    }
    dev::store_element<T, CPLX>(p, h, m, n, sr, si);
  }
}

// Hadamard products (M = N = K = 1, one folded batch axis of stride 1 in
// every operand, no broadcast copies, 16-byte aligned bases):
// D[h] = op_D(alpha op_A(A[h]) op_B(B[h]) + beta op_C(C[h])), VW = 16 /
// element size elements per lane per step in 16-byte loads and stores
// (2^27 elements: f32 2.4 -> 5.85 TB/s, f64 4.0 -> 5.6 against the general
// elementwise kernel).
template <typename T, bool CPLX, int VW>
__global__ void k_hadamard(ContractArgs p) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int64_t H = p.H;
  // M = N = K = 1: the one-element M, N and K scatter vectors hold offset 0.
  const S* __restrict__ A = (const S*)p.A;
  const S* __restrict__ B = (const S*)p.B;
  const S* __restrict__ C = p.read_c ? (const S*)p.C : nullptr;
  S* __restrict__ D = (S*)p.D;
  const T alr = (T)p.alpha[0], ali = (T)p.alpha[1], ber = (T)p.beta[0], bei = (T)p.beta[1];
  const bool ca = p.conj_a, cb = p.conj_b, cc = p.conj_c, cd = p.conj_d, rc = p.read_c;
  auto one = [&](S av, S bv, S cv) {
    T xr = St::re(av), yr = St::re(bv), sr, si;
    if constexpr (CPLX) {
      const T xi = ca ? -St::im(av) : St::im(av), yi = cb ? -St::im(bv) : St::im(bv);
      sr = xr * yr - xi * yi;
      si = xr * yi + xi * yr;
    } else {
      sr = xr * yr;
      si = 0;
    }
    T wr = alr * sr - ali * si, wi = alr * si + ali * sr;
    if (rc) {
      const T cr = St::re(cv), ci = cc ? -St::im(cv) : St::im(cv);
      wr += ber * cr - bei * ci;
      wi += ber * ci + bei * cr;
    }
    if (cd) wi = -wi;
    return St::make(wr, wi);
  };
  const int64_t step = (int64_t)gridDim.x * blockDim.x * VW;
  int64_t h = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * VW;
  static_assert(VW > 1, "16-byte vectors");
  // Streaming (non-temporal) loads and stores: every byte is touched once
  // (2^27 elements: f64 5.25 -> 5.62 TB/s, f32 5.29 -> 5.85 against plain
  // loads and stores).
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  {
    struct alignas(16) V {
      S x[VW];
    };
    for (; h + VW <= H; h += step) {
      const V av = __builtin_bit_cast(V, __builtin_nontemporal_load((const u4*)(A + h)));
      const V bv = __builtin_bit_cast(V, __builtin_nontemporal_load((const u4*)(B + h)));
      V cv{};
      if (rc) cv = __builtin_bit_cast(V, __builtin_nontemporal_load((const u4*)(C + h)));
      V dv;
#pragma unroll
      for (int u = 0; u < VW; u++) dv.x[u] = one(av.x[u], bv.x[u], cv.x[u]);
      __builtin_nontemporal_store(__builtin_bit_cast(u4, dv), (u4*)(D + h));
    }
    // The ragged end: at most VW - 1 elements, taken by the first lanes.
    const int64_t tail = H - H / VW * VW;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < tail) {
      const int64_t e = H / VW * VW + t;
      D[e] = one(A[e], B[e], rc ? C[e] : S{});
    }
  }
}

// Split-K reduction: element (h, m, n) sums the ksplit partials of the
// workspace in split order (deterministic) and does the write-back.
template <typename T, bool CPLX>
__global__ void k_splitk_reduce(ContractArgs p) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int64_t per = p.M * p.N, total = per * p.H;
  const S* W = (const S*)p.work;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    // Element e of one split's workspace: [h][n][m] when D is m-major, else
    // [h][m][n] (the layout the partial tiles were written in).
    int64_t m, n, h;
    if (p.d_mmajor) {
      m = e % p.M;
      const int64_t r = e / p.M;
      n = r % p.N;
      h = r / p.N;
    } else {
      n = e % p.N;
      const int64_t r = e / p.N;
      m = r % p.M;
      h = r / p.M;
    }
    T sr = 0, si = 0;
    for (int s = 0; s < p.ksplit; s++) {
      const S v = W[(int64_t)s * total + e];
      sr += St::re(v);
      si += St::im(v);
    }
    dev::store_element<T, CPLX>(p, h, m, n, sr, si);
  }
}

// The same over a two-dimensional grid: blockIdx.y is one line of the
// workspace (h and the slower output index), so a lane finds its element
// without 64-bit divisions (the line's split is uniform across the block).
template <typename T, bool CPLX>
__global__ void k_splitk_reduce_lines(ContractArgs p) {
  typedef Store<T, CPLX> St;
  typedef typename St::S S;
  const int64_t total = p.M * p.N * p.H;
  const int64_t inner = p.d_mmajor ? p.M : p.N, outer = p.d_mmajor ? p.N : p.M;
  const int64_t line = blockIdx.y, o = line % outer, h = line / outer;
  const S* W = (const S*)p.work + line * inner;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < inner; i += (int64_t)gridDim.x * blockDim.x) {
    T sr = 0, si = 0;
    for (int s = 0; s < p.ksplit; s++) {
      const S v = W[(int64_t)s * total + i];
      sr += St::re(v);
      si += St::im(v);
    }
    if (p.d_mmajor) dev::store_element<T, CPLX>(p, h, i, o, sr, si);
    else dev::store_element<T, CPLX>(p, h, o, i, sr, si);
  }
}

struct ZeroLayout {
  int rank;
  int64_t dims[8];
  int64_t strides[8];
  int64_t total;
};

// Zero n bytes starting at a 16-byte aligned address (dense layouts). Each
// workgroup clears one contiguous chunk, eight 16-byte stores per lane per
// step (8 GiB: 5.9 TB/s, against 5.4 for a grid-stride loop; the c64 adjoint
// of the benchmark VJP, 34 GB, 1.5 % faster than with four).
__global__ void k_zero_bytes(uint4* base, int64_t n16, char* tail, int64_t ntail) {
  const int64_t per = (n16 + gridDim.x - 1) / gridDim.x;
  const int64_t b = (int64_t)blockIdx.x * per, end = b + per < n16 ? b + per : n16;
  for (int64_t e = b + threadIdx.x; e < end; e += 8 * (int64_t)blockDim.x) {
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (e + u * blockDim.x < end) base[e + u * blockDim.x] = uint4{0, 0, 0, 0};
  }
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

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

// Launches the 16-byte-vector Hadamard kernel and returns true, or returns
// false (nothing launched) when the strides or alignment do not allow it:
// the general elementwise kernel measured faster than the one-element form
// (c64 2^27: 5.4 against 4.9 TB/s).
template <typename T, bool CPLX>
bool launch_hadamard(const ContractArgs& a, hipStream_t s) {
  typedef typename Store<T, CPLX>::S S;
  constexpr int VW = 16 / (int)sizeof(S);
  auto at = [](const void* base) { return (uintptr_t)base % 16 == 0; };
  const bool unit = a.h_stride[0] == 1 && a.h_stride[1] == 1 && a.h_stride[3] == 1 && (!a.read_c || a.h_stride[2] == 1);
  const bool aligned = at(a.A) && at(a.B) && at(a.D) && (!a.read_c || at(a.C));
  const int threads = 256;
  if constexpr (VW > 1) {
    if (unit && aligned) {
      int64_t blocks = ((a.H + VW - 1) / VW + threads - 1) / threads;
      if (blocks > 65536) blocks = 65536;
      if (blocks < 1) blocks = 1;
      hipLaunchKernelGGL((k_hadamard<T, CPLX, VW>), dim3(blocks), dim3(threads), 0, s, a);
      return true;
    }
  }
  return false;
}

hipError_t launch_elementwise(const ContractArgs& a, DType t, hipStream_t s) {
  const int64_t total = a.M * a.N * a.H;
  if (total == 0) return hipSuccess;
  // One product per output along one folded batch axis: the Hadamard kernel.
  if (a.M == 1 && a.N == 1 && a.K == 1 && a.h_affine && a.dr == nullptr) {
    bool done = false;
    switch (t) {
      case DType::F64: done = launch_hadamard<double, false>(a, s); break;
      case DType::F32: done = launch_hadamard<float, false>(a, s); break;
      case DType::C64: done = launch_hadamard<double, true>(a, s); break;
      case DType::C32: done = launch_hadamard<float, true>(a, s); break;
    }
    if (done) return hipGetLastError();
  }
  const int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case DType::F64: hipLaunchKernelGGL((k_elementwise<double, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::F32: hipLaunchKernelGGL((k_elementwise<float, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C64: hipLaunchKernelGGL((k_elementwise<double, true>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C32: hipLaunchKernelGGL((k_elementwise<float, true>), dim3(blocks), dim3(threads), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_splitk_reduce(const ContractArgs& a, DType t, hipStream_t s) {
  const int64_t total = a.M * a.N * a.H;
  if (total == 0) return hipSuccess;
  const int threads = 256;
  const int64_t inner = a.d_mmajor ? a.M : a.N, lines = total / inner;
  if (lines <= 65535) {  // one grid row per workspace line
    int64_t bx = (inner + threads - 1) / threads;
    if (bx > 1024) bx = 1024;
    const dim3 grid((unsigned)bx, (unsigned)lines);
    switch (t) {
      case DType::F64: hipLaunchKernelGGL((k_splitk_reduce_lines<double, false>), grid, dim3(threads), 0, s, a); break;
      case DType::F32: hipLaunchKernelGGL((k_splitk_reduce_lines<float, false>), grid, dim3(threads), 0, s, a); break;
      case DType::C64: hipLaunchKernelGGL((k_splitk_reduce_lines<double, true>), grid, dim3(threads), 0, s, a); break;
      case DType::C32: hipLaunchKernelGGL((k_splitk_reduce_lines<float, true>), grid, dim3(threads), 0, s, a); break;
    }
    return hipGetLastError();
  }
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case DType::F64: hipLaunchKernelGGL((k_splitk_reduce<double, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::F32: hipLaunchKernelGGL((k_splitk_reduce<float, false>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C64: hipLaunchKernelGGL((k_splitk_reduce<double, true>), dim3(blocks), dim3(threads), 0, s, a); break;
    case DType::C32: hipLaunchKernelGGL((k_splitk_reduce<float, true>), dim3(blocks), dim3(threads), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_zero_layout(void* base, DType t, int rank, const int64_t* dims, const int64_t* strides,
                              hipStream_t s) {
  if (rank > 8) return hipErrorInvalidValue;
  const int64_t es = elem_size(t);
  // Normalise: positive strides (moving the origin), extent-1 axes dropped,
  // sorted by stride, adjacent axes folded. The set of addressed elements is
  // unchanged, and a dense layout becomes one axis of stride 1.
  struct Ax {
    int64_t e, s;
  } ax[8];
  int n = 0;
  char* origin = (char*)base;
  for (int i = 0; i < rank; i++) {
    if (dims[i] == 0) return hipSuccess;
    if (dims[i] == 1) continue;
    int64_t s = strides[i];
    if (s < 0) {
      origin += (dims[i] - 1) * s * es;
      s = -s;
    }
    ax[n++] = {dims[i], s};
  }
  for (int i = 1; i < n; i++)
    for (int j = i; j > 0 && ax[j].s < ax[j - 1].s; j--) {
      Ax tmp = ax[j];
      ax[j] = ax[j - 1];
      ax[j - 1] = tmp;
    }
  int m = 0;
  for (int i = 0; i < n; i++) {
    if (m > 0 && ax[i].s == ax[m - 1].s * ax[m - 1].e) ax[m - 1].e *= ax[i].e;
    else ax[m++] = ax[i];
  }
  if (m == 0 || (m == 1 && ax[0].s == 1)) {
    const int64_t bytes = (m == 0 ? 1 : ax[0].e) * es;
    const uintptr_t addr = (uintptr_t)origin;
    const int64_t head = addr % 16 ? (int64_t)(16 - addr % 16) : 0;
    if (head > 0 || bytes < 64) {
      // Unaligned start or tiny: fall through to the element loop below.
    } else {
      const int64_t n16 = bytes / 16, ntail = bytes - n16 * 16;
      int64_t blocks = (n16 + 1023) / 1024;
      if (blocks > 16384) blocks = 16384;
      if (blocks < 1) blocks = 1;
      hipLaunchKernelGGL(k_zero_bytes, dim3(blocks), dim3(256), 0, s, (uint4*)origin, n16, origin + n16 * 16, ntail);
      return hipGetLastError();
    }
  }
  ZeroLayout z;
  z.rank = m;
  z.total = 1;
  for (int i = 0; i < m; i++) {
    z.dims[i] = ax[i].e;
    z.strides[i] = ax[i].s;
    z.total *= ax[i].e;
  }
  base = origin;
  const int threads = 256;
  int64_t blocks = (z.total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  switch (t) {
    case DType::F32: hipLaunchKernelGGL(k_zero_layout<float>, dim3(blocks), dim3(threads), 0, s, (float*)base, z); break;
    case DType::F64:
    case DType::C32:
      hipLaunchKernelGGL(k_zero_layout<double>, dim3(blocks), dim3(threads), 0, s, (double*)base, z);
      break;
    case DType::C64:
      hipLaunchKernelGGL(k_zero_layout<dev::Cx<double>>, dim3(blocks), dim3(threads), 0, s, (dev::Cx<double>*)base, z);
      break;
  }
  return hipGetLastError();
}

}  // namespace tp
