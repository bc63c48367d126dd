// Output-only kernels: the output pass D = op_D(beta * op_C(C)) used for an
// empty contraction or alpha == 0 (crates/tprims-kernel/src/pack/writeback.rs:495
// scale_only; crates/tprims-contract/src/strategy/elementwise.rs:210), and the
// zero fill of a layout used by the adjoints. Both are HBM-bound element passes.
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
    const int64_t off = p.hd[h] + p.dm[m] + p.dn[n];
    const int64_t coff = p.read_c ? p.hc[h] + p.cm[m] + p.cn[n] : 0;
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
