// Shared host-side vocabulary of libtprims (MI355X build): element types,
// status codes and the exception that carries a status across the library
// until the C boundary turns it into a return value.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

namespace tp {

// Status codes of the C ABI (crates/tprims-capi/src/status.rs:11-42).
enum Status : int32_t {
  OK = 0,
  ERR_INVALID_ARGUMENT = 1,
  ERR_SHAPE = 2,
  ERR_DTYPE = 3,
  ERR_DEVICE = 4,
  ERR_READ_ONLY = 5,
  ERR_ALIASED = 6,
  ERR_WOULD_MATERIALIZE = 7,
  ERR_NUMERICAL = 8,
  BUSY = 9,
  ERR_WOULD_DEADLOCK = 10,
  ERR_PANIC = 12,
  ERR_INTERNAL = 13,
  ERR_LABELS = 14,
  ERR_UNSUPPORTED = 15,
};

const char* describe(int32_t status);

struct Error : std::runtime_error {
  int32_t status;
  Error(int32_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(int32_t s, const std::string& m) { throw Error(s, m); }

// Storage type; the numbering is TAPP's datatype enumeration.
enum class DType : int32_t { F32 = 0, F64 = 1, C32 = 2, C64 = 3 };

inline int elem_size(DType t) {
  switch (t) {
    case DType::F32: return 4;
    case DType::F64: return 8;
    case DType::C32: return 8;
    case DType::C64: return 16;
  }
  return 0;
}
inline bool is_complex(DType t) { return t == DType::C32 || t == DType::C64; }
inline bool is_double(DType t) { return t == DType::F64 || t == DType::C64; }
inline const char* dtype_name(DType t) {
  switch (t) {
    case DType::F32: return "f32";
    case DType::F64: return "f64";
    case DType::C32: return "c32";
    case DType::C64: return "c64";
  }
  return "?";
}

// Plan-configuration enumerations (tprims/hip.h).
enum CplxMethod : int32_t { CPLX_NATIVE = 0, CPLX_1M = 1, CPLX_3M = 2 };
enum Precision : int32_t { PREC_STORAGE = 0, PREC_SPLIT_BF16_2 = 1, PREC_SPLIT_BF16_3 = 2, PREC_FP8 = 3 };

}  // namespace tp
