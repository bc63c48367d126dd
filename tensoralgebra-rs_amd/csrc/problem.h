// The validated description of one contraction
//   D = op_D(alpha * sum_K op_A(A) * op_B(B) + beta * op_C(C))
// built by the one lowering of labels into roles. Restates
// crates/tprims-contract/src/api/problem.rs (Problem, Roles, RoleAxis) and
// crates/tprims-contract/src/api/validate.rs (lower, reduce, injective, span_of).
#pragma once
#include <cstdint>
#include <vector>

#include "common.h"

namespace tp {

struct Layout {
  std::vector<int64_t> dims;
  std::vector<int64_t> strides;
  int64_t offset = 0;  // element index of the all-zero index tuple
  int rank() const { return (int)dims.size(); }
};

struct Operand {
  Layout layout;
  bool conj = false;
};

// One reduced logical axis; strides[4] = A, B, C, D (zero where absent).
struct RoleAxis {
  int64_t label = 0;
  int64_t extent = 0;
  int64_t s[4] = {0, 0, 0, 0};
  bool in_a = false, in_b = false;
};

struct Roles {
  std::vector<RoleAxis> m, n, k, h;
  // Output-only labels (TAPP case 5, a broadcast). Rejected for products; the
  // adjoint problems of differentiation use them.
  std::vector<RoleAxis> r;
};

struct Span {
  bool present = false;  // false: the operand addresses nothing (an extent is 0)
  int64_t lo = 0, hi = 0;  // lowest and highest addressed element (inclusive)
};

enum OperandId { OP_A = 0, OP_B = 1, OP_C = 2, OP_D = 3 };

struct Problem {
  DType dtype = DType::F64;
  Operand a, b, c, d;
  bool c_separate = true;  // TAPP always passes a separate C
  std::vector<int64_t> la, lb, lc, ld;
  Roles roles;
  Span spans[4];
  bool c_matches_d = true;
  bool k_empty = false;
  bool out_empty = false;

  int64_t macs() const;  // product of every M, N, K, H extent
};

// The lowering (validate.rs:207-308). allow_broadcast (adjoint problems only)
// admits output-only labels into roles.r instead of refusing them with
// ERR_UNSUPPORTED, and proves output injectivity with the addressed span
// instead of the reference's coarser stride*extent reach.
Problem lower(DType dtype, const Operand& a, const std::vector<int64_t>& la, const Operand& b,
              const std::vector<int64_t>& lb, const Operand* c, const std::vector<int64_t>* lc,
              const Operand& d, const std::vector<int64_t>& ld, bool allow_broadcast = false);

// Conservative sorted-stride injectivity test (validate.rs:122-143).
bool injective_layout(const std::vector<int64_t>& dims, const std::vector<int64_t>& strides);

}  // namespace tp
