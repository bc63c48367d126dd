// The one lowering of operand metadata into roles; restates
// crates/tprims-contract/src/api/validate.rs:47-308 step by step:
//   1. per-operand collapse of repeated labels onto a diagonal (strides add),
//   2. a label table over A, B, C, D with extent agreement,
//   3. classification into H (A,B,D), M (A,D), N (B,D), K (A or B, not D),
//      and output-only labels refused (or kept as broadcasts for adjoints),
//   4. checked address spans, 5. conservative D injectivity, 6. C == D proof.
#include "problem.h"

#include <algorithm>
#include <cstdlib>
#include <string>

namespace tp {

namespace {

struct Mode {
  int64_t label, extent, stride;
};

const char* opname(int i) {
  static const char* n[4] = {"A", "B", "C", "D"};
  return n[i];
}

// validate.rs:47-94
std::vector<Mode> reduce(int which, const Layout& l, const std::vector<int64_t>& labels) {
  if ((size_t)l.rank() != labels.size())
    fail(ERR_LABELS, std::string(opname(which)) + " has " + std::to_string(l.rank()) + " modes but " +
                         std::to_string(labels.size()) + " labels");
  std::vector<Mode> out;
  __int128 product = 1;
  for (size_t k = 0; k < labels.size(); k++) {
    int64_t e = l.dims[k], s = l.strides[k];
    if (e < 0) fail(ERR_SHAPE, "negative extent " + std::to_string(e) + " on axis " + std::to_string(k));
    auto it = std::find_if(out.begin(), out.end(), [&](const Mode& m) { return m.label == labels[k]; });
    if (it != out.end()) {
      if (it->extent != e)
        fail(ERR_SHAPE, "label " + std::to_string(labels[k]) + " has extents " + std::to_string(it->extent) +
                            " and " + std::to_string(e));
      __int128 sum = (__int128)it->stride + s;
      if (sum > INT64_MAX || sum < INT64_MIN) fail(ERR_SHAPE, "overflow: repeated-label stride sum");
      it->stride = (int64_t)sum;
    } else {
      product *= e;
      if (product > INT64_MAX) fail(ERR_SHAPE, "overflow: extent product");
      out.push_back({labels[k], e, s});
    }
  }
  return out;
}

// validate.rs:96-119: addressed element range over the original axes.
Span span_of(DType t, const Layout& l) {
  Span sp;
  __int128 lo = l.offset, hi = l.offset;
  for (int k = 0; k < l.rank(); k++) {
    if (l.dims[k] == 0) return sp;
    __int128 reach = (__int128)(l.dims[k] - 1) * l.strides[k];
    if (reach < 0) lo += reach; else hi += reach;
  }
  const __int128 limit = INT64_MAX;
  const __int128 es = elem_size(t);
  __int128 alo = lo < 0 ? -lo : lo;
  if (alo * es > limit || (hi + 1) * es > limit) fail(ERR_SHAPE, "overflow: address range");
  sp.present = true;
  sp.lo = (int64_t)lo;
  sp.hi = (int64_t)hi;
  return sp;
}

// validate.rs:122-143: after sorting by |stride|, each stride must clear the
// reach of the smaller ones, reach = stride * extent (conservative). With
// `tight`, the reach is the span actually addressed, 1 + sum (extent-1)|stride|,
// which still proves injectivity and also admits diagonal outputs followed by
// further axes (the adjoints write such diagonals).
bool injective_modes(std::vector<std::pair<int64_t, int64_t>> axes /* extent, stride */, bool tight = false) {
  for (auto& a : axes)
    if (a.first == 0) return true;
  std::vector<std::pair<__int128, __int128>> v;
  for (auto& a : axes)
    if (a.first > 1) v.push_back({(__int128)a.second < 0 ? -(__int128)a.second : (__int128)a.second, a.first});
  std::stable_sort(v.begin(), v.end(), [](auto& x, auto& y) { return x.first < y.first; });
  __int128 reach = 1;
  for (auto& [s, e] : v) {
    if (s < reach) return false;
    reach = tight ? reach + s * (e - 1) : s * e;
  }
  return true;
}

struct Label {
  int64_t id, extent;
  int64_t s[4] = {0, 0, 0, 0};
  bool in[4] = {false, false, false, false};
};

void merge(std::vector<Label>& table, const std::vector<Mode>& modes, int slot) {
  for (const Mode& m : modes) {
    auto it = std::find_if(table.begin(), table.end(), [&](const Label& l) { return l.id == m.label; });
    if (it == table.end()) {
      Label l;
      l.id = m.label;
      l.extent = m.extent;
      table.push_back(l);
      it = table.end() - 1;
    } else if (it->extent != m.extent) {
      fail(ERR_SHAPE, "label " + std::to_string(m.label) + " has extents " + std::to_string(it->extent) + " and " +
                          std::to_string(m.extent));
    }
    it->s[slot] = m.stride;
    it->in[slot] = true;
  }
}

}  // namespace

bool injective_layout(const std::vector<int64_t>& dims, const std::vector<int64_t>& strides) {
  std::vector<std::pair<int64_t, int64_t>> v;
  for (size_t i = 0; i < dims.size(); i++) v.push_back({dims[i], strides[i]});
  return injective_modes(v);
}

int64_t Problem::macs() const {
  int64_t p = 1;
  for (auto* role : {&roles.m, &roles.n, &roles.k, &roles.h})
    for (auto& ax : *role) p *= ax.extent;
  return p;
}

Problem lower(DType dtype, const Operand& a, const std::vector<int64_t>& la, const Operand& b,
              const std::vector<int64_t>& lb, const Operand* c, const std::vector<int64_t>* lc,
              const Operand& d, const std::vector<int64_t>& ld, bool allow_broadcast) {
  if ((c != nullptr) != (lc != nullptr)) fail(ERR_LABELS, "a separate C needs its labels");
  Problem p;
  p.dtype = dtype;
  p.a = a;
  p.b = b;
  p.d = d;
  p.la = la;
  p.lb = lb;
  p.ld = ld;
  p.c_separate = c != nullptr;
  p.c = c ? *c : d;
  p.lc = lc ? *lc : ld;

  auto ra = reduce(OP_A, a.layout, la);
  auto rb = reduce(OP_B, b.layout, lb);
  auto rd = reduce(OP_D, d.layout, ld);
  auto rc = c ? reduce(OP_C, c->layout, *lc) : rd;

  std::vector<Label> table;
  merge(table, ra, OP_A);
  merge(table, rb, OP_B);
  merge(table, rc, OP_C);
  merge(table, rd, OP_D);

  for (auto& l : table)
    if (l.in[OP_C] != l.in[OP_D]) fail(ERR_LABELS, "C and D are labelled differently");

  for (auto& l : table) {
    RoleAxis ax;
    ax.label = l.id;
    ax.extent = l.extent;
    for (int i = 0; i < 4; i++) ax.s[i] = l.s[i];
    ax.in_a = l.in[OP_A];
    ax.in_b = l.in[OP_B];
    bool ia = l.in[OP_A], ib = l.in[OP_B], id = l.in[OP_D];
    if (ia && ib && id) p.roles.h.push_back(ax);
    else if (ia && !ib && id) p.roles.m.push_back(ax);
    else if (!ia && ib && id) p.roles.n.push_back(ax);
    else if ((ia || ib) && !id) p.roles.k.push_back(ax);
    else if (allow_broadcast) p.roles.r.push_back(ax);
    else fail(ERR_UNSUPPORTED, "label " + std::to_string(l.id) + " appears only in the output (a broadcast)");
  }

  p.spans[OP_A] = span_of(dtype, a.layout);
  p.spans[OP_B] = span_of(dtype, b.layout);
  p.spans[OP_C] = span_of(dtype, p.c.layout);
  p.spans[OP_D] = span_of(dtype, d.layout);

  {
    std::vector<std::pair<int64_t, int64_t>> v;
    for (auto& m : rd) v.push_back({m.extent, m.stride});
    if (!injective_modes(v, allow_broadcast))
      fail(ERR_ALIASED, "the output addresses one element through two index tuples");
  }

  if (c) {
    bool same = c->layout.offset == d.layout.offset;
    for (auto& m : rd) {
      bool found = false;
      for (auto& x : rc)
        if (x.label == m.label && x.stride == m.stride) found = true;
      same = same && found;
    }
    p.c_matches_d = same;
  } else {
    p.c_matches_d = true;
  }

  p.k_empty = false;
  for (auto& x : p.roles.k)
    if (x.extent == 0) p.k_empty = true;
  p.out_empty = false;
  for (auto* role : {&p.roles.m, &p.roles.n, &p.roles.h, &p.roles.r})
    for (auto& x : *role)
      if (x.extent == 0) p.out_empty = true;
  return p;
}

const char* describe(int32_t status) {
  switch (status) {
    case OK: return "success";
    case ERR_INVALID_ARGUMENT: return "null pointer, invalid handle or invalid argument";
    case ERR_SHAPE: return "inconsistent or overflowing extents or strides";
    case ERR_DTYPE: return "unsupported or mismatched data type";
    case ERR_DEVICE: return "unsupported device";
    case ERR_READ_ONLY: return "output is read-only";
    case ERR_ALIASED: return "output aliases an operand or overlaps itself";
    case ERR_WOULD_MATERIALIZE: return "the operation would copy an operand";
    case ERR_NUMERICAL: return "numerical failure";
    case BUSY: return "executor has calls in flight";
    case ERR_WOULD_DEADLOCK: return "executor destroyed from one of its own workers";
    case ERR_PANIC: return "an exception was caught at the ABI boundary";
    case ERR_INTERNAL: return "internal error";
    case ERR_LABELS: return "invalid index labels";
    case ERR_UNSUPPORTED: return "operation not supported by this implementation";
    default: return "unknown error";
  }
}

}  // namespace tp
