// GPU planning: the index analysis of crates/tprims-contract/src/plan/analysis.rs
// (drop extent-1 axes, sort each role, fold compatible axes, build scatter
// vectors) plus the decisions that replace the CPU orientation, row-block and
// blocking rules (plan/orientation.rs:196,307; abi/types.rs:356):
//  - which axis each operand's gathers run along (coalescing in HBM),
//  - the K order, blocked so that one K slab reads whole cache lines of both
//    operands when their contiguous axes are different K axes,
//  - the workgroup tile.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>
#include <memory>

#include "kernels.h"

namespace tp {

std::vector<int64_t> build_scatter(const std::vector<int64_t>& extents, const std::vector<int64_t>& strides) {
  int64_t total = 1;
  for (auto e : extents) total *= e;
  std::vector<int64_t> out;
  if (total <= 0) return out;
  out.reserve((size_t)total);
  const size_t n = extents.size();
  std::vector<int64_t> counter(n, 0);
  int64_t offset = 0;
  for (int64_t t = 0; t < total; t++) {
    out.push_back(offset);
    for (size_t d = 0; d < n; d++) {
      counter[d]++;
      offset += strides[d];
      if (counter[d] < extents[d]) break;
      offset -= strides[d] * extents[d];
      counter[d] = 0;
    }
  }
  return out;
}

std::vector<Axis> fold_axes(const std::vector<Axis>& axes) {
  std::vector<Axis> out;
  for (const Axis& ax : axes) {
    if (!out.empty()) {
      Axis& p = out.back();
      if (ax.sa == p.sa * p.extent && ax.sb == p.sb * p.extent && ax.sc == p.sc * p.extent &&
          ax.sd == p.sd * p.extent) {
        p.extent *= ax.extent;
        continue;
      }
    }
    out.push_back(ax);
  }
  return out;
}

namespace {

uint64_t uabs(int64_t x) { return x < 0 ? (uint64_t)(-(x + 1)) + 1 : (uint64_t)x; }

// analysis.rs:197-207: extent-1 axes dropped, stable sort by key, fold.
template <class Key>
std::vector<Axis> role_axes(const std::vector<RoleAxis>& role, Key key) {
  std::vector<Axis> v;
  for (auto& r : role)
    if (r.extent != 1) v.push_back({r.extent, r.s[OP_A], r.s[OP_B], r.s[OP_C], r.s[OP_D]});
  std::stable_sort(v.begin(), v.end(), [&](const Axis& x, const Axis& y) { return key(x) < key(y); });
  return fold_axes(v);
}

int64_t product(const std::vector<Axis>& axes) {
  __int128 p = 1;
  for (auto& a : axes) {
    p *= a.extent;
    if (p * 8 > (__int128)LLONG_MAX) fail(ERR_SHAPE, "overflow: role extent product");
  }
  return (int64_t)p;
}

template <class Pick>
std::vector<int64_t> scatter_for(const std::vector<Axis>& axes, Pick pick) {
  std::vector<int64_t> e, s;
  for (auto& a : axes) {
    e.push_back(a.extent);
    s.push_back(pick(a));
  }
  return build_scatter(e, s);
}

// Coalescing cost of a gather direction: the element stride between lanes,
// penalised when the run along it is too short to fill a cache line, and
// prohibitive for a zero stride (a broadcast gives no contiguity).
double cost(int64_t stride, int64_t run) {
  if (stride == 0) return 1e18;
  double s = (double)uabs(stride);
  return run >= 4 ? s : s * 8.0;
}

int64_t largest_divisor_le(int64_t n, int64_t cap) {
  if (cap < 1) cap = 1;
  for (int64_t d = std::min(n, cap); d >= 1; d--)
    if (n % d == 0) return d;
  return 1;
}

void push_nonunit(std::vector<Axis>& v, const Axis& a) {
  if (a.extent > 1) v.push_back(a);
}

Axis scaled(const Axis& a, int64_t extent, int64_t by) {
  return {extent, a.sa * by, a.sb * by, a.sc * by, a.sd * by};
}

// Where A's and D's contiguous axes are different M axes, order M as D's axis
// in blocks of up to 8, then A's in blocks of up to 8, then the rest of both,
// then the other axes: a tile's rows cover 8 x 8 blocks that write runs of 8
// elements of D and read runs of 8 of A, where either axis alone leading
// leaves the other operand one element per cache line.
void block_rows(std::vector<Axis>& m) {
  int ia = -1, id = -1;
  for (size_t i = 0; i < m.size(); i++) {
    if (uabs(m[i].sa) == 1 && ia < 0) ia = (int)i;
    if (uabs(m[i].sd) == 1 && id < 0) id = (int)i;
  }
  if (ia < 0 || id < 0 || ia == id) return;
  const Axis xa = m[ia], xd = m[id];
  const int64_t dLo = largest_divisor_le(xd.extent, 8), aLo = largest_divisor_le(xa.extent, 8);
  if (dLo < 2 || aLo < 2) return;
  std::vector<Axis> out;
  push_nonunit(out, scaled(xd, dLo, 1));
  push_nonunit(out, scaled(xa, aLo, 1));
  push_nonunit(out, scaled(xd, xd.extent / dLo, dLo));
  push_nonunit(out, scaled(xa, xa.extent / aLo, aLo));
  for (size_t i = 0; i < m.size(); i++)
    if ((int)i != ia && (int)i != id) out.push_back(m[i]);
  m = out;
}

}  // namespace

bool tile_allowed(int id, DType t, int prec, int method) {
  if (prec == PREC_FP8) return id == (is_complex(t) ? STILE_128x64 : STILE_256x256);
  if (prec != PREC_STORAGE)
    return id == STILE_64x64 || (id == STILE_128x64 && is_complex(t)) ||
           ((id == STILE_128x128 || id == STILE_256x128) && !is_complex(t));
  if (id >= 10) return false;
  // c64 keeps the 64x64 tile and gets the 128x128 DMA tile instead of
  // 128x64; 128x128 (DMA only) exists for every dtype and complex schedule.
  if (id == TILE_128x64 && t == DType::C64) return false;
  return tile_cfg(id) != nullptr;
}

namespace {

int choose_tile(const Plan& p) {
  const int prec = p.cfg.precision;
  if (p.cfg.tile != 0) {
    if (!tile_allowed(p.cfg.tile, p.prob.dtype, prec, p.cfg.complex_method)) fail(ERR_UNSUPPORTED, "tile id not available for this plan");
    return p.cfg.tile;
  }
  std::vector<std::pair<int, double>> cands;
  if (prec == PREC_FP8) {
    return is_complex(p.prob.dtype) ? STILE_128x64 : STILE_256x256;
  } else if (prec != PREC_STORAGE) {
    cands = {{STILE_256x128, 1.3}, {STILE_128x128, 1.15}, {STILE_128x64, 1.2}, {STILE_64x64, 1.0}};
  } else {
    cands = {{TILE_128x128, 1.3}, {TILE_128x64, 1.1}, {TILE_64x64, 1.0}, {TILE_32x32, 0.7}};
  }
  // Estimated time of the MFMA work on the busiest CU: its tiles (the tile
  // count spread over 256 CUs) times the tile area, over the SIMDs those
  // tiles' waves can occupy (4 per CU) and the tile's measured efficiency. A
  // big tile wins on big problems; on small ones it would leave CUs idle and
  // compute its out-of-range rows, and a smaller tile finishes first.
  int best = 0;
  double best_cost = 0;
  for (auto [id, eff] : cands) {
    if (!tile_allowed(id, p.prob.dtype, prec, p.cfg.complex_method)) continue;
    const TileCfg* t = tile_cfg(id);
    const double tiles = (double)((p.M + t->bm - 1) / t->bm) * (double)((p.N + t->bn - 1) / t->bn) * (double)p.H;
    const double per_cu = std::ceil(tiles / 256.0);
    const double simds = std::min(4.0, per_cu * (t->threads / 64));
    const double cost = per_cu * t->bm * t->bn / (simds * eff);
    if (best == 0 || cost < best_cost * (1 - 1e-9)) {
      best_cost = cost;
      best = id;
    }
  }
  return best;
}

// Where the MFMA estimate is not what bounds the time, the 64 x 64 tile (four
// to eight waves per workgroup, four workgroups per CU) keeps more gathers in
// flight than one 128 x 128 (or 32 x 32) workgroup. Measured over the TCCG
// corpus at 64 MiB per tensor, every dtype, tiles forced one by one on one
// GPU, and on contiguous 6144 x 6144 products; each rule applies with at
// least 1024 tiles of 64 x 64:
//  - K <= 32 with operands not in 16-byte runs: a tile is little more than
//    its gathers and its write-back (CCSD(T) abcijk-ikmc-mjab, K = 24, f64
//    0.189 -> 0.147 ms; InTensLi ajbdc-ckbad-jk c64 0.275 -> 0.176). In
//    16-byte runs the 128 x 128 tile stays (f64 6144 x 6144 x 24: 0.094
//    against 0.139 on 64 x 64);
//  - K <= 128 instead of 128 x 64 or 128 x 128, where the operands are not
//    in 16-byte runs or M or N is at most 128, except f64 with A running
//    along K (TCCG ajbc-ckba-jk, N = 72, K = 52, f32 0.212 -> 0.135; f64
//    abjc-kbac-jk 0.120 on 128 x 128 against 0.149);
//  - complex instead of 128 x 128 unless A and B both run along K (c32
//    ijkl-imjn-nlmk 2.85 -> 2.39 ms, c64 5.49 -> 4.58; with both along K
//    c32 ijkl-minj-nlmk keeps 128 x 128: 4.29 against 5.21, and so does the
//    headline product, whose A is a stride-2 view: c32 100 against 74
//    TFLOP/s); c32 in 16-byte runs takes 64 x 64 in any orientation
//    (column-major 8192 x 8192 x 512 with A along K: 128 against 113
//    TFLOP/s; 4096^3: 135 against 124).
int gather_bound_tile(const Plan& p, int best) {
  if (p.cfg.precision != PREC_STORAGE || best == TILE_64x64 ||
      !tile_allowed(TILE_64x64, p.prob.dtype, PREC_STORAGE, p.cfg.complex_method))
    return best;
  const double tiles64 = (double)((p.M + 63) / 64) * (double)((p.N + 63) / 64) * (double)p.H;
  if (tiles64 < 1024) return best;
  // An operand runs along K where its smallest nonzero stride is a K axis's
  // (a strided view such as the headline's stride-2 A included).
  auto along_k = [&](int64_t Axis::*s, const std::vector<Axis>& rows) {
    auto least = [&](const std::vector<Axis>& ax) {
      uint64_t m = UINT64_MAX;
      for (const Axis& x : ax)
        if (x.*s != 0) m = std::min(m, uabs(x.*s));
      return m;
    };
    return least(p.k_ax) < least(rows);
  };
  const bool a_k = along_k(&Axis::sa, p.m_ax), b_k = along_k(&Axis::sb, p.n_ax);
  if (p.K <= 32 && !p.ab_vec && (best == TILE_32x32 || best == TILE_128x64 || best == TILE_128x128))
    return TILE_64x64;
  if (p.K <= 128 && (best == TILE_128x64 || best == TILE_128x128) && (!p.ab_vec || std::min(p.M, p.N) <= 128) &&
      !(p.prob.dtype == DType::F64 && a_k))
    return TILE_64x64;
  if (is_complex(p.prob.dtype) && best == TILE_128x128 &&
      (!(a_k && b_k) || (p.prob.dtype == DType::C32 && p.ab_vec)))
    return TILE_64x64;
  return best;
}

}  // namespace

void configure(Plan* plan, const PlanConfig& cfg_in) {
  Plan& p = *plan;
  const Problem& pr = p.prob;
  PlanConfig cfg = cfg_in;
  p.source = cfg.tile != 0 || cfg.split_k != 0 ? 1 : 0;
  if (p.strategy == 0 && cfg.tile == 0 && cfg.split_k == 0) {
    int t = 0, s = 0;
    if (tuning_lookup(tuning_key(p, cfg), &t, &s) && tile_allowed(t, pr.dtype, cfg.precision, cfg.complex_method)) {
      cfg.tile = t;
      cfg.split_k = s;
      p.source = 2;
    }
  }
  p.cfg = cfg;
  if (cfg.precision != PREC_STORAGE && is_double(pr.dtype))
    fail(ERR_UNSUPPORTED, "split-BF16 / FP8 precision applies to f32 and c32 only");
  if (cfg.complex_method < CPLX_NATIVE || cfg.complex_method > CPLX_3M)
    fail(ERR_INVALID_ARGUMENT, "unknown complex method");
  if (cfg.precision != PREC_STORAGE && cfg.complex_method != CPLX_NATIVE)
    fail(ERR_UNSUPPORTED, "split-BF16 and FP8 run the native complex schedule only");

  p.tile = choose_tile(p);
  const TileCfg* t = tile_cfg(p.tile);
  const int64_t BK = t->bk;
  const int64_t es = elem_size(pr.dtype);

  // Gather directions.
  const double inf = 1e30;
  double a_m = p.m_ax.empty() ? inf : cost(p.m_ax[0].sa, p.m_ax[0].extent);
  double b_n = p.n_ax.empty() ? inf : cost(p.n_ax[0].sb, p.n_ax[0].extent);
  int kA = -1, kB = -1;
  double a_k = inf, b_k = inf;
  for (size_t i = 0; i < p.k_ax.size(); i++) {
    double ca = cost(p.k_ax[i].sa, p.k_ax[i].extent), cb = cost(p.k_ax[i].sb, p.k_ax[i].extent);
    if (ca < a_k) {
      a_k = ca;
      kA = (int)i;
    }
    if (cb < b_k) {
      b_k = cb;
      kB = (int)i;
    }
  }
  p.a_kmajor = a_k <= a_m;
  p.b_kmajor = b_k <= b_n;
  if (!p.a_kmajor) kA = -1;
  if (!p.b_kmajor) kB = -1;
  double d_m = p.m_ax.empty() ? inf : cost(p.m_ax[0].sd, p.m_ax[0].extent);
  double d_n = p.n_ax.empty() ? inf : cost(p.n_ax[0].sd, p.n_ax[0].extent);
  p.d_mmajor = d_m <= d_n;

  // K order.
  std::vector<Axis> korder;
  if (kA >= 0 && kB >= 0 && kA != kB) {
    const Axis &xa = p.k_ax[kA], &xb = p.k_ax[kB];
    int64_t lineA = std::max<int64_t>(1, 128 / std::max<int64_t>(1, (int64_t)uabs(xa.sa) * es));
    int64_t lineB = std::max<int64_t>(1, 128 / std::max<int64_t>(1, (int64_t)uabs(xb.sb) * es));
    // One slab of BK k holds aLo consecutive indices of A's contiguous K axis
    // times bLo of B's: A reads runs of aLo elements, B runs of bLo. Of the
    // divisor pairs with aLo * bLo <= BK, take the one whose worse operand
    // uses the larger share of its cache line, then the fuller slab (extents
    // 72 x 72 in f64: 4 x 4, where taking A's longest run first gives 12 x 1
    // and one line per element of B).
    int64_t aLo = 1, bLo = 1;
    double best_share = -1.0;
    for (int64_t da = 1; da <= std::min(BK, xa.extent); da++) {
      if (xa.extent % da != 0) continue;
      for (int64_t db = 1; da * db <= BK && db <= xb.extent; db++) {
        if (xb.extent % db != 0) continue;
        const double share = std::min((double)std::min(da, lineA) / lineA, (double)std::min(db, lineB) / lineB);
        if (share > best_share + 1e-12 || (share > best_share - 1e-12 && da * db > aLo * bLo)) {
          best_share = share;
          aLo = da;
          bLo = db;
        }
      }
    }
    int64_t bMid = largest_divisor_le(xb.extent / bLo, std::max<int64_t>(1, lineB / bLo));
    push_nonunit(korder, scaled(xa, aLo, 1));
    push_nonunit(korder, scaled(xb, bLo, 1));
    push_nonunit(korder, scaled(xb, bMid, bLo));
    push_nonunit(korder, scaled(xa, xa.extent / aLo, aLo));
    push_nonunit(korder, scaled(xb, xb.extent / (bLo * bMid), bLo * bMid));
    for (size_t i = 0; i < p.k_ax.size(); i++)
      if ((int)i != kA && (int)i != kB) korder.push_back(p.k_ax[i]);
  } else {
    int first = kA >= 0 ? kA : kB;
    if (first >= 0) korder.push_back(p.k_ax[first]);
    for (size_t i = 0; i < p.k_ax.size(); i++)
      if ((int)i != first) korder.push_back(p.k_ax[i]);
  }

  p.am = scatter_for(p.m_ax, [](const Axis& a) { return a.sa; });
  p.cm = scatter_for(p.m_ax, [](const Axis& a) { return a.sc; });
  p.dm = scatter_for(p.m_ax, [](const Axis& a) { return a.sd; });
  p.bn = scatter_for(p.n_ax, [](const Axis& a) { return a.sb; });
  p.cn = scatter_for(p.n_ax, [](const Axis& a) { return a.sc; });
  p.dn = scatter_for(p.n_ax, [](const Axis& a) { return a.sd; });
  p.ak = scatter_for(korder, [](const Axis& a) { return a.sa; });
  p.bk = scatter_for(korder, [](const Axis& a) { return a.sb; });
  p.ha = scatter_for(p.h_ax, [](const Axis& a) { return a.sa; });
  p.hb = scatter_for(p.h_ax, [](const Axis& a) { return a.sb; });
  p.hc = scatter_for(p.h_ax, [](const Axis& a) { return a.sc; });
  p.hd = scatter_for(p.h_ax, [](const Axis& a) { return a.sd; });
  p.dr = scatter_for(p.r_ax, [](const Axis& a) { return a.sd; });
  // Empty roles would give empty vectors; keep one zero entry so device
  // pointers stay valid (the kernels never index past M, N, K, H).
  for (auto* v : {&p.am, &p.cm, &p.dm, &p.bn, &p.cn, &p.dn, &p.ak, &p.bk, &p.ha, &p.hb, &p.hc, &p.hd, &p.dr})
    if (v->empty()) v->push_back(0);

  auto kvec = [&](const std::vector<int64_t>& kv, const std::vector<int64_t>& rv, const std::vector<int64_t>& hv) {
    if (p.K % 8 != 0) return 0;
    const int64_t q = 16 / es;  // elements per 16 bytes
    for (int64_t v : rv)
      if (v % q != 0) return 0;
    for (int64_t v : hv)
      if (v % q != 0) return 0;
    bool identity = true;
    for (int64_t k = 0; k < p.K; k++) {
      identity = identity && kv[k] == k;
      if (k % 8 == 0 ? kv[k] % q != 0 : kv[k] != kv[k - 1] + 1) return 0;
    }
    return identity ? 2 : 1;
  };
  p.a_kvec = kvec(p.ak, p.am, p.ha);
  p.b_kvec = kvec(p.bk, p.bn, p.hb);

  auto fits32 = [&](const std::vector<int64_t>& rv, const std::vector<int64_t>& kv) {
    const auto [rlo, rhi] = std::minmax_element(rv.begin(), rv.end());
    const auto [klo, khi] = std::minmax_element(kv.begin(), kv.end());
    return *rlo >= 0 && *klo >= 0 && (__int128)(*rhi + *khi + 1) * es <= (__int128)(1 << 30);
  };
  p.off32 = fits32(p.am, p.ak) && fits32(p.bn, p.bk);
  // The 128x128 tile runs on the DMA kernel only, which needs 32-bit offsets
  // (the K order is the same for both tiles: BK 16).
  if (p.tile == TILE_128x128 && !p.off32 && cfg.tile == 0) p.tile = TILE_64x64;

  // Split K when fewer than 128 tiles would run (half of the 256 CUs): as many
  // splits as keep every tile's splits in one round over the 256 CUs, each
  // split at least 6 slabs, at most 16 (f64 576^3, 81 tiles: 3 splits 25.4 us
  // per call, 4 splits 30.5 (a second round on 68 CUs), 6 splits 27.6).
  // cfg.split_k = 1 keeps one K range, n > 1 asks for n splits (at most one
  // per slab). For real dtypes with K <= 1024, where 32x32 tiles alone would
  // give at least 192 workgroups, the K-group 32x32 tile replaces split K
  // (chosen below, once the gather form is known): f64 24^6 20.1 against 21.0
  // us per call, f32 640 x 1024 x 640 20.5 against 23.3; with longer K the
  // 64x64 tile's better MFMA rate per slab wins (f64 520 x 2048 x 520: 45.6
  // split against 68), and so does split K for c32 (24^6: 29.6 against 31.2).
  p.ksplit = 1;
  p.kslabs = 0;
  auto tiles_of = [&](int id) {
    const TileCfg* tc = tile_cfg(id);
    return ((p.M + tc->bm - 1) / tc->bm) * ((p.N + tc->bn - 1) / tc->bn) * p.H;
  };
  // Split K would run and the K-group tiles may replace it: K-group (LDS-DMA)
  // where 32x32 tiles give at least 192 workgroups, register-direct (f64 in
  // 16-byte runs) at any count (202 x 300 x 130: 6.9 against 12.5 us split
  // three ways; three 96 x 800 x 96 entries 11.9 against 14.1).
  bool kgroup_candidate = false, direct_candidate = false;
  if (cfg.split_k < 0 || cfg.split_k > 64) fail(ERR_INVALID_ARGUMENT, "split_k must be in [0, 64]");
  if (p.strategy == 0 && cfg.precision == PREC_STORAGE && cfg.split_k != 1) {
    const TileCfg* tc = tile_cfg(p.tile);
    const int64_t tiles = tiles_of(p.tile);
    const int64_t nk = (p.K + tc->bk - 1) / tc->bk;
    int64_t s = 1;
    if (cfg.split_k > 1) s = std::min<int64_t>(cfg.split_k, nk);
    else if (tiles < 128 && nk >= 16 && p.tile != TILE_32x32K && p.tile != TILE_32x32R && p.tile != TILE_32x48R)
      s = std::min<int64_t>(std::min<int64_t>(256 / tiles, nk / 6), 16);
    if (s >= 2) {
      p.ksplit = (int)s;
      p.kslabs = (nk + s - 1) / s;
      direct_candidate = cfg.split_k == 0 && cfg.tile == 0 && !is_complex(pr.dtype) && p.K <= 1024;
      kgroup_candidate = direct_candidate && tiles_of(TILE_32x32K) >= 192;
    }
  }
  p.ak32.clear();
  p.bk32.clear();
  if (p.off32) {
    const size_t padded = (size_t)((p.K + 63) / 64 * 64);
    p.ak32.assign(padded, 0x40000000u);
    p.bk32.assign(padded, 0x40000000u);
    for (int64_t k = 0; k < p.K; k++) {
      p.ak32[k] = (uint32_t)(p.ak[k] * es);
      p.bk32[k] = (uint32_t)(p.bk[k] * es);
    }
  }
  // 16-byte runs (ContractArgs::ab_vec): q = 16 / es consecutive elements
  // along each operand's major index, every offset a multiple of q. In block
  // scatter terms: the major scatter is regular with stride 1 at block size
  // q, with no ragged last block.
  p.ab_vec = false;
  p.a_vec16 = false;
  if (p.off32 && es < 16) {  // f32, f64, c32 (a c64 element is already one 16-byte DMA)
    const int64_t q = 16 / es;
    auto runs = [&](const std::vector<int64_t>& v, int64_t n) {
      if (n % q != 0) return false;
      for (int64_t s : build_block_scatter(v.data(), n, q))
        if (s != 1) return false;
      return true;
    };
    auto mult = [&](const std::vector<int64_t>& v, int64_t n) {
      for (int64_t i = 0; i < n; i++)
        if (v[i] % q != 0) return false;
      return true;
    };
    auto operand = [&](bool kmajor, const std::vector<int64_t>& rv, int64_t rn, const std::vector<int64_t>& kv,
                       const std::vector<int64_t>& hv) {
      if (!mult(hv, p.H)) return false;
      if (kmajor) {
        if (!runs(kv, p.K) || !mult(rv, rn)) return false;
        for (int64_t k = 0; k < p.K; k += q)
          if (kv[k] % q != 0) return false;
        return true;
      }
      if (!runs(rv, rn) || !mult(kv, p.K)) return false;
      for (int64_t i = 0; i < rn; i += q)
        if (rv[i] % q != 0) return false;
      return true;
    };
    p.a_vec16 = operand(p.a_kmajor, p.am, p.M, p.ak, p.ha);
    p.ab_vec = p.a_vec16 && operand(p.b_kmajor, p.bn, p.N, p.bk, p.hb);
  }
  if (cfg.tile == 0 && p.strategy == 0 && p.ksplit == 1) p.tile = gather_bound_tile(p, p.tile);
  // f64 in 16-byte runs takes the register-direct K-group tile (f64 24^6:
  // 16.4 against 20.6 us per call on the LDS-DMA one; 576 x 1152 x 576 27.8
  // against 34.2).
  const bool direct = direct_candidate && direct_supported(p.off32, p.ab_vec, pr.dtype);
  if (direct || (kgroup_candidate && kgroup_supported(p.off32, p.ab_vec, p.a_kmajor, p.b_kmajor, pr.dtype))) {
    p.tile = direct ? TILE_32x32R : TILE_32x32K;
    // More 32x32 tiles than CUs put two on some CUs: 32x48 tiles (B gathered
    // along K), one per CU, when they all fit in one round (f64 24^6: 324
    // tiles of 32x32, 216 of 32x48).
    if (p.tile == TILE_32x32R && p.b_kmajor && tiles_of(TILE_32x32R) > 256 && tiles_of(TILE_32x48R) <= 256)
      p.tile = TILE_32x48R;
    p.ksplit = 1;
    p.kslabs = 0;
  }
  // Split-BF16 that splits A inside the GEMM runs its 128x128 tile (two
  // workgroups per CU, whose splitting and MFMAs overlap) rather than
  // 256x128 (one workgroup per CU): f32 transform step, 3 terms, 135 -> 157
  // TFLOP/s. Both tiles have 32-k steps, so the K order above stands.
  if (cfg.tile == 0 && p.tile == STILE_256x128 && p.off32 &&
      (cfg.precision == PREC_SPLIT_BF16_2 || cfg.precision == PREC_SPLIT_BF16_3) &&
      split_a_in_gemm(false, p.a_kmajor, p.a_vec16, p.N, STILE_128x128))
    p.tile = STILE_128x128;
}

std::vector<int64_t> build_block_scatter(const int64_t* scat, int64_t len, int64_t blk) {
  std::vector<int64_t> out;
  if (len <= 0 || blk <= 0) return out;
  out.reserve((size_t)((len + blk - 1) / blk));
  for (int64_t lo = 0; lo < len; lo += blk) {
    const int64_t hi = std::min(lo + blk, len);
    int64_t stride = 0;
    if (hi - lo > 1) {
      stride = scat[lo + 1] - scat[lo];
      for (int64_t i = lo + 1; i < hi; i++)
        if (scat[i] - scat[i - 1] != stride) {
          stride = SCATTER_IRREGULAR;
          break;
        }
    }
    out.push_back(stride);
  }
  return out;
}

Plan* make_plan(const Problem& pr, const PlanConfig& cfg) {
  std::unique_ptr<Plan> p(new Plan);
  p->prob = pr;
  const auto& r = pr.roles;
  p->m_ax = role_axes(r.m, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sa)); });
  p->n_ax = role_axes(r.n, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sb)); });
  // The operand's own contiguous axis leads its role (A's for M, B's for N)
  // where D's contiguous axis is another one and the operand is at least a
  // quarter of D's size (4 K >= N for A, 4 K >= M for B): a tile's rows then
  // read whole runs of A (B), and the rows of D they write are completed in
  // L2 by the neighbouring tiles (TCCG ajbdc-ckbad-jk, all extents 24: A was
  // read one element per cache line, f64 0.185 -> 0.126 ms; ijkl-imjn-lnkm
  // f64 25.4 -> 55.4 TFLOP/s). Where D is much the larger tensor its writes
  // keep the lead (CCSD(T) abcijk-ikma-mjbc: 0.196 against 0.48 ms with A's
  // axis first).
  auto lead = [](std::vector<Axis>& ax, int64_t Axis::*s) {
    for (size_t i = 1; i < ax.size(); i++)
      if (uabs(ax[i].*s) == 1) {
        std::rotate(ax.begin(), ax.begin() + i, ax.begin() + i + 1);
        ax = fold_axes(ax);
        return;
      }
  };
  p->k_ax = role_axes(r.k, [](const Axis& x) { return std::make_pair(uabs(x.sa), uabs(x.sb)); });
  {
    const int64_t M = product(p->m_ax), N = product(p->n_ax), K = product(p->k_ax);
    // D written along one role leaves the other role's order free: there the
    // operand's contiguous axis leads whatever the sizes.
    auto has_unit = [](const std::vector<Axis>& ax, int64_t Axis::*s) {
      for (const Axis& x : ax)
        if (uabs(x.*s) == 1) return true;
      return false;
    };
    const bool d_along_n = has_unit(p->n_ax, &Axis::sd), d_along_m = has_unit(p->m_ax, &Axis::sd);
    if (4 * K >= N || d_along_n) lead(p->m_ax, &Axis::sa);
    if (4 * K >= M || d_along_m) lead(p->n_ax, &Axis::sb);
    // With a short K the reads of A and the writes of D are the work: block
    // the rows so that both run (TCCG InTensLi ajbdc-ckbad-jk, extents 24,
    // on the 64 x 64 tile: f64 0.104 -> 0.076 ms, f32 0.060 -> 0.047; abjc-
    // cbka-kj, K = 72, f32 0.112 -> 0.080).
    if (K <= 128) block_rows(p->m_ax);
  }
  p->h_ax = role_axes(r.h, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sa)); });
  p->r_ax = role_axes(r.r, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sc)); });
  p->M = product(p->m_ax);
  p->N = product(p->n_ax);
  p->K = product(p->k_ax);
  p->H = product(p->h_ax);
  p->R = product(p->r_ax);
  if (pr.out_empty) p->strategy = 2;
  else if (pr.k_empty) p->strategy = 1;
  else if (p->M * p->N <= 16 && p->K <= 256 && p->H * p->M * p->N >= 4096) p->strategy = 3;
  else p->strategy = 0;
  if (p->strategy == 2) {
    // Nothing is ever launched; keep the vectors trivial.
    p->cfg = cfg;
    p->tile = TILE_64x64;
    for (auto* v : {&p->am, &p->cm, &p->dm, &p->bn, &p->cn, &p->dn, &p->ak, &p->bk, &p->ha, &p->hb, &p->hc, &p->hd,
                    &p->dr})
      v->assign(1, 0);
    return p.release();
  }
  configure(p.get(), cfg);
  return p.release();
}

std::vector<int64_t> Plan::flat_tables() const {
  const std::vector<const std::vector<int64_t>*> parts = {&am, &ak, &bk, &bn, &cm, &cn, &dm,
                                                          &dn, &ha, &hb, &hc, &hd, &dr};
  size_t total = 0;
  for (auto* v : parts) total += v->size();
  // The 32-bit K offsets follow, two per 64-bit word (each array starts on a word).
  const size_t w32a = (ak32.size() + 1) / 2, w32b = (bk32.size() + 1) / 2;
  std::vector<int64_t> flat;
  flat.reserve(total + w32a + w32b);
  for (auto* v : parts) flat.insert(flat.end(), v->begin(), v->end());
  flat.resize(total + w32a + w32b, 0);
  if (!ak32.empty()) std::memcpy(flat.data() + total, ak32.data(), ak32.size() * 4);
  if (!bk32.empty()) std::memcpy(flat.data() + total + w32a, bk32.data(), bk32.size() * 4);
  return flat;
}

DeviceScatter Plan::tables_at(int64_t* base) const {
  DeviceScatter ds{};
  ds.base = base;
  const std::vector<const std::vector<int64_t>*> parts = {&am, &ak, &bk, &bn, &cm, &cn, &dm,
                                                          &dn, &ha, &hb, &hc, &hd, &dr};
  const int64_t* q = base;
  const int64_t** slots[] = {&ds.am, &ds.ak, &ds.bk, &ds.bn, &ds.cm, &ds.cn, &ds.dm,
                             &ds.dn, &ds.ha, &ds.hb, &ds.hc, &ds.hd, &ds.dr};
  for (size_t i = 0; i < parts.size(); i++) {
    *slots[i] = q;
    q += parts[i]->size();
  }
  const size_t w32a = (ak32.size() + 1) / 2;
  ds.ak32 = ak32.empty() ? nullptr : (const uint32_t*)q;
  ds.bk32 = bk32.empty() ? nullptr : (const uint32_t*)(q + w32a);
  return ds;
}

const std::vector<int64_t>& Plan::host_tables() {
  std::lock_guard<std::mutex> g(mu);
  if (host_image.empty()) host_image = flat_tables();
  return host_image;
}

const DeviceScatter& Plan::on_device(int device) {
  std::lock_guard<std::mutex> g(mu);
  auto it = dev.find(device);
  if (it != dev.end()) return it->second;
  const std::vector<int64_t> flat = flat_tables();
  const size_t words = flat.size();
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != device) (void)hipSetDevice(device);
  int64_t* base = nullptr;
  hipError_t e = hipMalloc(&base, words * sizeof(int64_t));
  if (e == hipSuccess) e = hipMemcpy(base, flat.data(), words * sizeof(int64_t), hipMemcpyHostToDevice);
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    if (base) (void)hipFree(base);
    fail(ERR_INTERNAL, std::string("uploading scatter vectors: ") + hipGetErrorString(e));
  }
  return dev.emplace(device, tables_at(base)).first->second;
}

ContractArgs Plan::args(int device, const void* A, const void* B, const void* C, void* D, const double alpha[2],
                        const double beta[2], bool read_c) {
  return args_with(on_device(device), A, B, C, D, alpha, beta, read_c);
}

ContractArgs Plan::args_with(const DeviceScatter& s, const void* A, const void* B, const void* C, void* D,
                             const double alpha[2], const double beta[2], bool read_c) const {
  ContractArgs a{};
  a.A = A;
  a.B = B;
  a.C = C;
  a.D = D;
  a.am = s.am;
  a.ak = s.ak;
  a.bk = s.bk;
  a.bn = s.bn;
  a.cm = s.cm;
  a.cn = s.cn;
  a.dm = s.dm;
  a.dn = s.dn;
  a.ha = s.ha;
  a.hb = s.hb;
  a.hc = s.hc;
  a.hd = s.hd;
  a.dr = r_ax.empty() ? nullptr : s.dr;
  a.M = M;
  a.N = N;
  a.K = K;
  a.H = H;
  a.R = R;
  a.alpha[0] = alpha[0];
  a.alpha[1] = alpha[1];
  a.beta[0] = beta[0];
  a.beta[1] = beta[1];
  a.conj_a = prob.a.conj && is_complex(prob.dtype);
  a.conj_b = prob.b.conj && is_complex(prob.dtype);
  a.conj_c = prob.c.conj && is_complex(prob.dtype);
  a.conj_d = prob.d.conj && is_complex(prob.dtype);
  a.a_kmajor = a_kmajor;
  a.b_kmajor = b_kmajor;
  a.d_mmajor = d_mmajor;
  a.read_c = read_c;
  a.a_kvec = (uintptr_t)A % 16 == 0 ? a_kvec : 0;
  a.b_kvec = (uintptr_t)B % 16 == 0 ? b_kvec : 0;
  a.off32 = off32;
  a.ab_vec = ab_vec && (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0;
  a.a_vec16 = a_vec16 && (uintptr_t)A % 16 == 0;
  a.ak32 = s.ak32;
  a.bk32 = s.bk32;
  a.mn_affine = m_ax.size() <= 1 && n_ax.size() <= 1;
  if (a.mn_affine) {
    a.cd_stride[0] = m_ax.empty() ? 0 : m_ax[0].sc;
    a.cd_stride[1] = n_ax.empty() ? 0 : n_ax[0].sc;
    a.cd_stride[2] = m_ax.empty() ? 0 : m_ax[0].sd;
    a.cd_stride[3] = n_ax.empty() ? 0 : n_ax[0].sd;
  }
  a.h_affine = h_ax.size() == 1;
  if (a.h_affine) {
    a.h_stride[0] = h_ax[0].sa;
    a.h_stride[1] = h_ax[0].sb;
    a.h_stride[2] = h_ax[0].sc;
    a.h_stride[3] = h_ax[0].sd;
  }
  return a;
}

std::string Plan::kernel_name() const {
  if (strategy == 2) return "none";
  if (strategy == 1) return std::string("output_pass.") + dtype_name(prob.dtype);
  if (strategy == 3) return std::string("elementwise.") + dtype_name(prob.dtype);
  std::string s = dtype_name(prob.dtype);
  if (cfg.precision == PREC_SPLIT_BF16_2) s += ".split_bf16x2";
  else if (cfg.precision == PREC_SPLIT_BF16_3) s += ".split_bf16x3";
  else if (cfg.precision == PREC_FP8) s += ".fp8_e4m3";
  else s += is_double(prob.dtype) ? ".mfma_f64" : ".mfma_f32";
  if (is_complex(prob.dtype)) {
    static const char* m[3] = {".native", ".1m", ".3m"};
    s += m[cfg.complex_method];
  }
  // The DMA-only tile runs as 64x64 on the register-staged kernel without 32-bit offsets.
  const bool fallback = tile == TILE_128x128 && !dma_supported(off32, prob.dtype, cfg.complex_method, tile);
  // The K-group tile runs without split K where its gathers fit, else as the
  // register-staged 32x32 tile.
  // The register-direct K-group tile runs as the LDS-DMA K-group tile where
  // its loads do not fit, and both as the register-staged 32x32 tile where
  // neither does.
  const bool rtile = tile == TILE_32x32R || tile == TILE_32x48R;
  const bool direct = rtile && ksplit <= 1 && direct_supported(off32, ab_vec, prob.dtype);
  const bool kgroup = (tile == TILE_32x32K || rtile) && !direct && ksplit <= 1 &&
                      kgroup_supported(off32, ab_vec, a_kmajor, b_kmajor, prob.dtype);
  // 32x48 runs as 32x32 with B gathered along N.
  const int shown = fallback                                  ? TILE_64x64
                    : rtile && !direct && !kgroup             ? TILE_32x32
                    : rtile && (!direct || !b_kmajor)         ? TILE_32x32K
                                                              : tile;
  const TileCfg* t = tile_cfg(shown);
  if (t) s += "." + std::to_string(t->bm) + "x" + std::to_string(t->bn) + "x" + std::to_string(cfg.precision == PREC_FP8 ? 128 : t->bk);
  if (cfg.precision == PREC_STORAGE &&
      ((!fallback && dma_supported(off32, prob.dtype, cfg.complex_method, tile)) || kgroup))
    s += ab_vec ? ".dma16" : ".dma";
  if (kgroup) s += ".kgroup4";
  if (direct) s += ".regs16.kgroup8";
  if (ksplit > 1) s += ".splitk" + std::to_string(ksplit);
  // Split-BF16 with an f32 A gathered along M in 16-byte runs: A is split
  // inside the GEMM (for a 16-byte aligned A, checked per call).
  if ((cfg.precision == PREC_SPLIT_BF16_2 || cfg.precision == PREC_SPLIT_BF16_3) && off32 &&
      split_a_in_gemm(is_complex(prob.dtype), a_kmajor, a_vec16, N, tile))
    s += ".a_in_gemm";
  return s;
}

Plan::~Plan() {
  delete adj[0];
  delete adj[1];
  for (auto& [d, s] : dev) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (prev != d) (void)hipSetDevice(d);
    (void)hipFree(s.base);
    if (prev != d) (void)hipSetDevice(prev);
  }
}

}  // namespace tp
