// CPU restatement of the reference's packed (TBLIS-style) driver: TEST AND
// BASELINE INFRASTRUCTURE ONLY. It is the "port" timed as the CPU baseline
// beside the GPU figure, and never part of the product path.
//
// Follows, in order:
//   lowering          crates/tprims-contract/src/api/validate.rs:207-308
//   role axes / fold  crates/tprims-contract/src/plan/analysis.rs:197-321
//   scatter vectors   crates/tprims-kernel/src/pack/scatter.rs:70-102
//   blocking          crates/tprims-kernel/src/abi/types.rs:356-389
//                     (kc 384 f32 / 256 f64; A block 512 KiB; B panel 3 MiB)
//   five-loop nest    crates/tprims-contract/src/driver/static_grid.rs:97-218
//                     (loops 5 NC, 4 KC with a shared B panel, 3 MC with A
//                      packed per block, 2/1 register tiles)
//   packing           crates/tprims-kernel/src/pack/pack.rs:152-230
//                     (planar re/im slivers, conj folded, zero-filled edges)
//   write-back        crates/tprims-kernel/src/pack/writeback.rs:114-300 and
//                     driver/tile.rs:312-351 (first K block: alpha*AB + beta*op_C(C);
//                     later blocks re-read D with beta 1 and op_C = op_D)
// The register tile is a portable MR x NR accumulator block the compiler
// vectorises (-O3 -march=native), in place of the reference's AVX-512
// assembly-level families. Threads: OpenMP over B slivers, A slivers and tiles.
#include <omp.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Axis {
  int64_t e, sa, sb, sc, sd;
};

uint64_t uabs(int64_t x) { return x < 0 ? (uint64_t)(-x) : (uint64_t)x; }

std::vector<Axis> fold(std::vector<Axis> v) {
  std::vector<Axis> out;
  for (auto& a : v) {
    if (!out.empty()) {
      Axis& p = out.back();
      if (a.sa == p.sa * p.e && a.sb == p.sb * p.e && a.sc == p.sc * p.e && a.sd == p.sd * p.e) {
        p.e *= a.e;
        continue;
      }
    }
    out.push_back(a);
  }
  return out;
}

template <class K>
std::vector<Axis> role(std::vector<Axis> v, K key) {
  std::vector<Axis> w;
  for (auto& a : v)
    if (a.e != 1) w.push_back(a);
  std::stable_sort(w.begin(), w.end(), [&](const Axis& x, const Axis& y) { return key(x) < key(y); });
  return fold(w);
}

template <class P>
std::vector<int64_t> scatter(const std::vector<Axis>& ax, P pick) {
  int64_t total = 1;
  for (auto& a : ax) total *= a.e;
  std::vector<int64_t> out((size_t)std::max<int64_t>(total, 0));
  std::vector<int64_t> ctr(ax.size(), 0);
  int64_t off = 0;
  for (int64_t t = 0; t < total; t++) {
    out[t] = off;
    for (size_t d = 0; d < ax.size(); d++) {
      ctr[d]++;
      off += pick(ax[d]);
      if (ctr[d] < ax[d].e) break;
      off -= pick(ax[d]) * ax[d].e;
      ctr[d] = 0;
    }
  }
  return out;
}

struct Desc {
  int n;
  const int64_t *dims, *strides, *labels;
};

// One sliver of a packed panel (crates/tprims-kernel/src/pack/pack.rs:57-230,
// PackFormat::Real and ::Planar): rows v0 .. v0 + vr of the row scatter (zero
// past vlen) at every k of the K scatter, k-major: real dst[p vr + r]; planar
// the vr real parts, then the vr imaginary parts of step p at dst[2 p vr].
// conj negates the imaginary parts.
template <typename R, bool CPLX>
void pack_sliver(R* dst, const R* src, int64_t base, const int64_t* vscat, int64_t v0, int64_t vlen, int vr,
                 const int64_t* kscat, int64_t kc, bool conj) {
  constexpr int W = CPLX ? 2 : 1;
  for (int64_t p = 0; p < kc; p++) {
    R* out = dst + p * W * vr;
    for (int r = 0; r < vr; r++) {
      const int64_t v = v0 + r;
      R re = 0, im = 0;
      if (v < vlen) {
        const int64_t off = base + vscat[v] + kscat[p];
        if (CPLX) {
          re = src[2 * off];
          im = conj ? -src[2 * off + 1] : src[2 * off + 1];
        } else {
          re = src[off];
        }
      }
      out[r] = re;
      if (CPLX) out[vr + r] = im;
    }
  }
}

template <typename R, bool CPLX, int MR, int NR>
void run(const Desc& da, const Desc& db, const Desc& dc, const Desc& dd, const R* A, const R* B, const R* C, R* D,
         const double* alpha, const double* beta, int ca, int cb, int cc, int cd, int threads) {
  // Lowering: collapse repeated labels, classify roles.
  struct L {
    int64_t id, e, s[4];
    bool in[4];
  };
  std::vector<L> tab;
  const Desc* ops[4] = {&da, &db, &dc, &dd};
  for (int o = 0; o < 4; o++) {
    if (o == 2 && !C) continue;
    for (int k = 0; k < ops[o]->n; k++) {
      int64_t id = ops[o]->labels[k];
      auto it = std::find_if(tab.begin(), tab.end(), [&](const L& l) { return l.id == id; });
      if (it == tab.end()) {
        L l{id, ops[o]->dims[k], {0, 0, 0, 0}, {false, false, false, false}};
        tab.push_back(l);
        it = tab.end() - 1;
      }
      it->s[o] += ops[o]->strides[k];
      it->in[o] = true;
    }
  }
  std::vector<Axis> m, n, k, h;
  for (auto& l : tab) {
    Axis a{l.e, l.s[0], l.s[1], C ? l.s[2] : l.s[3], l.s[3]};
    if (l.in[0] && l.in[1] && l.in[3]) h.push_back(a);
    else if (l.in[0] && l.in[3]) m.push_back(a);
    else if (l.in[1] && l.in[3]) n.push_back(a);
    else k.push_back(a);
  }
  m = role(m, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sa)); });
  n = role(n, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sb)); });
  k = role(k, [](const Axis& x) { return std::make_pair(uabs(x.sa), uabs(x.sb)); });
  h = role(h, [](const Axis& x) { return std::make_pair(uabs(x.sd), uabs(x.sa)); });
  auto am = scatter(m, [](const Axis& a) { return a.sa; });
  auto cm = scatter(m, [](const Axis& a) { return a.sc; });
  auto dm = scatter(m, [](const Axis& a) { return a.sd; });
  auto bn = scatter(n, [](const Axis& a) { return a.sb; });
  auto cn = scatter(n, [](const Axis& a) { return a.sc; });
  auto dn = scatter(n, [](const Axis& a) { return a.sd; });
  auto ak = scatter(k, [](const Axis& a) { return a.sa; });
  auto bk = scatter(k, [](const Axis& a) { return a.sb; });
  auto ha = scatter(h, [](const Axis& a) { return a.sa; });
  auto hb = scatter(h, [](const Axis& a) { return a.sb; });
  auto hc = scatter(h, [](const Axis& a) { return a.sc; });
  auto hd = scatter(h, [](const Axis& a) { return a.sd; });
  const int64_t M = am.size(), N = bn.size(), K = ak.size(), H = hd.size();
  if (M == 0 || N == 0 || H == 0) return;

  constexpr int W = CPLX ? 2 : 1;  // reals per element
  const int64_t KC = sizeof(R) == 4 ? 384 : 256;
  int64_t MC = (512 * 1024) / (KC * W * (int64_t)sizeof(R));
  MC = std::max<int64_t>(MR, MC / MR * MR);
  int64_t NC = (3 * 1024 * 1024) / (KC * W * (int64_t)sizeof(R));
  NC = std::max<int64_t>(NR, NC / NR * NR);
  const double ar = alpha[0], ai = CPLX ? alpha[1] : 0, br = beta[0], bi = CPLX ? beta[1] : 0;
  const bool read_c = C && (br != 0 || bi != 0);

  std::vector<R> bpack((size_t)(NC * KC * W + 64));
  std::vector<R> ashared((size_t)(MC * KC * W + 64));

  auto load = [](const R* p, int64_t off, bool conj, R& re, R& im) {
    if (CPLX) {
      re = p[2 * off];
      im = conj ? -p[2 * off + 1] : p[2 * off + 1];
    } else {
      re = p[off];
      im = 0;
    }
  };

  // Empty K: the output pass D = op_D(beta op_C(C)), C never read when beta
  // is 0 (crates/tprims-kernel/src/pack/writeback.rs:495 scale_only).
  if (K == 0) {
    for (int64_t hh = 0; hh < H; hh++)
      for (int64_t j = 0; j < N; j++)
        for (int64_t i = 0; i < M; i++) {
          double vr = 0, vi = 0;
          if (read_c) {
            R c_r, c_i;
            load(C, hc[hh] + cm[i] + cn[j], cc, c_r, c_i);
            vr = br * c_r - bi * c_i;
            vi = br * c_i + bi * c_r;
          }
          if (cd) vi = -vi;
          const int64_t doff = hd[hh] + dm[i] + dn[j];
          if (CPLX) {
            D[2 * doff] = (R)vr;
            D[2 * doff + 1] = (R)vi;
          } else {
            D[doff] = (R)vr;
          }
        }
    return;
  }

#pragma omp parallel num_threads(threads)
  for (int64_t hh = 0; hh < H; hh++) {
    for (int64_t jc = 0; jc < N; jc += NC) {
      const int64_t nc = std::min(NC, N - jc);
      const int64_t nsl = (nc + NR - 1) / NR;
      for (int64_t pc = 0; pc < K; pc += KC) {
        const int64_t kc = std::min(KC, K - pc);
        const bool first = pc == 0;
        // Pack B: slivers of NR columns.
#pragma omp for schedule(static)
        for (int64_t sl = 0; sl < nsl; sl++)
          pack_sliver<R, CPLX>(bpack.data() + sl * kc * NR * W, B, hb[hh], bn.data(), jc + sl * NR, N, NR,
                               bk.data() + pc, kc, cb);
        for (int64_t ic = 0; ic < M; ic += MC) {
          const int64_t mc = std::min(MC, M - ic);
          const int64_t msl = (mc + MR - 1) / MR;
#pragma omp for schedule(static)
          for (int64_t sl = 0; sl < msl; sl++)
            pack_sliver<R, CPLX>(ashared.data() + sl * kc * MR * W, A, ha[hh], am.data(), ic + sl * MR, M, MR,
                                 ak.data() + pc, kc, ca);
#pragma omp for schedule(static) collapse(2)
          for (int64_t js = 0; js < nsl; js++)
            for (int64_t is = 0; is < msl; is++) {
              const R* Ap = ashared.data() + is * kc * MR * W;
              const R* Bp = bpack.data() + js * kc * NR * W;
              R cr[MR][NR], ci[MR][NR];
              for (int ii = 0; ii < MR; ii++)
                for (int jj = 0; jj < NR; jj++) cr[ii][jj] = ci[ii][jj] = 0;
              for (int64_t p = 0; p < kc; p++) {
                const R* a = Ap + p * MR * W;
                const R* b = Bp + p * NR * W;
                if (CPLX) {
                  const R* aim = a + MR;
                  const R* bim = b + NR;
                  for (int jj = 0; jj < NR; jj++)
                    for (int ii = 0; ii < MR; ii++) {
                      cr[ii][jj] += a[ii] * b[jj] - aim[ii] * bim[jj];
                      ci[ii][jj] += a[ii] * bim[jj] + aim[ii] * b[jj];
                    }
                } else {
                  for (int jj = 0; jj < NR; jj++)
                    for (int ii = 0; ii < MR; ii++) cr[ii][jj] += a[ii] * b[jj];
                }
              }
              for (int jj = 0; jj < NR; jj++) {
                int64_t j = jc + js * NR + jj;
                if (j >= N) break;
                for (int ii = 0; ii < MR; ii++) {
                  int64_t i = ic + is * MR + ii;
                  if (i >= M) break;
                  double xr = cr[ii][jj], xi = ci[ii][jj];
                  double vr = ar * xr - ai * xi, vi = ar * xi + ai * xr;
                  const int64_t doff = hd[hh] + dm[i] + dn[j];
                  if (first) {
                    if (read_c) {
                      R c_r, c_i;
                      load(C, hc[hh] + cm[i] + cn[j], cc, c_r, c_i);
                      vr += br * c_r - bi * c_i;
                      vi += br * c_i + bi * c_r;
                    }
                  } else {
                    R d_r, d_i;
                    load(D, doff, cd, d_r, d_i);
                    vr += d_r;
                    vi += d_i;
                  }
                  if (cd) vi = -vi;
                  if (CPLX) {
                    D[2 * doff] = (R)vr;
                    D[2 * doff + 1] = (R)vi;
                  } else {
                    D[doff] = (R)vr;
                  }
                }
              }
            }
        }
      }
    }
  }
}

}  // namespace

extern "C" int packed_contract(int dtype, const double* alpha, const void* A, int na, const int64_t* da,
                               const int64_t* sa, const int64_t* la, int ca, const void* B, int nb, const int64_t* db,
                               const int64_t* sb, const int64_t* lb, int cb, const double* beta, const void* C, int nc,
                               const int64_t* dc, const int64_t* sc, const int64_t* lc, int cc, void* D, int nd,
                               const int64_t* dd, const int64_t* sd, const int64_t* ld, int cd, int threads) {
  Desc xa{na, da, sa, la}, xb{nb, db, sb, lb}, xc{nc, dc, sc, lc}, xd{nd, dd, sd, ld};
  switch (dtype) {
    case 0: run<float, false, 16, 6>(xa, xb, xc, xd, (const float*)A, (const float*)B, (const float*)C, (float*)D,
                                     alpha, beta, 0, 0, 0, 0, threads); break;
    case 1: run<double, false, 8, 6>(xa, xb, xc, xd, (const double*)A, (const double*)B, (const double*)C,
                                     (double*)D, alpha, beta, 0, 0, 0, 0, threads); break;
    case 2: run<float, true, 16, 4>(xa, xb, xc, xd, (const float*)A, (const float*)B, (const float*)C, (float*)D,
                                    alpha, beta, ca, cb, cc, cd, threads); break;
    case 3: run<double, true, 8, 4>(xa, xb, xc, xd, (const double*)A, (const double*)B, (const double*)C,
                                    (double*)D, alpha, beta, ca, cb, cc, cd, threads); break;
    default: return 1;
  }
  return 0;
}

// A whole packed panel (ceil(vlen / vr) slivers, pack.rs:57 pack_panel with
// the Real format for real dtypes and Planar for complex) for tests of the
// layout; dtype as packed_contract. Returns 1 for an unknown dtype.
extern "C" int packed_pack_panel(int dtype, const void* src, const int64_t* vscat, int64_t vlen, const int64_t* kscat,
                                 int64_t kc, int vr, int conj, void* out) {
  const int64_t nsl = (vlen + vr - 1) / vr;
  for (int64_t sl = 0; sl < nsl; sl++) {
    switch (dtype) {
      case 0: pack_sliver<float, false>((float*)out + sl * kc * vr, (const float*)src, 0, vscat, sl * vr, vlen, vr,
                                        kscat, kc, conj); break;
      case 1: pack_sliver<double, false>((double*)out + sl * kc * vr, (const double*)src, 0, vscat, sl * vr, vlen,
                                         vr, kscat, kc, conj); break;
      case 2: pack_sliver<float, true>((float*)out + sl * kc * vr * 2, (const float*)src, 0, vscat, sl * vr, vlen,
                                       vr, kscat, kc, conj); break;
      case 3: pack_sliver<double, true>((double*)out + sl * kc * vr * 2, (const double*)src, 0, vscat, sl * vr, vlen,
                                        vr, kscat, kc, conj); break;
      default: return 1;
    }
  }
  return 0;
}
