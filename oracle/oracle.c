/* The label oracle: TEST INFRASTRUCTURE ONLY.
 *
 * Only the test suite, the smoke check and the bench's CPU-baseline leg may
 * load this file's library, and only as the checker. The product path (the
 * HIP library) never calls it.
 *
 * A restatement in C of the reference's brute-force oracle
 * `contract_reference` (crates/tprims-testkit/src/oracle.rs:92-204), computed
 * in FP64 (real) or FP64 complex whatever the storage type of the tested
 * operands, so every GPU result can be stated against an FP64 oracle:
 *
 *   D[out] = op_D(alpha * sum_{sum labels} op_A(A) * op_B(B) + beta * op_C(C))
 *
 * - labels of D are the output labels; every other label is summed
 *   (oracle.rs:124-135), so an index isolated in A or B is a reduction;
 * - a label repeated within one operand selects its diagonal: every mode of
 *   that label reads the same assigned value (oracle.rs:137-143);
 * - beta == 0 never reads C (oracle.rs:176);
 * - conjugation of D is applied last, to the whole sum (oracle.rs:183-185);
 * - a repeated label in D writes only D's diagonal.
 *
 * Data are doubles; a complex operand is interleaved (re, im) pairs and its
 * strides and offset count complex elements. Returns 0, or 2 when a label has
 * two extents (OracleError::Extent, oracle.rs:58).
 *
 * `oracle_max_rel_err` restates `max_rel_err`
 * (crates/tprims-testkit/src/fixtures.rs:81-87): the largest |got - want|
 * normalised by max(1, max |want|).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 64

typedef struct {
  const double* data;
  int n;
  const int64_t* dims;
  const int64_t* strides;
  int64_t offset;
  const int64_t* labels;
  int conj;
  int slot[MAXL]; /* position of each mode's label in the assignment vector */
} operand;

static int add_labels(int64_t* ids, int64_t* ext, int* nl, const operand* o) {
  for (int k = 0; k < o->n; k++) {
    int found = -1;
    for (int j = 0; j < *nl; j++)
      if (ids[j] == o->labels[k]) found = j;
    if (found >= 0) {
      if (ext[found] != o->dims[k]) return 2;
    } else {
      if (*nl >= MAXL) return 3;
      ids[*nl] = o->labels[k];
      ext[*nl] = o->dims[k];
      (*nl)++;
    }
  }
  return 0;
}

static int in_list(const int64_t* l, int n, int64_t x) {
  for (int k = 0; k < n; k++)
    if (l[k] == x) return 1;
  return 0;
}

static void bind(operand* o, const int64_t* order, int n) {
  for (int k = 0; k < o->n; k++)
    for (int j = 0; j < n; j++)
      if (order[j] == o->labels[k]) o->slot[k] = j;
}

static int64_t offset_of(const operand* o, const int64_t* assign) {
  int64_t off = o->offset;
  for (int k = 0; k < o->n; k++) off += assign[o->slot[k]] * o->strides[k];
  return off;
}

int oracle_contract(int cplx, const double* alpha,
                    const double* A, int na, const int64_t* da, const int64_t* sa, int64_t oa, const int64_t* la, int ca,
                    const double* B, int nb, const int64_t* db, const int64_t* sb, int64_t ob, const int64_t* lb, int cb,
                    const double* beta,
                    const double* C, int nc, const int64_t* dc, const int64_t* sc, int64_t oc, const int64_t* lc, int cc,
                    double* D, int nd, const int64_t* dd, const int64_t* sd, int64_t od, const int64_t* ld, int cd) {
  operand a = {A, na, da, sa, oa, la, ca, {0}};
  operand b = {B, nb, db, sb, ob, lb, cb, {0}};
  operand c = {C, nc, dc, sc, oc, lc, cc, {0}};
  operand d = {D, nd, dd, sd, od, ld, cd, {0}};
  int64_t ids[MAXL], ext[MAXL];
  int nl = 0, e;
  if ((e = add_labels(ids, ext, &nl, &a))) return e;
  if ((e = add_labels(ids, ext, &nl, &b))) return e;
  if (C && (e = add_labels(ids, ext, &nl, &c))) return e;
  if ((e = add_labels(ids, ext, &nl, &d))) return e;

  /* Assignment vector: output labels first, then summed labels. */
  int64_t order[MAXL], oext[MAXL];
  int nout = 0, nsum = 0;
  for (int j = 0; j < nl; j++)
    if (in_list(ld, nd, ids[j])) {
      order[nout] = ids[j];
      oext[nout++] = ext[j];
    }
  for (int j = 0; j < nl; j++)
    if (!in_list(ld, nd, ids[j])) {
      order[nout + nsum] = ids[j];
      oext[nout + nsum++] = ext[j];
    }
  bind(&a, order, nout + nsum);
  bind(&b, order, nout + nsum);
  if (C) bind(&c, order, nout + nsum);
  bind(&d, order, nout + nsum);

  int64_t out_total = 1, sum_total = 1;
  for (int j = 0; j < nout; j++) out_total *= oext[j];
  for (int j = 0; j < nsum; j++) sum_total *= oext[nout + j];

  const int w = cplx ? 2 : 1;
  const double ar = alpha[0], ai = cplx ? alpha[1] : 0.0;
  const double br = beta[0], bi = cplx ? beta[1] : 0.0;
  const int read_c = C && (br != 0.0 || bi != 0.0);
  int64_t assign[MAXL];
  memset(assign, 0, sizeof assign);
  for (int64_t o = 0; o < out_total; o++) {
    double accr = 0.0, acci = 0.0;
    for (int j = 0; j < nsum; j++) assign[nout + j] = 0;
    for (int64_t s = 0; s < sum_total; s++) {
      const double* pa = A + w * offset_of(&a, assign);
      const double* pb = B + w * offset_of(&b, assign);
      if (cplx) {
        double xr = pa[0], xi = ca ? -pa[1] : pa[1];
        double yr = pb[0], yi = cb ? -pb[1] : pb[1];
        accr += xr * yr - xi * yi;
        acci += xr * yi + xi * yr;
      } else {
        accr += pa[0] * pb[0];
      }
      for (int j = 0; j < nsum; j++) {
        if (++assign[nout + j] < oext[nout + j]) break;
        assign[nout + j] = 0;
      }
    }
    double vr = ar * accr - ai * acci, vi = ar * acci + ai * accr;
    if (read_c) {
      const double* pc = C + w * offset_of(&c, assign);
      double xr = pc[0], xi = cplx ? (cc ? -pc[1] : pc[1]) : 0.0;
      vr += br * xr - bi * xi;
      vi += br * xi + bi * xr;
    }
    if (cd) vi = -vi;
    double* pd = D + w * offset_of(&d, assign);
    pd[0] = vr;
    if (cplx) pd[1] = vi;
    for (int j = 0; j < nout; j++) {
      if (++assign[j] < oext[j]) break;
      assign[j] = 0;
    }
  }
  return 0;
}

/* Normwise relative error of n elements (interleaved when cplx). */
double oracle_max_rel_err(int cplx, const double* got, const double* want, int64_t n) {
  const int w = cplx ? 2 : 1;
  double scale = 1.0, worst = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double m = cplx ? hypot(want[2 * i], want[2 * i + 1]) : fabs(want[i]);
    if (m > scale) scale = m;
  }
  for (int64_t i = 0; i < n; i++) {
    double dr = got[w * i] - want[w * i];
    double di = cplx ? got[w * i + 1] - want[w * i + 1] : 0.0;
    double e = cplx ? hypot(dr, di) : fabs(dr);
    if (isnan(e)) return NAN;
    if (e / scale > worst) worst = e / scale;
  }
  return worst;
}
