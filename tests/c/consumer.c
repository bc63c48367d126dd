/* A C host of libtprims through the TAPP C ABI alone: the matrix product of
 * the reference's C consumer (crates/tprims-capi/tests/c/standard_consumer.c:
 * D[i,j] = sum_k A[i,k] B[k,j] on column-major host arrays, expected
 * [22, 28, 49, 64]) on a created executor and on executor 0, then the
 * scatter extensions (crates/tprims-kernel/src/pack/scatter.rs:62-68,119-128).
 * Host arrays are staged through the executor's device memory, so running it
 * needs a GPU; compiling it checks the headers as plain C.
 * Exit status 0 and "ok" on success; the failing step otherwise. */
#include <stdio.h>
#include <string.h>

#include "tapp.h"
#include "tprims/tprims.h"

static int failed(const char* what, TAPP_error e) {
  char msg[256];
  TAPP_explain_error(e, sizeof msg, msg);
  fprintf(stderr, "%s: status %d (%s; %s)\n", what, (int)e, msg, tprims_last_error());
  return 1;
}

#define STEP(what, call)                          \
  do {                                            \
    TAPP_error e_ = (call);                       \
    if (!TAPP_check_success(e_)) return failed(what, e_); \
  } while (0)

static int run_product(TAPP_tensor_product plan, TAPP_executor exec) {
  const double a[6] = {1, 2, 3, 4, 5, 6}; /* A(i,k) at i + 2 k */
  const double b[6] = {1, 2, 3, 4, 5, 6}; /* B(k,j) at k + 3 j */
  double d[4] = {-1, -1, -1, -1};
  const double alpha = 1.0, beta = 0.0;
  TAPP_status status = -1;
  STEP("execute", TAPP_execute_product(plan, exec, &status, &alpha, a, b, &beta, TAPP_IN_PLACE, d));
  if (status != 0) {
    fprintf(stderr, "status handle not cleared\n");
    return 1;
  }
  if (!(d[0] == 22 && d[1] == 28 && d[2] == 49 && d[3] == 64)) {
    fprintf(stderr, "values %g %g %g %g\n", d[0], d[1], d[2], d[3]);
    return 1;
  }
  STEP("destroy status", TAPP_destroy_status(status));
  return 0;
}

int main(void) {
  TAPP_handle handle;
  TAPP_executor exec;
  STEP("create handle", TAPP_create_handle(&handle));
  STEP("create executor", TAPP_create_executor(&exec));

  const int64_t ext_a[2] = {2, 3}, str_a[2] = {1, 2};
  const int64_t ext_b[2] = {3, 2}, str_b[2] = {1, 3};
  const int64_t ext_d[2] = {2, 2}, str_d[2] = {1, 2};
  TAPP_tensor_info ta, tb, td;
  STEP("info A", TAPP_create_tensor_info(&ta, TAPP_F64, 2, ext_a, str_a));
  STEP("info B", TAPP_create_tensor_info(&tb, TAPP_F64, 2, ext_b, str_b));
  STEP("info D", TAPP_create_tensor_info(&td, TAPP_F64, 2, ext_d, str_d));

  const int64_t idx_a[2] = {'i', 'k'}, idx_b[2] = {'k', 'j'}, idx_d[2] = {'i', 'j'};
  TAPP_tensor_product plan;
  STEP("create product", TAPP_create_tensor_product(&plan, handle, TAPP_IDENTITY, ta, idx_a, TAPP_IDENTITY, tb, idx_b,
                                                    TAPP_IDENTITY, td, idx_d, TAPP_IDENTITY, td, idx_d,
                                                    TAPP_DEFAULT_PREC));
  if (run_product(plan, exec)) return 1;
  if (run_product(plan, 0)) return 1;

  /* The plan's M scatter and the library's scatter builders. */
  int64_t buf[8], len = 0;
  STEP("product scatter", tprims_product_scatter(plan, TPRIMS_SCATTER_AK, buf, 8, &len));
  if (len != 3 || buf[0] != 0 || buf[1] != 2 || buf[2] != 4) {
    fprintf(stderr, "A's K scatter: length %lld\n", (long long)len);
    return 1;
  }
  const int64_t ext[2] = {2, 3}, str[2] = {1, 10};
  STEP("build scatter", tprims_build_scatter(ext, str, 2, buf, 8, &len));
  const int64_t want[6] = {0, 1, 10, 11, 20, 21};
  if (len != 6 || memcmp(buf, want, sizeof want) != 0) {
    fprintf(stderr, "build_scatter\n");
    return 1;
  }
  int64_t blocks[4], nb = 0;
  STEP("block scatter", tprims_build_block_scatter(want, 6, 3, blocks, 4, &nb));
  if (nb != 2 || blocks[0] != TPRIMS_SCATTER_IRREGULAR || blocks[1] != TPRIMS_SCATTER_IRREGULAR) {
    fprintf(stderr, "build_block_scatter\n");
    return 1;
  }

  STEP("destroy product", TAPP_destroy_tensor_product(plan));
  STEP("destroy A", TAPP_destroy_tensor_info(ta));
  STEP("destroy B", TAPP_destroy_tensor_info(tb));
  STEP("destroy D", TAPP_destroy_tensor_info(td));
  STEP("destroy executor", TAPP_destroy_executor(exec));
  STEP("destroy handle", TAPP_destroy_handle(handle));
  printf("ok\n");
  return 0;
}
