/* Fixed cost of one TAPP_execute_product call: the 2x2x2 f64 matrix product
 * D[i,j] = sum_k A[i,k] B[k,j] (the shape of the reference's C-ABI fixed-cost
 * benchmark, benchmarks/c/README.md:31) on device pointers, repeated many
 * times, through
 *   - executor 0 (the null stream; every call waits for its kernel),
 *   - a TAPP_create_executor executor (its own stream; every call waits),
 *   - the same executor in asynchronous mode (calls return after enqueueing;
 *     one wait at the end),
 * plus, for scale, one hipPointerGetAttributes and one synchronisation of an
 * idle stream. Prints one JSON object (nanoseconds per call). Needs a GPU. */
#define _POSIX_C_SOURCE 199309L
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "tapp.h"
#include "tprims/tprims.h"

static double now_ns(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec * 1e9 + (double)t.tv_nsec;
}

static int check(const char* what, TAPP_error e) {
  if (e != 0) fprintf(stderr, "%s: status %d (%s)\n", what, (int)e, tprims_last_error());
  return e != 0;
}

int main(int argc, char** argv) {
  int reps = argc > 1 ? atoi(argv[1]) : 20000;
  if (reps < 10) reps = 10;
  double *a = NULL, *b = NULL, *d = NULL;
  const double ha[4] = {1, 2, 3, 4}, hb[4] = {5, 6, 7, 8};
  if (hipMalloc((void**)&a, 32) || hipMalloc((void**)&b, 32) || hipMalloc((void**)&d, 32)) return 2;
  if (hipMemcpy(a, ha, 32, hipMemcpyHostToDevice) || hipMemcpy(b, hb, 32, hipMemcpyHostToDevice)) return 2;

  TAPP_handle h;
  TAPP_tensor_info ia, ib, id;
  const int64_t ext[2] = {2, 2}, str[2] = {1, 2};
  const int64_t la[2] = {'i', 'k'}, lb[2] = {'k', 'j'}, ld[2] = {'i', 'j'};
  if (check("handle", TAPP_create_handle(&h))) return 1;
  if (check("info", TAPP_create_tensor_info(&ia, TAPP_F64, 2, ext, str))) return 1;
  if (check("info", TAPP_create_tensor_info(&ib, TAPP_F64, 2, ext, str))) return 1;
  if (check("info", TAPP_create_tensor_info(&id, TAPP_F64, 2, ext, str))) return 1;
  TAPP_tensor_product plan;
  if (check("plan", TAPP_create_tensor_product(&plan, h, TAPP_IDENTITY, ia, la, TAPP_IDENTITY, ib, lb, TAPP_IDENTITY,
                                               id, ld, TAPP_IDENTITY, id, ld, TAPP_DEFAULT_PREC)))
    return 1;
  const double alpha = 1.0, beta = 0.0;
  TAPP_executor ex;
  if (check("executor", TAPP_create_executor(&ex))) return 1;

  double ns[3], enqueue_ns = 0;
  for (int mode = 0; mode < 3; mode++) {
    TAPP_executor e = mode == 0 ? 0 : ex;
    if (mode == 2 && check("async", tprims_tapp_executor_set_async(ex, 1))) return 1;
    for (int w = 0; w < 200; w++)
      if (check("warm-up", TAPP_execute_product(plan, e, NULL, &alpha, a, b, &beta, NULL, d))) return 1;
    if (check("sync", tprims_tapp_executor_synchronize(e))) return 1;
    const double t0 = now_ns();
    for (int r = 0; r < reps; r++)
      if (TAPP_execute_product(plan, e, NULL, &alpha, a, b, &beta, NULL, d) != 0) return check("execute", 1);
    if (check("sync", tprims_tapp_executor_synchronize(e))) return 1;
    ns[mode] = (now_ns() - t0) / reps;
    if (mode == 2) {
      /* Host time alone: 64 calls into an idle stream (the GPU cannot hold
       * the host back yet), best of 20 bursts. */
      enqueue_ns = 1e30;
      for (int burst = 0; burst < 20; burst++) {
        const double b0 = now_ns();
        for (int r = 0; r < 64; r++)
          if (TAPP_execute_product(plan, e, NULL, &alpha, a, b, &beta, NULL, d) != 0) return check("execute", 1);
        const double t = (now_ns() - b0) / 64;
        if (t < enqueue_ns) enqueue_ns = t;
        if (check("sync", tprims_tapp_executor_synchronize(e))) return 1;
      }
    }
  }
  /* A destroyed executor with queued work is refused (TPRIMS_BUSY), then
   * accepted once synchronized: checked at the end. */
  double hd[4];
  if (hipMemcpy(hd, d, 32, hipMemcpyDeviceToHost)) return 2;
  /* column-major [1 3; 2 4] [5 7; 6 8] = [23 31; 34 46] */
  const int ok = hd[0] == 23 && hd[1] == 34 && hd[2] == 31 && hd[3] == 46;

  hipPointerAttribute_t at;
  double t0 = now_ns();
  for (int r = 0; r < reps; r++) (void)hipPointerGetAttributes(&at, a);
  const double attr_ns = (now_ns() - t0) / reps;
  hipStream_t s;
  (void)hipStreamCreate(&s);
  t0 = now_ns();
  for (int r = 0; r < reps; r++) (void)hipStreamSynchronize(s);
  const double sync_ns = (now_ns() - t0) / reps;

  for (int r = 0; r < 256; r++)
    if (TAPP_execute_product(plan, ex, NULL, &alpha, a, b, &beta, NULL, d) != 0) return check("execute", 1);
  const TAPP_error de = TAPP_destroy_executor(ex);
  const int busy = de == TPRIMS_BUSY;
  if (busy) {
    if (check("sync", tprims_tapp_executor_synchronize(ex))) return 1;
    if (check("destroy", TAPP_destroy_executor(ex))) return 1;
  } else if (check("destroy", de)) {
    return 1;
  }

  printf("{\"reps\": %d, \"correct\": %s, \"default_executor_sync_ns\": %.0f, \"created_executor_sync_ns\": %.0f, "
         "\"created_executor_async_ns\": %.0f, \"async_enqueue_host_ns\": %.0f, \"pointer_attributes_ns\": %.0f, "
         "\"idle_stream_sync_ns\": %.0f, \"destroy_with_queued_work_busy\": %s}\n",
         reps, ok ? "true" : "false", ns[0], ns[1], ns[2], enqueue_ns, attr_ns, sync_ns, busy ? "true" : "false");
  (void)hipStreamDestroy(s);
  TAPP_destroy_tensor_product(plan);
  TAPP_destroy_tensor_info(ia);
  TAPP_destroy_tensor_info(ib);
  TAPP_destroy_tensor_info(id);
  TAPP_destroy_handle(h);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(d);
  return ok ? 0 : 3;
}
