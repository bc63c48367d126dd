"""Benchmark of the contraction path on one MI355X.

Workload (the complex benchmark contraction, c64 by default):
  D[i,k,n] = sum_{j,l} conj(A[j,i,l,n]) * B[l,k,j,n] + 0.5 * C[i,k,n]
  i = k = 512, j = l = 64, n = 16; A is every 2nd element of a 2x-padded
  buffer (a non-unit-stride view), B, C, D column-major; Re/Im ~ N(0, 1).
A "step" is one execution of the planned contraction over resident device
buffers (TAPP_execute_product on an asynchronous executor).

Prints one JSON line: TFLOP/s (8 flops per complex multiply-add), the MFMA
roofline of the contraction kernel (HIP events on the executor's stream), the
relative error against an FP64 evaluation of sampled outputs, variants
(c32 plain FP32, c32 split-BF16, 1m, 3m) and the CPU baseline (the C++ port of
the reference's packed driver, timed on a bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tensoralgebra-rs_amd")]

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2516.6}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", default="c64", choices=["c64", "c32"])
    ap.add_argument("--method", default="native", choices=["native", "1m", "3m"])
    ap.add_argument("--precision", default="storage", choices=["storage", "split2", "split3"])
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scale", type=int, default=1, help="divide i,k by this (quick runs)")
    return ap.parse_args()


class Workload:
    """Buffers and the plan of the complex benchmark contraction."""

    def __init__(self, torch, tp, dtype, method, precision, scale=1, seed=0):
        i = k = 512 // scale
        j = l = 64
        n = 16
        self.dims = dict(i=i, k=k, j=j, l=l, n=n)
        cdt = torch.complex128 if dtype == "c64" else torch.complex64
        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev).manual_seed(seed)
        # A: (j, i, l, n) view with strides 2 * column-major inside a 2x buffer.
        self.Abuf = torch.randn(2 * j * i * l * n, dtype=cdt, device=dev, generator=g)
        self.B = torch.randn(l * k * j * n, dtype=cdt, device=dev, generator=g)
        self.C = torch.randn(i * k * n, dtype=cdt, device=dev, generator=g)
        self.D = torch.zeros(i * k * n, dtype=cdt, device=dev)
        sa = [2, 2 * j, 2 * j * i, 2 * j * i * l]
        sb = tp.col_major([l, k, j, n])
        sd = tp.col_major([i, k, n])
        J, I, L, N, K = 1, 2, 3, 4, 5
        DT = tp.DType.C64 if dtype == "c64" else tp.DType.C32
        A = tp.OperandSpec(tp.LayoutSpec([j, i, l, n], sa), tp.Op.CONJUGATE)
        B = tp.OperandSpec(tp.LayoutSpec([l, k, j, n], sb))
        D = tp.OperandSpec(tp.LayoutSpec([i, k, n], sd))
        prob = tp.Problem.from_labels(DT, A, B, tp.CSpec.separate(D), D,
                                      tp.Labels([J, I, L, N], [L, K, J, N], [I, K, N], [I, K, N]))
        cm = {"native": tp.ComplexMethod.NATIVE, "1m": tp.ComplexMethod.ONE_M, "3m": tp.ComplexMethod.THREE_M}[method]
        pr = {"storage": tp.Precision.STORAGE, "split2": tp.Precision.SPLIT_BF16_2,
              "split3": tp.Precision.SPLIT_BF16_3}[precision]
        self.plan = tp.Plan(prob, tp.PlanConfig(complex_method=cm, precision=pr))
        self.sa, self.sb, self.sd = sa, sb, sd
        self.flops = 8 * i * k * j * l * n
        # Algorithmic bytes: addressed elements of A, B, C (beta != 0) and D (write).
        es = 16 if dtype == "c64" else 8
        self.bytes = es * (j * i * l * n + l * k * j * n + 2 * i * k * n)

    def run(self, ex):
        self.plan.execute_raw(1.0, self.Abuf.data_ptr(), self.B.data_ptr(), 0.5, self.C.data_ptr(),
                              self.D.data_ptr(), ex)

    def sampled_rel_err(self, torch, samples=256, seed=1):
        """Normwise relative error of D against an FP64 evaluation of sampled outputs."""
        import numpy as np

        d = self.dims
        rng = np.random.default_rng(seed)
        ii = rng.integers(0, d["i"], samples)
        kk = rng.integers(0, d["k"], samples)
        nn = rng.integers(0, d["n"], samples)
        j, l = d["j"], d["l"]
        jj, ll = np.meshgrid(np.arange(j), np.arange(l), indexing="ij")
        A = self.Abuf.cpu().numpy().astype(np.complex128)
        B = self.B.cpu().numpy().astype(np.complex128)
        C = self.C.cpu().numpy().astype(np.complex128)
        D = self.D.cpu().numpy().astype(np.complex128)
        want = np.empty(samples, np.complex128)
        got = np.empty(samples, np.complex128)
        for s in range(samples):
            i, k, n = ii[s], kk[s], nn[s]
            aoff = jj * self.sa[0] + i * self.sa[1] + ll * self.sa[2] + n * self.sa[3]
            boff = ll * self.sb[0] + k * self.sb[1] + jj * self.sb[2] + n * self.sb[3]
            doff = i * self.sd[0] + k * self.sd[1] + n * self.sd[2]
            want[s] = np.sum(np.conj(A[aoff]) * B[boff]) + 0.5 * C[doff]
            got[s] = D[doff]
        scale = max(1.0, np.abs(want).max())
        return float(np.abs(got - want).max() / scale)


def time_steps(torch, w, ex, stream, steps, warmup):
    for _ in range(warmup):
        w.run(ex)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        w.run(ex)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / steps
    return ms, wall


def cpu_baseline(args):
    """The C++ port of the reference's packed driver on the host cores, timed
    on one batch entry (n = 1 of 16) of the same contraction."""
    try:
        import oracle.packed as packed
    except Exception as e:  # the port is optional for the GPU figure
        return {"value": None, "unit": "TFLOP/s", "cores": 0, "kind": "port", "sample": "unavailable: %s" % e}
    return packed.bench_complex_sample(dtype=args.dtype, seconds=10.0)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    import tprims_amd as tp

    stream = torch.cuda.Stream()
    ex = tp.Executor(0, stream.cuda_stream, asynchronous=True)
    w = Workload(torch, tp, args.dtype, args.method, args.precision, args.scale)
    ms, wall = time_steps(torch, w, ex, stream, args.steps, args.warmup)
    tflops = w.flops / (ms * 1e-3) / 1e12
    err = w.sampled_rel_err(torch)
    rep = w.plan.report()
    unit_peak = "f64" if args.dtype == "c64" else ("bf16" if args.precision != "storage" else "f32")
    peak = PEAK_TFLOPS[unit_peak]
    if args.precision == "split3":
        peak = PEAK_TFLOPS["bf16"] / 6
    elif args.precision == "split2":
        peak = PEAK_TFLOPS["bf16"] / 3
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (args.dtype, rep["kernel"]))
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    out = {
        "metric": "TFLOP/s + % MFMA peak; split-BF16 vs plain FP32 GPU at stated rel. err",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic N(0,1), seeded",
        "config": {"workload": "D[i,k,n]=sum_{j,l} conj(A[j,i,l,n]) B[l,k,j,n] + 0.5 C[i,k,n]; "
                               "i=k=%d j=l=64 n=16; A stride-2 view of a 2x-padded buffer" % (512 // args.scale),
                   "method": args.method, "precision": args.precision, "kernel": rep["kernel"],
                   "m": rep["m"], "n": rep["n"], "k": rep["k"], "batch": rep["batch"]},
        "rel_err_vs_fp64": err,
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(tflops / peak, 4), "traffic": traffic,
                     "alg_bytes": w.bytes, "hbm_gbs_alg": round(w.bytes / (ms * 1e-3) / 1e9, 1)},
    }
    del w
    torch.cuda.empty_cache()
    if not args.no_variants and args.scale == 1:
        variants = {}
        for dt, m, p in [("c32", "native", "storage"), ("c32", "native", "split3"), ("c32", "native", "split2"),
                         ("c64", "1m", "storage"), ("c64", "3m", "storage"), ("c32", "3m", "storage")]:
            if (dt, m, p) == (args.dtype, args.method, args.precision):
                continue
            try:
                v = Workload(torch, tp, dt, m, p)
                vms, _ = time_steps(torch, v, ex, stream, max(3, args.steps // 2), 2)
                vt = v.flops / (vms * 1e-3) / 1e12
                variants["%s.%s.%s" % (dt, m, p)] = {"tflops": round(vt, 3), "ms": round(vms, 4),
                                                      "rel_err_vs_fp64": v.sampled_rel_err(torch),
                                                      "kernel": v.plan.report()["kernel"]}
                del v
                torch.cuda.empty_cache()
            except Exception as e:  # report, do not hide
                variants["%s.%s.%s" % (dt, m, p)] = {"error": str(e)}
        out["variants"] = variants
        f32 = variants.get("c32.native.storage", {}).get("tflops")
        s3 = variants.get("c32.native.split3", {}).get("tflops")
        if f32 and s3:
            out["split_bf16x3_vs_fp32"] = round(s3 / f32, 3)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
