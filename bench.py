"""Benchmark of the contraction path on one MI355X.

Headline workload (the complex benchmark contraction, c64 native):
  D[i,k,n] = sum_{j,l} conj(A[j,i,l,n]) * B[l,k,j,n] + 0.5 * C[i,k,n]
  i = k = 512, j = l = 64, n = 16; A is every 2nd element of a 2x-padded
  buffer (a non-unit-stride view); B, C, D column-major; Re/Im ~ N(0, 1).
A "step" is one execution of the planned contraction over device-resident
buffers (TAPP_execute_product on an asynchronous executor); the time is taken
with HIP events on the executor's stream.

The JSON line also carries, under "configs", the other benchmark
contractions measured the same way:
  f64_24      D[a,b,c,d] = sum_{e,f} A[a,e,b,f] B[f,c,e,d], all extents 24
  f64_gemm4096  f64 D[m,n] = sum_k A[m,k] B[k,n], 4096^3, column-major
  transform   f32 D[p,q,r,s] = sum_t A[p,q,r,t] U[t,s], p=q=r=64, t=s=512,
              rows N(0,1) * 10^U(-3,3): FP32, split-BF16 3-term, 2-term
  gemm4096    f32 4096^3, same distribution and precisions
  grouped     512 MPS contractions E[a,s,b] = sum_{c,t} L[a,c] M[c,t,s,b] (c64):
              one persistent launch (autotuned tile, hipGraph) vs one launch each
  tapp_batched  512 c64 64^3 GEMMs of one plan through TAPP_execute_batched_product
              (one launch) vs one call per item
  elementwise f64 Hadamard D[i] = A[i] B[i], 2^27 elements (HBM-bound)
  vjp         D[i,k,b] = sum_{j,r} A[i,i,j,r,b] B[j,k,b] forward + dA, dB (f32, c64)
with relative errors against FP64 evaluations of sampled outputs, and the CPU
baseline (the C++ port of the reference's packed driver on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tensoralgebra-rs_amd")]

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2516.6, "fp8": 5033.2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # The GPU reaches its steady clock after about 60 ms of this load: with 3
    # untimed steps the headline reads ~66.2 TFLOP/s, with 30 ~68.4.
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--dtype", default="c64", choices=["c64", "c32"])
    ap.add_argument("--method", default="native", choices=["native", "1m", "3m"])
    ap.add_argument("--precision", default="storage", choices=["storage", "split2", "split3", "fp8"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one 1 GiB device-to-device copy (the known-size kernel that the HBM traffic "
                         "counters are scaled by) and exit")
    ap.add_argument("--no-vendor", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--corpus", action="store_true",
                    help="time the 49-case TCCG corpus (the reference tcbench's, 64 MiB per tensor) in every dtype "
                         "and print one JSON line: TFLOP/s per case, min/median/geomean/max per dtype")
    ap.add_argument("--only-config", default=None,
                    choices=["f64_24", "f64_gemm4096", "transform", "gemm4096", "grouped", "tapp_batched",
                             "elementwise", "vjp", "fixed_cost"],
                    help="run one benchmark contraction of the configs table (with --steps and --warmup), print "
                         "its JSON object and exit (for counter passes)")
    return ap.parse_args()


def timed(torch, fn, stream, steps, warmup):
    """Average ms per call of fn() between HIP events on `stream`."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def rel_err(np, got, want):
    scale = max(1.0, float(np.abs(want).max()))
    return float(np.abs(got - want).max() / scale)


def elem_err(np, got, want):
    """Elementwise relative error |got - want| / |want| over the sampled
    outputs (nonzero want): median and 99th percentile. Unlike the normwise
    figure, it sees an error confined to small-magnitude rows."""
    got, want = np.asarray(got), np.asarray(want)
    nz = np.abs(want) > 0
    r = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    return {"p50": float(np.percentile(r, 50)), "p99": float(np.percentile(r, 99)), "samples": int(nz.sum())}


class Complex:
    """The headline complex contraction."""

    def __init__(self, torch, tp, dtype, method, precision, seed=0):
        i = k = 512
        j = l = 64
        n = 16
        self.dims = dict(i=i, k=k, j=j, l=l, n=n)
        cdt = torch.complex128 if dtype == "c64" else torch.complex64
        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev).manual_seed(seed)
        self.Abuf = torch.randn(2 * j * i * l * n, dtype=cdt, device=dev, generator=g)
        self.B = torch.randn(l * k * j * n, dtype=cdt, device=dev, generator=g)
        self.C = torch.randn(i * k * n, dtype=cdt, device=dev, generator=g)
        self.D = torch.zeros(i * k * n, dtype=cdt, device=dev)
        self.sa = [2, 2 * j, 2 * j * i, 2 * j * i * l]
        self.sb = tp.col_major([l, k, j, n])
        self.sd = tp.col_major([i, k, n])
        J, I, L, N, K = 1, 2, 3, 4, 5
        DT = tp.DType.C64 if dtype == "c64" else tp.DType.C32
        A = tp.OperandSpec(tp.LayoutSpec([j, i, l, n], self.sa), tp.Op.CONJUGATE)
        B = tp.OperandSpec(tp.LayoutSpec([l, k, j, n], self.sb))
        D = tp.OperandSpec(tp.LayoutSpec([i, k, n], self.sd))
        prob = tp.Problem.from_labels(DT, A, B, tp.CSpec.separate(D), D,
                                      tp.Labels([J, I, L, N], [L, K, J, N], [I, K, N], [I, K, N]))
        cm = {"native": tp.ComplexMethod.NATIVE, "1m": tp.ComplexMethod.ONE_M, "3m": tp.ComplexMethod.THREE_M}[method]
        pr = {"storage": tp.Precision.STORAGE, "split2": tp.Precision.SPLIT_BF16_2,
              "split3": tp.Precision.SPLIT_BF16_3, "fp8": tp.Precision.FP8_E4M3}[precision]
        self.plan = tp.Plan(prob, tp.PlanConfig(complex_method=cm, precision=pr))
        self.flops = 8 * i * k * j * l * n
        es = 16 if dtype == "c64" else 8
        self.bytes = es * (j * i * l * n + l * k * j * n + 2 * i * k * n)

    def run(self, ex):
        self.plan.execute_raw(1.0, self.Abuf.data_ptr(), self.B.data_ptr(), 0.5, self.C.data_ptr(),
                              self.D.data_ptr(), ex)

    def sampled_rel_err(self, np, samples=256, seed=1):
        d = self.dims
        rng = np.random.default_rng(seed)
        ii, kk, nn = (rng.integers(0, d[x], samples) for x in ("i", "k", "n"))
        jj, ll = np.meshgrid(np.arange(d["j"]), np.arange(d["l"]), indexing="ij")
        A = self.Abuf.cpu().numpy().astype(np.complex128)
        B = self.B.cpu().numpy().astype(np.complex128)
        C = self.C.cpu().numpy().astype(np.complex128)
        D = self.D.cpu().numpy().astype(np.complex128)
        want, got = np.empty(samples, complex), np.empty(samples, complex)
        for s in range(samples):
            i, k, n = ii[s], kk[s], nn[s]
            aoff = jj * self.sa[0] + i * self.sa[1] + ll * self.sa[2] + n * self.sa[3]
            boff = ll * self.sb[0] + k * self.sb[1] + jj * self.sb[2] + n * self.sb[3]
            doff = i * self.sd[0] + k * self.sd[1] + n * self.sd[2]
            want[s] = np.sum(np.conj(A[aoff]) * B[boff]) + 0.5 * C[doff]
            got[s] = D[doff]
        self.elem = elem_err(np, got, want)
        return rel_err(np, got, want)


def config_f64_24(torch, tp, np, ex, stream, steps):
    e = 24
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(2)
    A = torch.rand(e ** 4, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    B = torch.rand(e ** 4, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    D = torch.zeros(e ** 4, dtype=torch.float64, device=dev)
    cm = tp.col_major([e] * 4)
    sp = lambda: tp.OperandSpec(tp.LayoutSpec([e] * 4, cm))
    prob = tp.Problem.from_labels(tp.DType.F64, sp(), sp(), tp.CSpec.separate(sp()), sp(),
                                  tp.Labels([0, 4, 1, 5], [5, 2, 4, 3], [0, 1, 2, 3], [0, 1, 2, 3]))
    plan = tp.Plan(prob)
    run = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
    ms = timed(torch, run, stream, max(steps, 50), 5)
    flops = 2 * e ** 6
    An = A.cpu().numpy().reshape([e] * 4, order="F")
    Bn = B.cpu().numpy().reshape([e] * 4, order="F")
    want = np.einsum("aebf,fced->abcd", An, Bn)
    got = D.cpu().numpy().reshape([e] * 4, order="F")
    out = {"workload": "f64 D[a,b,c,d]=sum_{e,f} A[a,e,b,f] B[f,c,e,d], extents 24", "ms": round(ms, 4),
           "tflops": round(flops / ms / 1e9, 3), "frac_fp64_peak": round(flops / ms / 1e9 / PEAK_TFLOPS["f64"], 4),
           "rel_err_vs_fp64": rel_err(np, got, want), "rel_err_elementwise": elem_err(np, got, want),
           "kernel": plan.report()["kernel"]}
    # The same contraction on a plan autotuned on these operands
    # (tprims_product_autotune; the process's tuning cache is cleared after).
    tuned = tp.Plan(prob)
    tuned.autotune(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex, reps=20)
    tp.tuning_cache_clear()
    run_t = lambda: tuned.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
    ms_t = timed(torch, run_t, stream, max(steps, 50), 5)
    got_t = D.cpu().numpy().reshape([e] * 4, order="F")
    out["autotuned"] = {"ms": round(ms_t, 4), "tflops": round(flops / ms_t / 1e9, 3),
                        "kernel": tuned.report()["kernel"], "rel_err_vs_fp64": rel_err(np, got_t, want)}
    return out


def config_f64_gemm(torch, tp, np, ex, stream, steps, n=4096):
    """f64 GEMM-shaped contraction D[m,n] = sum_k A[m,k] B[k,n], column-major."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(7)
    A = torch.randn(n * n, dtype=torch.float64, device=dev, generator=g)
    B = torch.randn(n * n, dtype=torch.float64, device=dev, generator=g)
    D = torch.zeros(n * n, dtype=torch.float64, device=dev)
    sp = lambda: tp.OperandSpec(tp.LayoutSpec([n, n], [1, n]))
    plan = tp.Plan(tp.Problem.from_labels(tp.DType.F64, sp(), sp(), tp.CSpec.separate(sp()), sp(),
                                          tp.Labels([0, 2], [2, 1], [0, 1], [0, 1])))
    run = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
    ms = timed(torch, run, stream, max(3, steps // 2), 2)
    rng = np.random.default_rng(8)
    si, sj = rng.integers(0, n, 256), rng.integers(0, n, 256)
    Ah = A.cpu().numpy().reshape(n, n)   # [k, m]
    Bh = B.cpu().numpy().reshape(n, n)   # [n, k]
    Dh = D.cpu().numpy().reshape(n, n)   # [n, m]
    want = np.array([Ah[:, i] @ Bh[j, :] for i, j in zip(si, sj)])
    got = np.array([Dh[j, i] for i, j in zip(si, sj)])
    tf = 2 * n ** 3 / ms / 1e9
    return {"workload": "f64 GEMM D[m,n]=sum_k A[m,k] B[k,n], %d^3, column-major" % n, "ms": round(ms, 4),
            "tflops": round(tf, 3), "frac_fp64_peak": round(tf / PEAK_TFLOPS["f64"], 4),
            "rel_err_vs_fp64": rel_err(np, got, want), "rel_err_elementwise": elem_err(np, got, want),
            "kernel": plan.report()["kernel"]}


def scaled_rows(torch, rows, cols, gen, dev):
    x = torch.randn(rows, cols, dtype=torch.float64, device=dev, generator=gen)
    scale = 10.0 ** (torch.rand(rows, 1, dtype=torch.float64, device=dev, generator=gen) * 6 - 3)
    return (x * scale)


def config_precision(torch, tp, np, ex, stream, steps, shape):
    """f32 contraction at three precisions; A rows and U rows scaled by 10^U(-3,3)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(3)
    if shape == "transform":
        M, K, N = 64 ** 3, 512, 512
        desc = "f32 D[p,q,r,s]=sum_t A[p,q,r,t] U[t,s], p=q=r=64, t=s=512"
        a_dims, a_lab, d_dims, d_lab = [64, 64, 64, 512], [0, 1, 2, 3], [64, 64, 64, 512], [0, 1, 2, 4]
        b_dims, b_lab = [512, 512], [3, 4]
    else:
        M = K = N = 4096
        desc = "f32 GEMM D[m,n]=sum_k A[m,k] B[k,n], 4096^3"
        a_dims, a_lab, d_dims, d_lab, b_dims, b_lab = [M, K], [0, 3], [M, N], [0, 4], [K, N], [3, 4]
    # Column-major A (M x K): rows are the M index; U/B (K x N) rows are k.
    A64 = scaled_rows(torch, M, K, g, dev)          # [m, k]
    B64 = scaled_rows(torch, K, N, g, dev)          # [k, n]
    A = A64.t().contiguous().to(torch.float32)      # column-major storage of A
    B = B64.t().contiguous().to(torch.float32)
    D = torch.zeros(M * N, dtype=torch.float32, device=dev)
    specs = lambda dims: tp.OperandSpec(tp.LayoutSpec(dims, tp.col_major(dims)))
    out = {"workload": desc + "; rows N(0,1)*10^U(-3,3)", "flops": 2 * M * N * K}
    rng = np.random.default_rng(4)
    si, sj = rng.integers(0, M, 2048), rng.integers(0, N, 2048)
    Ah = A.cpu().numpy().astype(np.float64).reshape(K, M)   # [k, m]
    Bh = B.cpu().numpy().astype(np.float64).reshape(N, K)   # [n, k]
    want = np.array([Ah[:, i] @ Bh[j, :] for i, j in zip(si, sj)])
    del A64, B64
    for name, prec in (("fp32", tp.Precision.STORAGE), ("split_bf16x3", tp.Precision.SPLIT_BF16_3),
                       ("split_bf16x2", tp.Precision.SPLIT_BF16_2), ("fp8_e4m3", tp.Precision.FP8_E4M3)):
        prob = tp.Problem.from_labels(tp.DType.F32, specs(a_dims), specs(b_dims), tp.CSpec.separate(specs(d_dims)),
                                      specs(d_dims), tp.Labels(a_lab, b_lab, d_lab, d_lab))
        plan = tp.Plan(prob, tp.PlanConfig(precision=prec))
        run = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
        D.fill_(float("nan"))  # every precision writes all of D itself (beta = 0)
        torch.cuda.synchronize()
        ms = timed(torch, run, stream, max(3, steps // 2), 5)
        Dh = D.cpu().numpy().reshape(N, M)
        got = np.array([Dh[j, i] for i, j in zip(si, sj)], dtype=np.float64)
        tf = out["flops"] / ms / 1e9
        peak = {tp.Precision.STORAGE: PEAK_TFLOPS["f32"], tp.Precision.SPLIT_BF16_3: PEAK_TFLOPS["bf16"] / 6,
                tp.Precision.SPLIT_BF16_2: PEAK_TFLOPS["bf16"] / 3, tp.Precision.FP8_E4M3: PEAK_TFLOPS["fp8"]}[prec]
        out[name] = {"ms": round(ms, 4), "tflops": round(tf, 3), "frac_of_derived_peak": round(tf / peak, 4),
                     "rel_err_vs_fp64": rel_err(np, got, want), "rel_err_elementwise": elem_err(np, got, want),
                     "kernel": plan.report()["kernel"]}
    out["split3_vs_fp32"] = round(out["fp32"]["ms"] / out["split_bf16x3"]["ms"], 3)
    return out


def config_grouped(torch, tp, np, ex, stream, steps, n_items=512):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    g = torch.Generator(device=dev).manual_seed(6)
    plans, A, B, C, D = [], [], [], [], []
    flops = 0
    for _ in range(n_items):
        a, b, c = (int(x) for x in rng.choice(np.arange(16, 257, 16), 3))
        s, t = (int(x) for x in rng.choice([2, 4], 2))
        L = tp.OperandSpec(tp.LayoutSpec([a, c], tp.col_major([a, c])))
        Mt = tp.OperandSpec(tp.LayoutSpec([c, t, s, b], tp.col_major([c, t, s, b])))
        E = tp.OperandSpec(tp.LayoutSpec([a, s, b], tp.col_major([a, s, b])))
        plans.append(tp.Plan(tp.Problem.from_labels(tp.DType.C64, L, Mt, tp.CSpec.separate(E), E,
                                                    tp.Labels([1, 3], [3, 4, 2, 5], [1, 2, 5], [1, 2, 5]))))
        A.append(torch.randn(a * c, dtype=torch.complex128, device=dev, generator=g))
        B.append(torch.randn(c * t * s * b, dtype=torch.complex128, device=dev, generator=g))
        D.append(torch.zeros(a * s * b, dtype=torch.complex128, device=dev))
        flops += 8 * a * c * t * s * b
    grp = tp.Group(plans, [1.0] * n_items, A, B, [0.0] * n_items, None, D, executor=ex)
    tile, tune_ms = grp.autotune(reps=3)
    t_group = timed(torch, grp.execute, stream, steps, 2)
    ref = [d.clone() for d in D]
    t_each = timed(torch, grp.execute_each, stream, max(3, steps // 2), 1)
    same = all(torch.allclose(x, y, rtol=1e-12, atol=1e-12) for x, y in zip(ref, D))
    grp.use_graph(True)
    t_graph = timed(torch, grp.execute, stream, steps, 2)
    return {"workload": "%d c64 MPS contractions E[a,s,b]=sum_{c,t} L[a,c] M[c,t,s,b], a,b,c in 16..256 step 16, "
                        "s,t in {2,4}" % n_items, "flops": flops, "tile": tile,
            "grouped_ms": round(t_group, 4), "grouped_tflops": round(flops / t_group / 1e9, 3),
            "graph_ms": round(t_graph, 4), "graph_tflops": round(flops / t_graph / 1e9, 3),
            "per_launch_ms": round(t_each, 4), "per_launch_tflops": round(flops / t_each / 1e9, 3),
            "speedup_vs_per_launch": round(t_each / min(t_group, t_graph), 3), "matches_per_launch": same}


def config_tapp_batched(torch, tp, np, ex, stream, steps, n=512, e=64):
    """TAPP_execute_batched_product: one c64 plan (e^3 GEMM) on n items, as one
    launch (items folded into the batch offsets) vs one call per item."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(9)
    sp = lambda: tp.OperandSpec(tp.LayoutSpec([e, e], [1, e]))
    plan = tp.Plan(tp.Problem.from_labels(tp.DType.C64, sp(), sp(), tp.CSpec.separate(sp()), sp(),
                                          tp.Labels([0, 2], [2, 1], [0, 1], [0, 1])))
    A = torch.randn(n, e * e, dtype=torch.complex128, device=dev, generator=g)
    B = torch.randn(n, e * e, dtype=torch.complex128, device=dev, generator=g)
    D = torch.zeros(n, e * e, dtype=torch.complex128, device=dev)
    pa, pb, pd = ([x[i].data_ptr() for i in range(n)] for x in (A, B, D))
    fused = lambda: plan.execute_batched_raw(1.0, pa, pb, 0.0, None, pd, ex)

    def each():
        for i in range(n):
            plan.execute_raw(1.0, pa[i], pb[i], 0.0, None, pd[i], ex)

    t_fused = timed(torch, fused, stream, steps, 2)
    got = D.clone()
    t_each = timed(torch, each, stream, max(2, steps // 4), 1)
    flops = 8 * e ** 3 * n
    i = n // 3
    want = (A[i].view(e, e).t().cpu() @ B[i].view(e, e).t().cpu()).t().reshape(-1)
    err = float((got[i].cpu() - want).abs().max() / want.abs().max())
    return {"workload": "%d c64 GEMMs %d^3 through TAPP_execute_batched_product" % (n, e), "flops": flops,
            "batched_ms": round(t_fused, 4), "batched_tflops": round(flops / t_fused / 1e9, 3),
            "per_call_ms": round(t_each, 4), "per_call_tflops": round(flops / t_each / 1e9, 3),
            "speedup": round(t_each / t_fused, 2), "rel_err_item": err, "kernel": plan.report()["kernel"]}


def config_elementwise(torch, tp, np, ex, stream, steps, n=1 << 27):
    """The elementwise (all-batch) strategy on a Hadamard product of 2^27 f64
    (D[i] = A[i] B[i]): HBM-bound, 24 bytes per element."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(10)
    A = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    B = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    D = torch.empty(n, dtype=torch.float64, device=dev)
    sp = lambda: tp.OperandSpec(tp.LayoutSpec([n], [1]))
    plan = tp.Plan(tp.Problem.from_labels(tp.DType.F64, sp(), sp(), tp.CSpec.separate(sp()), sp(),
                                          tp.Labels([0], [0], [0], [0])))
    run = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
    ms = timed(torch, run, stream, max(3, steps // 2), 2)
    ok = bool(torch.equal(D, A * B))
    gbs = 24 * n / ms / 1e6
    return {"workload": "f64 Hadamard D[i]=A[i]B[i], 2^27 elements", "ms": round(ms, 4),
            "hbm_gbs": round(gbs, 1), "frac_hbm_peak": round(gbs / 8000.0, 4), "exact": ok,
            "kernel": plan.report()["kernel"]}


def config_vjp(torch, tp, np, ex, stream, steps):
    dev = torch.device("cuda:0")
    out = {"workload": "D[i,k,b]=sum_{j,r} A[i,i,j,r,b] B[j,k,b], i=k=256, j=128, r=8, b=32: forward + dA, dB"}
    for dt, tdt in (("f32", torch.float32), ("c64", torch.complex128)):
        i, k, j, r, b = 256, 256, 128, 8, 32
        g = torch.Generator(device=dev).manual_seed(7)
        A = torch.randn(i, i, j, r, b, dtype=tdt, device=dev, generator=g)
        B = torch.randn(j, k, b, dtype=tdt, device=dev, generator=g)
        D = torch.empty(i, k, b, dtype=tdt, device=dev)
        G = torch.randn(i, k, b, dtype=tdt, device=dev, generator=g)
        spec = lambda t: tp.OperandSpec(tp.LayoutSpec.of(t))
        DT = tp.DType.F32 if dt == "f32" else tp.DType.C64
        lab = tp.Labels([1, 1, 2, 3, 4], [2, 5, 4], [1, 5, 4], [1, 5, 4])
        plan = tp.Plan(tp.Problem.from_labels(DT, spec(A), spec(B), tp.CSpec.separate(spec(D)), spec(D), lab))
        fwd = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
        t_fwd = timed(torch, fwd, stream, max(3, steps // 2), 2)
        dA = torch.empty_like(A)
        dB = torch.empty_like(B)
        bwd = lambda: plan.vjp(1.0, A, B, G, dA, dB, executor=ex)
        t_bwd = timed(torch, bwd, stream, max(3, steps // 4), 1)
        # check one diagonal column of dA and dB against FP64 sums
        Gc, Bc = G.cpu().to(torch.complex128), B.cpu().to(torch.complex128)
        want_a = torch.einsum("kb,jkb->jb", Gc[3], Bc.conj())  # dA[3,3,:,r,:] for every r
        got_a = dA[3, 3, :, 5, :].cpu().to(torch.complex128)
        Ac = A[:, :, :, :, :].diagonal(dim1=0, dim2=1).cpu().to(torch.complex128)  # [j,r,b,i]
        want_b = torch.einsum("jrbi,ib->jb", Ac.conj(), Gc[:, 7, :])
        got_b = dB[:, 7, :].cpu().to(torch.complex128)
        e_a = rel_err(np, got_a.numpy(), want_a.numpy())
        e_b = rel_err(np, got_b.numpy(), want_b.numpy())
        fl = (2 if dt == "f32" else 8) * i * k * j * r * b
        out[dt] = {"forward_ms": round(t_fwd, 4), "forward_tflops": round(fl / t_fwd / 1e9, 3),
                   "vjp_ms": round(t_bwd, 4), "dA_bytes": A.numel() * A.element_size(),
                   "vjp_dense_dA_GBs": round((A.numel() * A.element_size()) / t_bwd / 1e6, 1),
                   "rel_err_dA": e_a, "rel_err_dB": e_b}
        del A, B, D, G, dA, dB
        torch.cuda.empty_cache()
    return out


def vendor_blas(torch):
    """Same-box vendor GEMM rates (torch.matmul: hipBLASLt / rocBLAS) beside the
    spec peaks: what the library reaches on this GPU at the same clocks."""
    out = {}
    for name, dt, n in (("zgemm_4096", torch.complex128, 4096), ("dgemm_4096", torch.float64, 4096),
                        ("sgemm_4096", torch.float32, 4096), ("bf16_gemm_8192", torch.bfloat16, 8192)):
        a = torch.randn(n, n, dtype=dt, device="cuda")
        b = torch.randn(n, n, dtype=dt, device="cuda")
        ms = timed(torch, lambda: a @ b, torch.cuda.current_stream(), 5, 2)
        out[name + "_tflops"] = round((8 if dt.is_complex else 2) * n ** 3 / ms / 1e9, 1)
        del a, b
        torch.cuda.empty_cache()
    return out


def copy_calibration(torch):
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": src.numel() * 8}))


def config_fixed_cost():
    """tests/c/fixed_cost.c: the host cost of one 2x2x2 TAPP_execute_product on
    device pointers (ns per call; the reference's C-ABI fixed cost is 661 ns on
    its CPU path, benchmarks/c/README.md:31)."""
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_capi_cpu import build_fixed_cost

    with tempfile.TemporaryDirectory() as tmp:
        exe = build_fixed_cost(tmp)
        out = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=120)
    if out.returncode != 0:
        return {"error": "exit %d: %s" % (out.returncode, out.stderr.strip()[-300:])}
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    rec["workload"] = "2x2x2 f64 TAPP_execute_product on device pointers, 20000 calls per executor mode"
    return rec


# The reference's published TCCG corpus figures (49 cases, 64 MiB, GF/s on one
# Ice Lake-SP core, docs/archive/tensorprimitives/results.md:88-93).
TCCG_REFERENCE_GFS = {"f32": {"min": 11.0, "median": 74.8, "geomean": 77.2, "max": 168.8},
                      "f64": {"min": 7.1, "median": 44.7, "geomean": 43.7, "max": 71.6},
                      "c32": {"min": 25.5, "median": 121.0, "geomean": 113.9, "max": 179.3},
                      "c64": {"min": 15.7, "median": 62.0, "geomean": 63.3, "max": 91.1}}


def corpus_run(torch, tp, ex, stream, steps, warmup):
    """Every TCCG case (tests/tccg.py) in every dtype at 64 MiB per tensor:
    inputs N(0,1) resident in HBM, the planner's default plan, HIP events on
    the executor's stream."""
    import math

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import tccg

    dts = [("f32", tp.DType.F32, torch.float32), ("f64", tp.DType.F64, torch.float64),
           ("c32", tp.DType.C32, torch.complex64), ("c64", tp.DType.C64, torch.complex128)]
    rows, summary = [], {}
    for name, DT, tdt in dts:
        vals = []
        for case in tccg.corpus():
            s = tccg.Sized(case, 64 * 1024 * 1024)
            spec = lambda x: tp.OperandSpec(tp.LayoutSpec(*s.layout(x)))
            lab = lambda x: [ord(ch) for ch in x]
            prob = tp.Problem.from_labels(DT, spec(s.a), spec(s.b), tp.CSpec.separate(spec(s.c)), spec(s.c),
                                          tp.Labels(lab(s.a), lab(s.b), lab(s.c), lab(s.c)))
            plan = tp.Plan(prob)
            g = torch.Generator(device="cuda").manual_seed(1)
            A = torch.randn(s.storage(s.a), dtype=tdt, device="cuda", generator=g)
            B = torch.randn(s.storage(s.b), dtype=tdt, device="cuda", generator=g)
            D = torch.empty(s.storage(s.c), dtype=tdt, device="cuda")
            fn = lambda: plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr(), ex)
            ms = timed(torch, fn, stream, steps, warmup)
            flops = (8 if DT.is_complex else 2) * s.macs()
            tf = flops / (ms * 1e-3) / 1e12
            vals.append(tf)
            rows.append({"case": case[0], "group": case[1], "dtype": name, "ms": round(ms, 4),
                         "tflops": round(tf, 3), "kernel": plan.report()["kernel"]})
            del A, B, D, plan
        vals.sort()
        med = vals[len(vals) // 2] if len(vals) % 2 else 0.5 * (vals[len(vals) // 2 - 1] + vals[len(vals) // 2])
        geo = math.exp(sum(math.log(v) for v in vals) / len(vals))
        summary[name] = {"min": round(vals[0], 3), "median": round(med, 3), "geomean": round(geo, 3),
                         "max": round(vals[-1], 3),
                         "reference_one_core_gflops": TCCG_REFERENCE_GFS[name]}
        torch.cuda.empty_cache()
        print("bench: corpus %s done" % name, file=sys.stderr, flush=True)
    return {"metric": "TCCG corpus TFLOP/s (49 cases, 64 MiB per tensor)", "unit": "TFLOP/s", "steps": steps,
            "warmup": warmup, "summary": summary, "cases": rows}


def cpu_baseline(args):
    """The C++ port of the reference's packed driver on the host cores, timed
    on one batch entry (n = 1 of 16) of the headline contraction, on every
    physical core and (one_thread) on one."""
    try:
        import oracle.packed as packed
    except Exception as e:
        return {"value": None, "unit": "TFLOP/s", "cores": 0, "kind": "port", "sample": "unavailable: %s" % e}
    rec = packed.bench_complex_sample(dtype=args.dtype, seconds=10.0)
    one = packed.bench_complex_sample(dtype=args.dtype, seconds=3.0, threads=1)
    rec["one_thread"] = {"value": one["value"], "sample": one["sample"]}
    return rec


def config_table(torch, tp, np, ex, stream, steps):
    """The benchmark contractions beside the headline, by name."""
    return (("f64_24", lambda: config_f64_24(torch, tp, np, ex, stream, steps)),
            ("f64_gemm4096", lambda: config_f64_gemm(torch, tp, np, ex, stream, steps)),
            ("transform", lambda: config_precision(torch, tp, np, ex, stream, steps, "transform")),
            ("gemm4096", lambda: config_precision(torch, tp, np, ex, stream, steps, "gemm")),
            ("grouped", lambda: config_grouped(torch, tp, np, ex, stream, steps)),
            ("tapp_batched", lambda: config_tapp_batched(torch, tp, np, ex, stream, steps)),
            ("elementwise", lambda: config_elementwise(torch, tp, np, ex, stream, steps)),
            ("vjp", lambda: config_vjp(torch, tp, np, ex, stream, steps)),
            ("fixed_cost", config_fixed_cost))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import tprims_amd as tp

    if args.copy_calibration:
        copy_calibration(torch)
        return
    stream = torch.cuda.Stream()
    ex = tp.Executor(0, stream.cuda_stream, asynchronous=True)
    table = config_table(torch, tp, np, ex, stream, args.steps)
    if args.only_config:
        print(json.dumps({"config": args.only_config, **dict(table)[args.only_config]()}))
        return
    if args.corpus:
        print(json.dumps(corpus_run(torch, tp, ex, stream, args.steps, min(args.warmup, 3))))
        return
    w = Complex(torch, tp, args.dtype, args.method, args.precision)
    t0 = time.perf_counter()
    ms = timed(torch, lambda: w.run(ex), stream, args.steps, args.warmup)
    tflops = w.flops / (ms * 1e-3) / 1e12
    err = w.sampled_rel_err(np)
    rep = w.plan.report()
    unit_peak = "f64" if args.dtype == "c64" else ("bf16" if args.precision != "storage" else "f32")
    peak = PEAK_TFLOPS[unit_peak]
    if args.precision == "split3":
        peak = PEAK_TFLOPS["bf16"] / 6
    elif args.precision == "split2":
        peak = PEAK_TFLOPS["bf16"] / 3
    elif args.precision == "fp8":
        peak = PEAK_TFLOPS["fp8"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_complex_benchmark.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(rep["kernel"], {}).get("bytes_per_launch")
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
                               "i=k=512 j=l=64 n=16; A stride-2 view of a 2x-padded buffer",
                   "method": args.method, "precision": args.precision, "kernel": rep["kernel"],
                   "m": rep["m"], "n": rep["n"], "k": rep["k"], "batch": rep["batch"]},
        "rel_err_vs_fp64": err,
        "rel_err_elementwise": w.elem,
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(tflops / peak, 4), "traffic": traffic,
                     "alg_bytes": w.bytes, "alg_hbm_gbs": round(w.bytes / (ms * 1e-3) / 1e9, 1)},
    }
    del w
    torch.cuda.empty_cache()
    if not args.no_variants:
        variants = {}
        for dt, m, p in [("c32", "native", "storage"), ("c32", "native", "split3"), ("c32", "native", "split2"),
                         ("c32", "native", "fp8"),
                         ("c64", "1m", "storage"), ("c64", "3m", "storage"), ("c32", "3m", "storage")]:
            if (dt, m, p) == (args.dtype, args.method, args.precision):
                continue
            key = "%s.%s.%s" % (dt, m, p)
            try:
                v = Complex(torch, tp, dt, m, p)
                vms = timed(torch, lambda: v.run(ex), stream, max(3, args.steps // 2), 2)
                vt = v.flops / (vms * 1e-3) / 1e12
                variants[key] = {"tflops": round(vt, 3), "ms": round(vms, 4), "rel_err_vs_fp64": v.sampled_rel_err(np),
                                 "rel_err_elementwise": v.elem, "kernel": v.plan.report()["kernel"]}
                del v
                torch.cuda.empty_cache()
            except Exception as e:
                variants[key] = {"error": str(e)}
        out["variants"] = variants
        f32 = variants.get("c32.native.storage", {}).get("tflops")
        s3 = variants.get("c32.native.split3", {}).get("tflops")
        if f32 and s3:
            out["split_bf16x3_vs_fp32"] = round(s3 / f32, 3)
    if not args.no_configs:
        cfgs = {}
        for name, fn in table:
            try:
                cfgs[name] = fn()
            except Exception as e:
                cfgs[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            torch.cuda.empty_cache()
            print("bench: %s done" % name, file=sys.stderr, flush=True)
        out["configs"] = cfgs
    if not args.no_vendor:
        out["same_box_vendor_blas"] = vendor_blas(torch)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
