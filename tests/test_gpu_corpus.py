"""GPU replay of the reference benchmark suite's dot_general corpus
(benchmarks/benchmarks/tprims/corpus/*.json; the shapes are extracted to
tests/golden/corpus_dot_general.json by tests/golden/make_corpus.py): every
entry (operand dims and strides, including negative ones, contracted and batch
axes, conjugation) runs through the library on the GPU with seeded inputs.

Plans are built by the library's Problem.from_dot_general, as the corpus's
callers build them. Entries the oracle finishes quickly are checked against it in full; larger
ones on a fixed sample of output elements, each evaluated in FP64 straight
from the definition D[out] = sum over the contracted index space of
op(A) op(B) (crates/tprims-testkit/src/oracle.rs:92-204 restated per element)."""
import json
import math
import os

import numpy as np
import pytest

import tprims_amd as tp
from harness import NP, T, check, run, tol

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "corpus_dot_general.json")) as f:
    ENTRIES = json.load(f)
DT = {"f32": tp.DType.F32, "f64": tp.DType.F64, "c32": tp.DType.C32, "c64": tp.DType.C64}
FULL_MACS, FULL_STORAGE = 2e7, 4e6


def labels(e):
    """Problem::from_dot_general's synthetic labels (api/problem.rs:396-449):
    batch axes first, then contracted, then A's free and B's free axes; the
    output is [A free, B free, batch]."""
    ra, rb = len(e["a"]["dims"]), len(e["b"]["dims"])
    nb, nc = len(e["lb"]), len(e["lc"])
    la, lb, ld = [0] * ra, [0] * rb, []
    nxt = nb + nc
    for x in range(ra):
        if x not in e["lc"] and x not in e["lb"]:
            la[x] = nxt
            ld.append(nxt)
            nxt += 1
    for x in range(rb):
        if x not in e["rc"] and x not in e["rb"]:
            lb[x] = nxt
            ld.append(nxt)
            nxt += 1
    for p, (l, r) in enumerate(zip(e["lb"], e["rb"])):
        la[l] = p
        lb[r] = p
        ld.append(p)
    for p, (l, r) in enumerate(zip(e["lc"], e["rc"])):
        la[l] = nb + p
        lb[r] = nb + p
    return la, lb, ld


def tensors(e):
    la, lb, ld = labels(e)
    conj = e.get("conj", [False, False])
    a = T(e["a"]["dims"], e["a"]["strides"], la, conj=bool(conj[0]))
    b = T(e["b"]["dims"], e["b"]["strides"], lb, conj=bool(conj[1]))
    d = T(e["c"]["dims"], e["c"]["strides"], ld)
    return a, b, d


def macs(e):
    return math.prod(e["c"]["dims"]) * math.prod(e["a"]["dims"][i] for i in e["lc"])


def rand(dtype, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, n)
    if dtype.is_complex:
        v = v + 1j * rng.uniform(-1, 1, n)
    return v.astype(NP[dtype])


def dot_general_plan(e, dt):
    """The plan as a caller of the reference's API builds it:
    Problem.from_dot_general (api/problem.rs:396) over the entry's layouts.
    The oracle side keeps the test's own label rule (labels() above), so the
    library's synthetic labels are checked against an independent statement."""
    conj = e.get("conj", [False, False])
    spec = lambda k, cj: tp.OperandSpec(tp.LayoutSpec(e[k]["dims"], e[k]["strides"]),
                                        tp.Op.CONJUGATE if cj else tp.Op.IDENTITY)
    dot = tp.DotGeneral(e["lc"], e["rc"], e["lb"], e["rb"])
    prob = tp.Problem.from_dot_general(dt, spec("a", conj[0]), spec("b", conj[1]), spec("c", False), dot)
    return tp.Plan(prob)


def name(e):
    return "%s/%s" % (e["corpus"], e.get("name", "?"))


SMALL = [e for e in ENTRIES if macs(e) <= FULL_MACS and max(T(e[k]["dims"], e[k]["strides"], []).storage()
                                                           for k in "abc") <= FULL_STORAGE]
LARGE = [e for e in ENTRIES if e not in SMALL]


def test_corpus_is_complete():
    assert len(ENTRIES) == 169 and len(SMALL) + len(LARGE) == len(ENTRIES)


@pytest.mark.gpu
@pytest.mark.parametrize("e", SMALL, ids=name)
def test_corpus_entry_vs_oracle(gpu, e):
    dt = DT[e["dtype"]]
    a, b, d = tensors(e)
    av, bv = rand(dt, max(1, a.storage()), 11), rand(dt, max(1, b.storage()), 12)
    dv = np.zeros(max(1, d.storage()), NP[dt])
    got, want, plan = run(dt, a, av, b, bv, d, dv, plan=dot_general_plan(e, dt))
    check(dt, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("e", LARGE, ids=name)
def test_corpus_entry_sampled(gpu, e):
    import torch

    dt = DT[e["dtype"]]
    a, b, d = tensors(e)
    plan = dot_general_plan(e, dt)
    dev = torch.device("cuda:0")
    tdt = {tp.DType.F32: torch.float32, tp.DType.F64: torch.float64, tp.DType.C32: torch.complex64,
           tp.DType.C64: torch.complex128}[dt]
    g = torch.Generator(device=dev).manual_seed(5)
    ta = torch.randn(max(1, a.storage()), dtype=tdt, device=dev, generator=g)
    tb = torch.randn(max(1, b.storage()), dtype=tdt, device=dev, generator=g)
    td = torch.zeros(max(1, d.storage()), dtype=tdt, device=dev)
    es = dt.size
    plan.execute_raw(1.0, ta.data_ptr() + a.origin() * es, tb.data_ptr() + b.origin() * es, 0.0, None,
                     td.data_ptr() + d.origin() * es)
    torch.cuda.synchronize()
    A, B, D = ta.cpu().numpy(), tb.cpu().numpy(), td.cpu().numpy()
    cd = np.complex128 if dt.is_complex else np.float64
    A, B = A.astype(cd), B.astype(cd)
    if a.conj:
        A = np.conj(A)
    if b.conj:
        B = np.conj(B)
    # The contracted index space as A and B offsets, shared by every sample.
    kdims = [e["a"]["dims"][i] for i in e["lc"]]
    grid = np.indices(kdims).reshape(len(kdims), -1) if kdims else np.zeros((0, 1), np.int64)
    ak = sum(grid[p] * e["a"]["strides"][i] for p, i in enumerate(e["lc"])) if kdims else np.zeros(1, np.int64)
    bk = sum(grid[p] * e["b"]["strides"][i] for p, i in enumerate(e["rc"])) if kdims else np.zeros(1, np.int64)
    rng = np.random.default_rng(6)
    la, lb, ld = labels(e)
    got, want = [], []
    for _ in range(24):
        idx = {lab: int(rng.integers(0, ext)) for lab, ext in zip(ld, e["c"]["dims"])}
        aoff = a.origin() + sum(idx.get(lab, 0) * s for lab, s, x in zip(la, e["a"]["strides"], range(len(la)))
                                if x not in e["lc"])
        boff = b.origin() + sum(idx.get(lab, 0) * s for lab, s, x in zip(lb, e["b"]["strides"], range(len(lb)))
                                if x not in e["rc"])
        doff = d.origin() + sum(idx[lab] * s for lab, s in zip(ld, e["c"]["strides"]))
        want.append(np.dot(A[aoff + ak], B[boff + bk]))
        got.append(D[doff])
    got, want = np.array(got), np.array(want)
    err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
    assert err <= tol(dt) * max(1.0, math.sqrt(len(ak)) / 32), "sampled rel err %.3e" % err
