"""CPU tests of the TCCG corpus restatement (tests/tccg.py) against the
reference's own tests of it (benchmarks/benchmarks/tcbench/corpus.rs:486-523),
and of the planner's reduced view of every case."""
import pytest

import tccg
import tprims_amd as tp


def test_corpus_is_the_full_49_case_set():
    # corpus.rs:490-494: 19 + 3 + 8 + 18 + 1, none colliding under normalisation
    cases = tccg.corpus()
    assert len(cases) == 49
    groups = [g for _, g, _, _, _ in cases]
    assert [groups.count(g) for g in ("ccsd", "ao2mo", "intensli", "ccsd_t", "transC")] == [19, 3, 8, 18, 1]


def test_sorted_order_entries_are_all_in_the_corpus():
    # corpus.rs:496-507: _sortedTCs re-labels existing cases
    norms = {tccg.normalize(c[0]) for c in tccg.corpus()}
    order = tccg.sorted_order()
    assert len(order) == 24
    for s in order:
        assert tccg.normalize(s) in norms, s


def test_normalisation_examples():
    # the renaming that makes "_sortedTCs" entries match corpus cases
    assert tccg.normalize("ij-ik-kj") == tccg.normalize("ab-ac-cb") == "ab-ac-cb"
    assert tccg.normalize("abcijk-ijma-mkbc") == "abcdef-dega-gfbc"  # a _sortedTCs entry
    assert tccg.normalize("ij-ik-kj") != tccg.normalize("ij-ki-kj")


def test_sizing_matches_tccg_rule():
    # corpus.rs:509-521: "ij-ik-kj" at 200 MiB: avg = 5120; i and k are
    # stride-1 somewhere and round up to 5136, j rounds to 5120
    case = [c for c in tccg.corpus() if c[0] == "ij-ik-kj"][0]
    s = tccg.Sized(case, 200 * 1024 * 1024)
    assert (s.extent["i"], s.extent["k"], s.extent["j"]) == (5136, 5136, 5120)


def test_mnk_is_consistent_with_macs():
    # corpus.rs:523-530
    for c in tccg.corpus():
        s = tccg.Sized(c, 8 * 1024 * 1024)
        m, n, k = s.mnk()
        assert m * n * k == s.macs(), c[0]


@pytest.mark.parametrize("dtype", [tp.DType.F64, tp.DType.C32], ids=lambda d: d.name)
def test_planner_reduces_every_case_to_its_matrix_view(dtype):
    """Every case at the 64 MiB benchmark size plans (no GPU needed) with
    M N K = the case's multiply-adds, no batch index, and the reference's
    flop count (2 per real, 8 per complex multiply-add)."""
    for c in tccg.corpus():
        s = tccg.Sized(c, 64 * 1024 * 1024)
        plan = plan_of(s, dtype)
        r = plan.report()
        assert r["m"] * r["n"] * r["k"] == s.macs() and r["batch"] == 1, c[0]
        assert (r["m"], r["n"], r["k"]) == s.mnk(), c[0]
        assert r["flops"] == (8 if dtype.is_complex else 2) * s.macs()


def plan_of(s, dtype, config=None):
    """The plan of a sized case: C[c] = A[a] B[b], column-major."""
    spec = lambda x: tp.OperandSpec(tp.LayoutSpec(*s.layout(x)))
    lab = lambda x: [ord(ch) for ch in x]
    prob = tp.Problem.from_labels(dtype, spec(s.a), spec(s.b), tp.CSpec.separate(spec(s.c)), spec(s.c),
                                  tp.Labels(lab(s.a), lab(s.b), lab(s.c), lab(s.c)))
    return tp.Plan(prob, config)
