"""CPU tests that pin the oracle: hand-checkable known answers from the
reference's own tests, an independent numpy restatement over every
conformance case, and properties the mathematics fixes (linearity,
conjugation symmetry, the reduction identity) that catch a dropped term, a
wrong sign or a transposed operand."""
import numpy as np
import pytest

import oracle
from cases import CASES
from harness import DTYPES, NP, T


def test_kat_single_complex_product():
    # crates/tprims-capi/tests/tapp_conformance.rs:88-111
    a = oracle.Operand(np.array([1 + 2j]), [1, 1], [1, 1], [0, 2])
    b = oracle.Operand(np.array([3 + 4j]), [1, 1], [1, 1], [2, 1])
    d = oracle.Operand(np.zeros(1, complex), [1, 1], [1, 1], [0, 1])
    assert oracle.contract(1.0, a, b, 0.0, None, d)[0] == -5 + 10j


def test_kat_small_real_product():
    # tapp_conformance.rs:114-140: [19, 43, 22, 50]
    a = oracle.Operand([1, 3, 2, 4], [2, 2], [1, 2], [0, 2])
    b = oracle.Operand([5, 7, 6, 8], [2, 2], [1, 2], [2, 1])
    d = oracle.Operand(np.zeros(4), [2, 2], [1, 2], [0, 1])
    np.testing.assert_array_equal(oracle.contract(1.0, a, b, 0.0, None, d), [19, 43, 22, 50])


def test_kat_c_consumer_examples():
    # tests/c/standard_consumer.c: [22, 28, 49, 64]; tests/c/abi_test.c: [35, 44, 44, 56]
    a = oracle.Operand(np.arange(1, 7.0), [2, 3], [1, 2], [0, 2])
    b = oracle.Operand(np.arange(1, 7.0), [3, 2], [1, 3], [2, 1])
    d = oracle.Operand(np.full(4, -1.0), [2, 2], [1, 2], [0, 1])
    np.testing.assert_array_equal(oracle.contract(1.0, a, b, 0.0, None, d), [22, 28, 49, 64])
    ra = oracle.Operand([1, 3, 5, 2, 4, 6.0], [2, 3], [3, 1], [0, 2])
    bb = oracle.Operand([1, 3, 5, 2, 4, 6.0], [3, 2], [1, 3], [2, 1])
    np.testing.assert_array_equal(oracle.contract(1.0, ra, bb, 0.0, None, d), [35, 44, 44, 56])


def test_kat_diagonal_is_trace():
    # D[] = sum_i A[i,i] B[i]: with B = 1 it is the trace, a closed form.
    A = np.arange(16.0)
    a = oracle.Operand(A, [4, 4], [1, 4], [0, 0])
    b = oracle.Operand(np.ones(4), [4], [1], [0])
    d = oracle.Operand(np.zeros(1), [], [], [])
    assert oracle.contract(1.0, a, b, 0.0, None, d)[0] == np.trace(A.reshape(4, 4))


def test_beta_zero_never_reads_c():
    a = oracle.Operand(np.ones(4), [2, 2], [1, 2], [0, 2])
    d = oracle.Operand(np.zeros(4), [2, 2], [1, 2], [0, 1])
    c = oracle.Operand(np.full(4, np.nan), [2, 2], [1, 2], [0, 1])
    assert np.all(np.isfinite(oracle.contract(1.0, a, a, 0.0, c, d)))


def test_output_diagonal_leaves_the_rest():
    a = oracle.Operand(np.arange(1, 5.0), [4], [1], [0])
    b = oracle.Operand(np.array([2.0]), [1], [1], [1])
    d0 = np.full(16, 7.0)
    d = oracle.Operand(d0, [4, 4], [1, 4], [0, 0])
    out = oracle.contract(1.0, a, b, 0.0, None, d).reshape(4, 4)
    assert np.all(np.diag(out) == 2 * np.arange(1, 5.0))
    assert np.all(out[~np.eye(4, dtype=bool)] == 7.0)


def test_seq_is_the_reference_sequence_class():
    # multiples of 1/2048 in [-1, 1), deterministic per seed, complex uses two draws
    x = oracle.seq(1000, 11)
    assert np.all(x >= -1) and np.all(x < 1)
    assert np.all(x * 2048 == np.round(x * 2048))
    np.testing.assert_array_equal(x, oracle.seq(1000, 11))
    z = oracle.seq(10, 11, cplx=True)
    np.testing.assert_array_equal(z.real, oracle.seq(20, 11)[0::2])
    np.testing.assert_array_equal(z.imag, oracle.seq(20, 11)[1::2])


def _gf2_apply(cols, v):
    """A 64x64 matrix over GF(2), given by the images of the 64 unit vectors,
    applied to the 64-bit vector v."""
    out = 0
    j = 0
    while v:
        if v & 1:
            out ^= cols[j]
        v >>= 1
        j += 1
    return out


def _gf2_pow(cols, n):
    result = [1 << j for j in range(64)]
    base = list(cols)
    while n:
        if n & 1:
            result = [_gf2_apply(base, c) for c in result]
        base = [_gf2_apply(base, c) for c in base]
        n >>= 1
    return result


def test_seq_generator_has_the_full_period_and_the_published_constants():
    """seq's state update (crates/tprims-capi/tests/common/mod.rs:86-91) is a
    linear map T over GF(2)^64. Its shift triple is pinned by the period: T
    has order exactly 2^64 - 1 (T^(2^64-1) = I and T^((2^64-1)/p) != I for
    every prime p dividing 2^64 - 1), which holds for the published triple
    (13, 7, 17) and fails for a wrong shift or order. The state is one
    xorshift step per value drawn from the scrambled seed; the scrambling
    constant is floor(2^64 / golden ratio), odd."""
    mask = (1 << 64) - 1
    # T's columns: run one step of the generator's update on each unit vector.
    def step(s):
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        return s
    cols = [step(1 << j) for j in range(64)]
    # The step actually used by oracle.seq: its first value from a seed equals
    # the map applied to the scrambled seed.
    state = ((5 * 0x9E3779B97F4A7C15) & mask) | 1
    assert oracle.seq(1, 5)[0] == (_gf2_apply(cols, state) >> 52) / 2048.0 - 1.0
    n = mask
    ident = [1 << j for j in range(64)]
    assert _gf2_pow(cols, n) == ident
    for p in (3, 5, 17, 257, 641, 65537, 6700417):
        assert n % p == 0
        assert _gf2_pow(cols, n // p) != ident, p
    import math
    assert 3 * 5 * 17 * 257 * 641 * 65537 * 6700417 == n
    golden = (math.isqrt(5 << 128) - (1 << 64)) // 2  # floor(2^64 (sqrt(5) - 1) / 2)
    assert golden == 0x9E3779B97F4A7C15 and golden & 1
    # Values: the state's top 12 bits scaled into [-1, 1), every multiple of
    # 1/2048 reached over one long run.
    x = oracle.seq(200000, 3)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert np.all(x * 2048 == np.round(x * 2048))
    assert len(np.unique(x)) == 4096


def test_max_rel_err_definition():
    # fixtures.rs:81-87: normalised by max(1, max |want|)
    assert oracle.max_rel_err(np.array([1.5, 0.0]), np.array([1.0, 0.0])) == 0.5
    assert oracle.max_rel_err(np.array([11.0]), np.array([10.0])) == 0.1
    assert oracle.max_rel_err(np.array([3 + 4j]), np.array([0j])) == 5.0


@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_numpy_restatement(name, dtype):
    kw = CASES[name](dtype)
    a, b, d = kw["a"], kw["b"], kw["d"]
    c = d if kw.get("in_place") else kw.get("c")
    cv = kw["dv"] if kw.get("in_place") else kw.get("cv")
    alpha, beta = kw.get("alpha", 1.0), kw.get("beta", 0.0)
    if cv is not None and np.any(np.isnan(np.asarray(cv, dtype=complex))):
        pytest.skip("NaN C is covered by the never-read test")
    args = (alpha, a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]), beta,
            c.oracle_op(cv) if cv is not None else None, d.oracle_op(kw["dv"]))
    got = oracle.contract(*args, cplx=dtype.is_complex)
    want = oracle.einsum_reference(*args)
    np.testing.assert_allclose(got.astype(complex), want, rtol=1e-12, atol=1e-12)


def test_oracle_linearity_and_conjugation():
    rng = np.random.default_rng(3)
    A = rng.standard_normal(12) + 1j * rng.standard_normal(12)
    B = rng.standard_normal(15) + 1j * rng.standard_normal(15)
    B2 = rng.standard_normal(15) + 1j * rng.standard_normal(15)
    mk = lambda x, dims, labels, conj=False: oracle.Operand(x, dims, [1, dims[0]], labels, conj=conj)
    d = oracle.Operand(np.zeros(20, complex), [4, 5], [1, 4], [0, 1])
    f = lambda Bv, conj=False: oracle.contract(1.0, mk(A, [4, 3], [0, 2], conj), mk(Bv, [3, 5], [2, 1]), 0.0, None, d)
    np.testing.assert_allclose(f(B + 2 * B2), f(B) + 2 * f(B2), atol=1e-12)
    # conj(A) B = conj(A conj(B))
    dconj = oracle.Operand(np.zeros(20, complex), [4, 5], [1, 4], [0, 1], conj=True)
    lhs = oracle.contract(1.0, mk(A, [4, 3], [0, 2], True), mk(B, [3, 5], [2, 1]), 0.0, None, d)
    rhs = oracle.contract(1.0, mk(A, [4, 3], [0, 2]), mk(B, [3, 5], [2, 1], True), 0.0, None, dconj)
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)
    # matrix product closed form
    np.testing.assert_allclose(f(B).reshape(5, 4).T, A.reshape(3, 4).T @ B.reshape(5, 3).T, atol=1e-12)


def test_oracle_plants_are_caught():
    """A wrong sign in the complex cross term or a transposed operand must
    move the result: checked by planting them in the inputs of the oracle."""
    a = oracle.Operand(np.array([1 + 2j, 3 - 1j]), [2], [1], [0])
    b = oracle.Operand(np.array([2 - 1j, 1 + 1j]), [2], [1], [0])
    d = oracle.Operand(np.zeros(1, complex), [], [], [])
    good = oracle.contract(1.0, a, b, 0.0, None, d)[0]
    assert good == (1 + 2j) * (2 - 1j) + (3 - 1j) * (1 + 1j)
    flipped = oracle.contract(1.0, a, oracle.Operand(np.conj(b.data), [2], [1], [0]), 0.0, None, d)[0]
    assert flipped != good


def test_port_packs_panels_in_the_reference_layout():
    """The CPU port's packer against the reference's packing examples
    (crates/tprims-kernel/src/pack/pack.rs:304-330, 331-338, 364-369,
    393-400, 402-425): Real slivers with zero fill, Planar (the real parts of
    a k step, then its imaginary parts), conjugation, an irregular row
    scatter, and the planar size of two k."""
    from oracle import packed
    real = packed.pack_panel(np.arange(8, dtype=np.float64), [0, 1, 2], [0, 4], 2)
    assert real.tolist() == [0.0, 1.0, 4.0, 5.0, 2.0, 0.0, 6.0, 0.0]
    cplx = np.array([complex(i, 10 + i) for i in range(4)])
    assert packed.pack_panel(cplx, [0, 1], [0, 2], 2).tolist() == [0.0, 1.0, 10.0, 11.0, 2.0, 3.0, 12.0, 13.0]
    two = np.array([1 + 2j, 3 + 4j])
    assert packed.pack_panel(two, [0, 1], [0], 2, conj=True).tolist() == [1.0, 3.0, -2.0, -4.0]
    assert packed.pack_panel(np.arange(6, dtype=np.float64), [0, 3, 1], [0], 3).tolist() == [0.0, 3.0, 1.0]
    assert len(packed.pack_panel(two, [0, 1], [0, 0], 2)) == 8
    assert packed.pack_panel(two.astype(np.complex64), [0, 1], [0], 2).tolist() == [1.0, 3.0, 2.0, 4.0]


@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_packed_port_matches_the_oracle(name, dtype):
    """The CPU baseline (the port of the reference's packed driver) gives the
    oracle's result on every conformance case and dtype before it is timed:
    f64 within 1e-11, f32 within 2e-4 (crates/tprims-capi/tests/common/mod.rs:57-63)."""
    from oracle import packed
    kw = CASES[name](dtype)
    a, b, d = kw["a"], kw["b"], kw["d"]
    c = d if kw.get("in_place") else kw.get("c")
    cv = kw["dv"] if kw.get("in_place") else kw.get("cv")
    alpha, beta = kw.get("alpha", 1.0), kw.get("beta", 0.0)
    if cv is not None and np.any(np.isnan(np.asarray(cv, dtype=complex))):
        pytest.skip("NaN C is covered by the never-read test")
    want = oracle.contract(alpha, a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]), beta,
                           c.oracle_op(cv) if cv is not None else None, d.oracle_op(kw["dv"]), cplx=dtype.is_complex)
    dv = np.array(kw["dv"], copy=True)
    dop = d.oracle_op(dv)
    cop = dop if kw.get("in_place") else (c.oracle_op(np.array(cv, copy=True)) if cv is not None else None)
    packed.contract(alpha, a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]), beta, cop, dop, threads=2)
    tol = 2e-4 if dtype.name in ("F32", "C32") else 1e-11
    assert oracle.max_rel_err(dop.data, want) <= tol, (name, oracle.max_rel_err(dop.data, want))
