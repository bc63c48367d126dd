"""The conformance case list: each case mirrors one of the reference's TAPP
conformance tests (crates/tprims-capi/tests/tapp_conformance.rs, cited per
case) with the same extents, strides, labels and `seq` seeds, so the inputs
are the reference suite's own. A case is a function of the dtype returning
the keyword arguments of harness.run().
"""
import numpy as np

import oracle
from harness import NP, T, values

I, J, K, R, H = ord("i"), ord("j"), ord("k"), ord("r"), ord("h")


def scalar(dtype, re, im=0.0):
    return complex(re, im) if dtype.is_complex else re


def zeros(dtype, t):
    return np.zeros(max(1, t.storage()), dtype=NP[dtype])


def plain(dtype, a, sa, b, sb, d, **kw):
    out = dict(a=a, av=values(dtype, a, sa), b=b, bv=values(dtype, b, sb), d=d, dv=zeros(dtype, d))
    out.update(kw)
    return out


def mixed_layout_gemm(dtype):
    # tapp_conformance.rs:34-58 (a column-major A, row-major B, column-major D)
    a = T([4, 3], [1, 4], [I, K])
    b = T([3, 5], [5, 1], [K, J])
    d = T([4, 5], [1, 4], [I, J])
    return plain(dtype, a, 1, b, 2, d)


def general_strides(dtype):
    # tapp_conformance.rs:153-166
    a = T([2, 4, 3], [12, 1, 4], [R, I, K])
    b = T([3, 5], [5, 1], [K, J])
    d = T([4, 5, 2], [5, 1, 25], [I, J, R])
    return plain(dtype, a, 11, b, 12, d)


def hadamard(dtype):
    # tapp_conformance.rs:170-181
    a = T([2, 4, 3], [1, 2, 8], [H, I, K])
    b = T([2, 3, 5], [1, 2, 6], [H, K, J])
    d = T([2, 4, 5], [1, 2, 8], [H, I, J])
    return plain(dtype, a, 21, b, 22, d)


def diagonal(dtype):
    # tapp_conformance.rs:185-196
    a = T([4, 4], [1, 4], [I, I])
    b = T([4, 5], [1, 4], [I, J])
    d = T([5], [1], [J])
    return plain(dtype, a, 31, b, 32, d)


def diagonal3(dtype):
    # tapp_conformance.rs:200-211
    a = T([4, 4, 4], [1, 4, 16], [I, I, I])
    b = T([4, 5], [1, 4], [I, J])
    d = T([5], [1], [J])
    return plain(dtype, a, 33, b, 34, d)


def output_diagonal(dtype):
    # tapp_conformance.rs:221-235: off-diagonal D must stay untouched
    a = T([4, 3], [1, 4], [I, K])
    b = T([3], [1], [K])
    d = T([4, 4], [1, 4], [I, I])
    c = plain(dtype, a, 41, b, 42, d)
    c["dv"] = values(dtype, d, 43)
    return c


def reduction_in_a(dtype):
    # tapp_conformance.rs:239-251
    a = T([4, 3, 2], [1, 4, 12], [I, K, R])
    b = T([3, 5], [1, 3], [K, J])
    d = T([4, 5], [1, 4], [I, J])
    return plain(dtype, a, 51, b, 52, d)


def reduction_in_b(dtype):
    # tapp_conformance.rs:255-266
    a = T([4, 3], [1, 4], [I, K])
    b = T([3, 5, 2], [1, 3, 15], [K, J, R])
    d = T([4, 5], [1, 4], [I, J])
    return plain(dtype, a, 61, b, 62, d)


def full_reduction(dtype):
    # tapp_conformance.rs:271-283: rank-0 output
    a = T([4, 3], [1, 4], [I, J])
    b = T([4, 3], [3, 1], [I, J])
    d = T([], [], [])
    return plain(dtype, a, 71, b, 72, d)


def any_int64_labels(dtype):
    # tapp_conformance.rs:288-301
    P, Q, S = 0, -1, -(2 ** 63) + 1
    a = T([5, 3], [1, 5], [P, S])
    b = T([3, 4], [4, 1], [S, Q])
    d = T([5, 4], [1, 5], [P, Q])
    return plain(dtype, a, 81, b, 82, d)


def conjugation(mask, kk=None):
    # tapp_conformance.rs:306-322 (and :361-400 with k = 600 for kk)
    def case(dtype):
        if kk is None:
            a = T([4, 3], [1, 4], [I, K], conj=bool(mask & 1))
            b = T([3, 5], [1, 3], [K, J], conj=bool(mask & 2))
            d = T([4, 5], [1, 4], [I, J], conj=bool(mask & 8))
            seeds, al, be = (91, 92, 93), (0.5, -0.25), (-2.0, 0.75)
        else:
            a = T([3, kk], [1, 3], [I, K], conj=bool(mask & 1))
            b = T([kk, 2], [1, kk], [K, J], conj=bool(mask & 2))
            d = T([3, 2], [1, 3], [I, J], conj=bool(mask & 8))
            seeds, al, be = (101, 102, 103), (1.25, -0.5), (0.75, 0.5)
        c = T(d.extents, d.strides, d.labels, conj=bool(mask & 4))
        out = plain(dtype, a, seeds[0], b, seeds[1], d, c=c, cv=values(dtype, d, seeds[2]),
                    alpha=scalar(dtype, *al), beta=scalar(dtype, *be))
        return out
    return case


def alpha_beta(dtype):
    # tapp_conformance.rs:418-432
    a = T([4, 3], [1, 4], [I, K])
    b = T([3, 5], [1, 3], [K, J])
    d = T([4, 5], [1, 4], [I, J])
    return plain(dtype, a, 121, b, 122, d, c=d, cv=values(dtype, d, 123), alpha=scalar(dtype, 2.0, -3.0),
                 beta=scalar(dtype, -1.0, 0.5))


def alpha_zero(dtype):
    # tapp_conformance.rs:435-450: D = C exactly
    a = T([4, 3], [1, 4], [I, K])
    b = T([3, 5], [1, 3], [K, J])
    d = T([4, 5], [1, 4], [I, J])
    return plain(dtype, a, 131, b, 132, d, c=d, cv=values(dtype, d, 133), alpha=0.0, beta=1.0)


def beta_zero_nan_c(kk=None):
    # tapp_conformance.rs:458-500: C is NaN and never read
    def case(dtype):
        if kk is None:
            a, b, d = T([4, 3], [1, 4], [I, K]), T([3, 5], [1, 3], [K, J]), T([4, 5], [1, 4], [I, J])
            s = (141, 142)
        else:
            a, b, d = T([3, kk], [1, 3], [I, K]), T([kk, 2], [1, kk], [K, J]), T([3, 2], [1, 3], [I, J])
            s = (151, 152)
        cv = np.full(d.storage(), np.nan, dtype=NP[dtype])
        if dtype.is_complex:
            cv = cv + 1j * np.nan
        return plain(dtype, a, s[0], b, s[1], d, c=d, cv=cv.astype(NP[dtype]), beta=0.0)
    return case


def accumulate_in_place(dtype):
    # tapp_conformance.rs:505-520: C is D's own pointer
    a, b, d = T([4, 3], [1, 4], [I, K]), T([3, 5], [1, 3], [K, J]), T([4, 5], [1, 4], [I, J])
    out = plain(dtype, a, 161, b, 162, d, beta=1.0, in_place=True)
    out["dv"] = values(dtype, d, 163)
    return out


def negative_strides(dtype):
    # correctness.rs:584-657 class: reversed axes in every operand
    a = T([5, 4], [-1, -5], [I, K])
    b = T([4, 6], [6, -1], [K, J])
    d = T([5, 6], [-6, 1], [I, J])
    return plain(dtype, a, 171, b, 172, d)


def empty_k(dtype):
    # an empty contraction: D = beta * C
    a, b, d = T([4, 0], [1, 4], [I, K]), T([0, 5], [1, 1], [K, J]), T([4, 5], [1, 4], [I, J])
    out = plain(dtype, a, 181, b, 182, d, c=d, cv=values(dtype, d, 183), beta=scalar(dtype, 0.5, 0.25))
    out["av"] = np.zeros(1, dtype=NP[dtype])
    out["bv"] = np.zeros(1, dtype=NP[dtype])
    return out


def padded_view(dtype):
    # A non-unit-stride view: every 2nd element of a 2x-padded buffer, with conj(A)
    # and a batch index (the shape class of the complex benchmark, small).
    j, i, l, n, k = 6, 20, 5, 3, 18
    sa = [2, 2 * j, 2 * j * i, 2 * j * i * l]
    a = T([j, i, l, n], sa, [J, I, K, H], conj=True)
    b = T([l, k, j, n], [1, l, l * k, l * k * j], [K, ord("m"), J, H])
    d = T([i, k, n], [1, i, i * k], [I, ord("m"), H])
    return plain(dtype, a, 191, b, 192, d, c=d, cv=values(dtype, d, 193), beta=scalar(dtype, 0.5))


def large_k(dtype):
    # crosses many K slabs and several tiles in M and N; odd extents everywhere
    a = T([70, 300], [1, 70], [I, K])
    b = T([300, 45], [45, 1], [K, J])
    d = T([70, 45], [1, 70], [I, J])
    return plain(dtype, a, 201, b, 202, d, alpha=scalar(dtype, 0.75, 0.5))


def c_layout_differs(dtype):
    # tapp_conformance.rs:545-561: C row-major, D column-major over the same indices
    a, b = T([4, 3], [1, 4], [I, K]), T([3, 5], [1, 3], [K, J])
    d, c = T([4, 5], [1, 4], [I, J]), T([4, 5], [5, 1], [I, J])
    return plain(dtype, a, 181, b, 182, d, c=c, cv=values(dtype, c, 183), beta=scalar(dtype, 1.5, -0.5))


def empty_k_conj(mask):
    # tapp_conformance.rs:821-842: k = 0 honours op_C (mask bit 0) and op_D (bit 1)
    def case(dtype):
        a, b = T([4, 0], [1, 4], [I, K]), T([0, 5], [1, 0], [K, J])
        d = T([4, 5], [1, 4], [I, J], conj=bool(mask & 2))
        c = T([4, 5], [1, 4], [I, J], conj=bool(mask & 1))
        out = plain(dtype, a, 1, b, 1, d, c=c, cv=values(dtype, d, 271), alpha=scalar(dtype, 2.0, 1.0),
                    beta=scalar(dtype, -1.5, 0.25))
        out["av"] = np.zeros(1, dtype=NP[dtype])
        out["bv"] = np.zeros(1, dtype=NP[dtype])
        return out
    return case


CASES = {
    "mixed_layout_gemm": mixed_layout_gemm,
    "general_strides": general_strides,
    "hadamard": hadamard,
    "diagonal": diagonal,
    "diagonal3": diagonal3,
    "output_diagonal": output_diagonal,
    "reduction_in_a": reduction_in_a,
    "reduction_in_b": reduction_in_b,
    "full_reduction": full_reduction,
    "any_int64_labels": any_int64_labels,
    "alpha_beta": alpha_beta,
    "alpha_zero": alpha_zero,
    "beta_zero_nan_c": beta_zero_nan_c(),
    "beta_zero_nan_c_deep_k": beta_zero_nan_c(600),
    "accumulate_in_place": accumulate_in_place,
    "negative_strides": negative_strides,
    "empty_k": empty_k,
    "padded_view": padded_view,
    "large_k": large_k,
    "c_layout_differs": c_layout_differs,
}
for _m in range(4):
    CASES["empty_k_conj_mask_%d" % _m] = empty_k_conj(_m)
for _m in range(16):
    CASES["conj_mask_%d" % _m] = conjugation(_m)
    CASES["conj_mask_%d_deep_k" % _m] = conjugation(_m, 600)
