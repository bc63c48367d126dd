"""The TCCG contraction corpus as the reference's tcbench builds it
(benchmarks/benchmarks/tcbench/corpus.rs): 49 specifications "c-a-b"
(C[c] = sum A[a] B[b], column-major, the first label of each operand
stride-1) from tests/golden/tccg_corpus.json, de-duplicated under TCCG's label
normalisation (corpus.rs:279-312), and sized by TCCG's rule (corpus.rs:397-424):
avg = (tensor_bytes / 8) ** (1 / max rank); a label that is stride-1 in some
operand is rounded up to a multiple of 24, every other label to the nearest
multiple of 4 (at least 4); 8 bytes per element whatever the dtype. Used by
the GPU corpus tests and by bench.py --corpus."""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def normalize(spec):
    """TCCG's normalisation (corpus.rs:279-312): C's labels renamed A, B, ...
    in order, then A's remaining labels; lower-cased."""
    c, a, b = (list(x) for x in spec.split("-"))
    nxt = [ord("A")]

    def rename(frm):
        to = chr(nxt[0])
        for v in (a, b, c):
            for i, x in enumerate(v):
                if x == frm:
                    v[i] = to
        nxt[0] += 1

    for ch in list(c):
        rename(ch)
    for ch in list(a):
        if ch not in c and ch.islower():
            rename(ch)
    return ("%s-%s-%s" % ("".join(c), "".join(a), "".join(b))).lower()


def _data():
    with open(os.path.join(HERE, "golden", "tccg_corpus.json")) as f:
        return json.load(f)


def corpus():
    """[(name, group, c, a, b)] in the reference's order, de-duplicated."""
    out, seen = [], set()
    for group, specs in _data()["groups"]:
        for spec in specs:
            n = normalize(spec)
            if n in seen:
                continue
            seen.add(n)
            c, a, b = spec.split("-")
            out.append((spec, group, c, a, b))
    return out


def sorted_order():
    return list(_data()["sorted_order"])


def col_major(ext):
    s, acc = [], 1
    for e in ext:
        s.append(acc)
        acc *= max(e, 1)
    return s


class Sized:
    """A corpus case at one size: extents per label and the operand layouts."""

    def __init__(self, case, tensor_bytes):
        name, group, c, a, b = case
        self.name, self.group, self.c, self.a, self.b = name, group, c, a, b
        max_ndim = max(len(a), len(b), len(c))
        avg = (tensor_bytes / 8.0) ** (1.0 / max_ndim)
        labels = []
        for s in (c, a, b):
            for ch in s:
                if ch not in labels:
                    labels.append(ch)
        stride1 = {s[0] for s in (c, a, b)}
        ext = {}
        for ch in labels:
            if ch in stride1:
                ext[ch] = int((avg + 23.0) / 24.0) * 24
            else:
                up = int((avg + 3.0) / 4.0) * 4
                down = max(int(avg / 4.0) * 4, 4)
                ext[ch] = up if abs(avg - up) < abs(avg - down) else down
        self.labels, self.extent = labels, ext

    def dims(self, s):
        return [self.extent[ch] for ch in s]

    def layout(self, s):
        d = self.dims(s)
        return d, col_major(d)

    def macs(self):
        return math.prod(self.extent.values())

    def mnk(self):
        """m, n, k of the equivalent matrix product (corpus.rs:158-177)."""
        m = n = k = 1
        for ch in self.labels:
            e = self.extent[ch]
            if ch in self.c and ch in self.a:
                m *= e
            elif ch in self.c and ch in self.b:
                n *= e
            else:
                k *= e
        return m, n, k

    def storage(self, s):
        return math.prod(self.dims(s))
