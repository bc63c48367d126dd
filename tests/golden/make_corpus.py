"""Extracts the dot_general shape corpus of the reference's benchmark suite
(benchmarks/benchmarks/tprims/corpus/*.json: operand dims and strides, the
contracted and batch axis lists, conjugation, call counts) into
corpus_dot_general.json beside this script. Only the shape data is kept; the
GPU corpus tests and the corpus replay read that file. Run from the
repository root in a checkout that has the reference at /root/reference."""
import glob
import json
import os

SRC = "/root/reference/benchmarks/benchmarks/tprims/corpus"
KEYS = ["name", "dtype", "a", "b", "c", "lc", "rc", "lb", "rb", "conj", "calls"]


def main():
    out = []
    for path in sorted(glob.glob(os.path.join(SRC, "*.json"))):
        corpus = os.path.splitext(os.path.basename(path))[0]
        with open(path) as f:
            data = json.load(f)
        for e in data["entries"]:
            if e.get("op") != "dot_general":
                continue
            rec = {"corpus": corpus}
            for k in KEYS:
                if k in e:
                    rec[k] = e[k]
            out.append(rec)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "corpus_dot_general.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote %d entries to %s" % (len(out), dst))


if __name__ == "__main__":
    main()
