"""Extracts the TCCG benchmark corpus of the reference's tcbench
(benchmarks/benchmarks/tcbench/corpus.rs: the contraction specifications
"c-a-b" of its CCSD, AO2MO, InTensLi, CCSD(T) and transpose-C groups, and
the 24-entry presentation order SORTED_ORDER) into tccg_corpus.json beside
this script: the specification strings only, as data. The corpus()
de-duplication and the sizing rule are restated in tests/tccg.py. Run from
the repository root in a checkout that has the reference at /root/reference."""
import json
import os
import re

SRC = "/root/reference/benchmarks/benchmarks/tcbench/corpus.rs"
GROUPS = [("ccsd", "CCSD"), ("ao2mo", "AO2MO"), ("intensli", "INTENSLI"), ("ccsd_t", "CCSD_T"),
          ("transC", "TRANS_C")]


def strings_of(text, const):
    m = re.search(r"const %s: &\[&str\] = &\[(.*?)\];" % const, text, re.S)
    if not m:
        raise SystemExit("constant %s not found" % const)
    return re.findall(r'"([a-z]+-[a-z]+-[a-z]+)"', m.group(1))


def main():
    text = open(SRC).read()
    groups = {g: strings_of(text, c) for g, c in GROUPS}
    sorted_order = strings_of(text.replace("pub const SORTED_ORDER", "const SORTED_ORDER"), "SORTED_ORDER")
    out = {"groups": [[g, groups[g]] for g, _ in GROUPS], "sorted_order": sorted_order}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tccg_corpus.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote %d specifications (+%d in the presentation order) to %s"
          % (sum(len(v) for v in groups.values()), len(sorted_order), dst))


if __name__ == "__main__":
    main()
