"""Writes tests/golden/thread_gauss_subset.json: Gauss-Legendre nodes and weights
held by the reference as tables (sources/HHG/thread_gauss.hpp:21-311,
gauss::container<N> = the N/2 positive nodes and their weights, 20 significant
digits, generated there from Boost in long double).  These numeric tables are
the only reference-held vectors on the hot path.  The script reads the header as
text and keeps only the numbers of a few orders; run it in a container that has
the reference mounted:  python tests/golden/make_gl_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/sources/HHG/thread_gauss.hpp"
ORDERS = [4, 8, 64, 128, 256, 512, 768, 960]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "thread_gauss_subset.json")


def main():
    text = open(SRC).read()
    out = {"source": "sources/HHG/thread_gauss.hpp", "orders": {}}
    for n in ORDERS:
        start = text.index("struct container<%d> {" % n)
        nxt = text.find("template<>", start + 1)
        body = text[start:nxt if nxt > 0 else len(text)]
        arrays = re.findall(r"=\s*\{(.*?)\}", body, re.S)
        nums = [[s.strip().rstrip("L") for s in a.split(",") if s.strip()] for a in arrays]
        assert len(nums) == 2 and len(nums[0]) == n // 2 == len(nums[1]), (n, [len(x) for x in nums])
        out["orders"][str(n)] = {"abscissa": nums[0], "weights": nums[1]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
