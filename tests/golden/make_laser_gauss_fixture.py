"""Writes tests/golden/laser_gauss30.json: the 30-point Gauss-Legendre table on
[-1, 1] that the reference's pulse moments use (sources/HHG/Laser/gauss.hpp:14-80;
the header stores 0.5 (x + 1) and 0.5 w of these numbers).  The script reads
the header as text and keeps the 15 distinct positive nodes and weights, run in
a container that has the reference mounted:
    python tests/golden/make_laser_gauss_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/sources/HHG/Laser/gauss.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "laser_gauss30.json")


def main():
    text = open(SRC).read()
    blocks = {}
    for name in ("abscissa", "weights"):
        start = text.index("g_array %s = {" % name)
        body = text[start:text.index("};", start)]
        blocks[name] = [float(v) for v in re.findall(r"0\.5 \* \(?(-?[0-9.]+e[-+][0-9]+)", body)]
        assert len(blocks[name]) == 30, (name, len(blocks[name]))
    nodes = sorted({abs(v) for v in blocks["abscissa"]})
    weights = blocks["weights"][:15]
    assert len(nodes) == 15
    # the header lists nodes by increasing |x| with their weights in the same order
    out = {"source": "sources/HHG/Laser/gauss.hpp", "positive_nodes": [repr(v) for v in nodes],
           "weights": [repr(v) for v in weights]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
