"""Writes tests/golden/gauss_legendre_ref.json from the reference's literal rules.

The reference hard-codes the 1..20 point Gauss-Legendre rules on [0,1] as 40-digit decimal
literals (Bembel/src/Quadrature/GaussLegendreRule.hpp:57-688).  This script reads that file as
text, parses the literals to doubles (Python's float() rounds correctly, as a C++ compiler does
for a literal) and stores them as hex floats.  The fixture is data: nodes and weights only.
Run in the development container, where the reference is mounted.
"""
import json
import os
import re
import sys

SRC = "/root/reference/Bembel/src/Quadrature/GaussLegendreRule.hpp"


def main():
    text = open(SRC).read()
    out = {}
    for m in re.finditer(r"struct GaussLegendreRule<(\d+)>\s*\{(.*?)\n\};", text, re.S):
        n = int(m.group(1))
        body = m.group(2)
        xs = re.search(r"xi_\s*=\s*\{(.*?)\};", body, re.S)
        ws = re.search(r"w_\s*=\s*\{(.*?)\};", body, re.S)
        if not xs or not ws:
            continue
        xv = [float(t) for t in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?|\d+", xs.group(1))]
        wv = [float(t) for t in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?|\d+", ws.group(1))]
        assert len(xv) == n and len(wv) == n, (n, len(xv), len(wv))
        out[str(n)] = {"x": [v.hex() for v in xv], "w": [v.hex() for v in wv]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss_legendre_ref.json")
    json.dump(out, open(dst, "w"), indent=0)
    print("rules:", sorted(int(k) for k in out))


if __name__ == "__main__":
    sys.exit(main())
