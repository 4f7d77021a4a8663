"""Writes tests/golden/nearfield_degree.json: near-field quadrature degrees of single-layer
operators as the reference computes them.

The reference's rule (Bembel/src/LinearOperator/LinearOperatorBase.hpp:69-89, operator order
OO = -1 for the single-layer operators) is

    dlog  = -(L * log 2)  if dist * 2^L < 1  else  log(dist)
    alpha = 2 - OO + 2 P
    num   = (alpha + P) * L * log 2 - (2 - P + OO) * dlog
    den   = (L + 2) * log 2 + dlog
    degree = (int)(0.5 * num / den)            (truncation toward zero)

and evaluateBilinearForm (DuffyTrick/evaluateBilinearForm.hpp:31-38) then takes
max(degree, P - OO + 1).  For touching pairs the real value is exactly (2 + P) L / 2, so it
sits on an integer whenever P L is even, and the truncated IEEE result depends on how the
double expression rounds (P = 1, L = 6 gives 8.999999999999998 -> 8, not 9).

Each entry records two independent evaluations:
  * "ieee": the expression above in IEEE double arithmetic, in the reference's operand order
    (Python floats, no fused multiply-add, libm log), truncated as C++ truncates;
  * "exact": the same expression in 50-digit arithmetic (mpmath).
Where the exact value is farther than 1e-9 from an integer both must truncate alike (the
script asserts it); where it is closer, the IEEE result is the reference's and "boundary" is
set.  The fixture is data only: inputs (P, L, dist as a hex float) and expected degrees.
Run: python tests/golden/make_degree_fixture.py
"""
import json
import math
import os
import random
import sys

import mpmath

OO = -1


def degree_ieee(P, dist, L, oo=OO):
    """IEEE double evaluation in the reference's operand order; returns (value, truncated)."""
    dlog = -(L * math.log(2.)) if dist * (1 << L) < 1 else math.log(dist)
    alpha = 2 - oo + 2 * P
    num = (alpha + P) * L * math.log(2.) - (2 - P + oo) * dlog
    den = (L + 2) * math.log(2.) + dlog
    v = 0.5 * num / den
    return v, int(v)  # int() truncates toward zero, as the C++ conversion does


def degree_exact(P, dist, L, oo=OO):
    mpmath.mp.dps = 50
    ln2 = mpmath.log(2)
    d = mpmath.mpf(dist)
    dlog = -L * ln2 if d * 2 ** L < 1 else mpmath.log(d)
    alpha = 2 - oo + 2 * P
    num = (alpha + P) * L * ln2 - (2 - P + oo) * dlog
    den = (L + 2) * ln2 + dlog
    return mpmath.mpf("0.5") * num / den


def entry(P, L, dist):
    v, t = degree_ieee(P, dist, L)
    ex = degree_exact(P, dist, L)
    near = abs(ex - mpmath.nint(ex)) < mpmath.mpf("1e-9")
    if not near:
        trunc = mpmath.floor(ex) if ex >= 0 else mpmath.ceil(ex)
        assert t == int(trunc), (P, L, dist, v, ex)
    return {"P": P, "L": L, "dist": float(dist).hex(), "ieee": v.hex(), "exact": mpmath.nstr(ex, 20),
            "boundary": bool(near), "degree": t, "with_farfield": max(t, P - OO + 1)}


def main():
    rng = random.Random(20240611)
    out = []
    for P in range(0, 4):
        for L in range(0, 9):
            out.append(entry(P, L, 0.0))  # touching pairs
            h = 2.0 ** -L
            # below, at and above the switch dist * 2^L = 1, and up to the regular-pair bound
            for f in (0.25, 0.5, 0.999999, 1.0, 1.000001, 1.5, 2.0, 3.0, 4.0, 7.5):
                out.append(entry(P, L, f * h))
            for _ in range(6):
                out.append(entry(P, L, h * 10 ** rng.uniform(-1, 1)))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nearfield_degree.json")
    with open(dst, "w") as f:
        json.dump({"operator_order": OO, "entries": out}, f, indent=0)
    print("entries:", len(out), "boundary cases:", sum(e["boundary"] for e in out))


if __name__ == "__main__":
    sys.exit(main())
