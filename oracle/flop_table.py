"""Fit the floating-point operation counts of the oracle's inviscid and
viscous right-hand sides on periodic n^3 boxes to exact cubic polynomials in
n, and print the table used by opencfd-sc_amd/ocfd/perfmodel.py.

The counting build (liboracle_count.so) counts every add/sub, mul, div, sqrt
and pow the restatement executes; since the restatement follows the
reference's expressions one for one, these are the reference's own counts.
For the hybrid scheme the count depends on the face types, so it is fitted
for all-linear (UD7L), all-WENO7 and all-WENO5 faces.

    python oracle/flop_table.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "opencfd-sc_amd"))

import pyoracle  # noqa: E402
from ocfd import cases  # noqa: E402

SIZES = (8, 10, 12, 14)

# (key, random_case kwargs, which routine)
VARIANTS = [
    ("WENO5/SW/nonchar", dict(scheme="WENO5", split=1, char=0), "inv"),
    ("WENO5/SW/char", dict(scheme="WENO5", split=1, char=1), "inv"),
    ("WENO5/LLF/char", dict(scheme="WENO5", split=2, char=1), "inv"),
    ("WENO7/LLF/char", dict(scheme="WENO7", split=2, char=1), "inv"),
    ("WENO7/SW/nonchar", dict(scheme="WENO7", split=1, char=0), "inv"),
    ("UD7L/SW/nonchar", dict(scheme="UD7L", split=1, char=0), "inv"),
    ("HYBRID-linear/LLF/char", dict(scheme="HYBRID", split=2, char=1, hybrid=(1e30, 1e30)), "inv"),
    ("HYBRID-WENO7/LLF/char", dict(scheme="HYBRID", split=2, char=1, hybrid=(-1.0, 1e30)), "inv"),
    ("HYBRID-WENO5/LLF/char", dict(scheme="HYBRID", split=2, char=1, hybrid=(-1.0, -1.0)), "inv"),
    ("HYBRID-linear/SW/nonchar", dict(scheme="HYBRID", split=1, char=0, hybrid=(1e30, 1e30)), "inv"),
    ("viscous/CD6", dict(scheme="WENO5", split=1, char=0), "visc"),
]


def count(n, kw, which):
    case = cases.random_case(shape=(n, n, n), seed=1, **kw)
    b = pyoracle.setup_oracle(case, counting=True)
    b.exchange_primitives()
    if kw.get("scheme") == "HYBRID":
        b.comput_rhybrid()
    if which == "visc":
        fn = lambda blk: blk.du_viscous()  # noqa: E731
    elif kw.get("char"):
        fn = lambda blk: blk.du_inviscous_character()  # noqa: E731
    else:
        fn = lambda blk: blk.du_inviscous()  # noqa: E731
    return sum(pyoracle.count_ops(b, fn))


def fit(kw, which):
    ns = np.array(SIZES, dtype=float)
    ys = np.array([count(n, kw, which) for n in SIZES], dtype=float)
    A = np.vander(ns, 4)
    c = np.linalg.solve(A, ys)
    return [float(round(x, 6)) for x in c]


def table():
    return {key: fit(kw, which) for key, kw, which in VARIANTS}


if __name__ == "__main__":
    t = table()
    print("INVISCID_FIT = {")
    for k, v in t.items():
        print("    %r: %r," % (k, v))
    print("}")
