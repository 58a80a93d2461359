"""Writes the golden fixtures in tests/golden/ from the reference's own code.

The outputs come from the reference's NWChem routines old_sort_4/6 and
old_sortacc_4/6 (trunk/old_sort.f), compiled from the reference checkout by
oracle/Makefile into oracle/_ref/libold_sort.so.  That library exists only
where the reference does, so the fixtures are committed: they pin the
oracle restatement and the HIP kernels on machines without the reference.

Inputs are regenerated, not stored: either the reference testers' integer
ramp (init_4d_array / init_6d_array, trunk/tester_futil.F:139-185) or the
seeded uniform(-1,1) stream (splitmix64, oracle.uniform).  Each case stores
the SHA-256 of the output's little-endian bytes; small cases also store the
output itself (golden_small.npz).

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# The 54 permutations NWChem's (T) code calls tce_sort_6 with, plus the
# identity (background/nwchem-perturbative-triples-usage:5-58; the list
# with the identity is trunk/generator6.py:15-17).
TRIPLES = [
    "123456", "261543", "265431", "326541", "362154", "362514", "362541", "421653",
    "431652", "432165", "432651", "436215", "436251", "436521", "463215", "463251",
    "463521", "521643", "531642", "532164", "532641", "541632", "542136", "542163",
    "542631", "543162", "543216", "543261", "543621", "621543", "625431", "631542",
    "632154", "632514", "632541", "641532", "642153", "642531", "643152", "643215",
    "643251", "643521", "651432", "652143", "652413", "652431", "653142", "653214",
    "653241", "653421", "654132", "654213", "654231", "654312", "654321",
]


def random_perms_6d(count: int = 64, seed: int = 2024):
    """The seeded random 6D permutations of the 16^6 / 20^6 sweeps."""
    rng = np.random.default_rng(seed)
    return ["".join(str(int(v) + 1) for v in rng.permutation(6)) for _ in range(count)]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def make_input(spec: dict, n: int, dims) -> np.ndarray:
    if spec["kind"] == "uniform":
        return O.uniform(n, spec["seed"])
    if spec["kind"] == "kat":
        # the tester passes dim1..dimR (slowest first for old_sort) straight
        # to init_Nd_array, which takes its first argument as the fastest
        return O.kat(list(dims))
    raise ValueError(spec)


def run_case(case: dict) -> np.ndarray:
    dims, perm = case["dims"], [int(c) for c in case["perm"]]
    n = int(np.prod(dims))
    x = make_input(case["input"], n, dims)
    acc = case["beta"] != 0.0
    y = make_input(case["y0"], n, dims) if acc else np.zeros(n)
    O.ref_old_sort(x, y, dims, perm, case["factor"], acc)
    return y


def cases():
    out = []
    u = lambda s: {"kind": "uniform", "seed": s}  # noqa: E731
    z = {"kind": "zero"}
    perms4 = ["".join(p) for p in itertools.permutations("1234")]
    # reference tester KATs: integer ramp, factors {1,-1,37} (generator.py:198)
    for p in perms4:
        for f in (1.0, -1.0, 37.0):
            out.append(dict(tag="kat4", dims=[5, 6, 7, 8], perm=p, factor=f, beta=0.0,
                            input={"kind": "kat"}, y0=z, small=True))
        out.append(dict(tag="kat4acc", dims=[5, 6, 7, 8], perm=p, factor=37.0, beta=1.0,
                        input={"kind": "kat"}, y0={"kind": "kat"}, small=True))
    for p in itertools.permutations("123456"):
        p = "".join(p)
        out.append(dict(tag="kat6", dims=[3, 4, 2, 3, 2, 3], perm=p, factor=1.0, beta=0.0,
                        input={"kind": "kat"}, y0=z))
        out.append(dict(tag="uni6acc", dims=[3, 4, 2, 3, 2, 3], perm=p, factor=-0.5, beta=1.0,
                        input=u(61), y0=u(62)))
    for p in perms4:
        for f, b in ((0.5, 0.0), (-1.0, 1.0), (37.0, 1.0)):
            out.append(dict(tag="uni4", dims=[3, 5, 4, 2], perm=p, factor=f, beta=b,
                            input=u(41), y0=u(42), small=True))
    # full-size configurations (digests only)
    for p in perms4:
        out.append(dict(tag="c30", dims=[30] * 4, perm=p, factor=1.0, beta=0.0,
                        input=u(1), y0=z))
        out.append(dict(tag="c30acc", dims=[30] * 4, perm=p, factor=0.5, beta=1.0,
                        input=u(1), y0=u(2)))
        out.append(dict(tag="c20", dims=[20] * 4, perm=p, factor=-1.0, beta=1.0,
                        input=u(3), y0=u(4)))
        out.append(dict(tag="cnc", dims=[10, 30, 20, 40], perm=p, factor=0.5, beta=0.0,
                        input=u(5), y0=z))
        out.append(dict(tag="cncacc", dims=[10, 30, 20, 40], perm=p, factor=-1.0, beta=1.0,
                        input=u(5), y0=u(6)))
    for p in ("4321", "1234", "2143", "3412", "1324"):
        out.append(dict(tag="c100", dims=[100] * 4, perm=p, factor=1.0, beta=0.0,
                        input=u(7), y0=z))
    out.append(dict(tag="c100acc", dims=[100] * 4, perm="4321", factor=0.5, beta=1.0,
                    input=u(7), y0=u(8)))
    for p in TRIPLES:
        out.append(dict(tag="t10", dims=[10] * 6, perm=p, factor=1.0, beta=0.0,
                        input=u(9), y0=z))
        out.append(dict(tag="t10acc", dims=[10] * 6, perm=p, factor=-1.0, beta=1.0,
                        input=u(9), y0=u(10)))
    rp = random_perms_6d()
    for p in rp[:8]:
        out.append(dict(tag="r16", dims=[16] * 6, perm=p, factor=1.0, beta=0.0,
                        input=u(11), y0=z))
        out.append(dict(tag="r16acc", dims=[16] * 6, perm=p, factor=0.5, beta=1.0,
                        input=u(11), y0=u(12)))
    for p in rp[:3]:
        out.append(dict(tag="r20", dims=[20] * 6, perm=p, factor=-1.0, beta=0.0,
                        input=u(13), y0=z))
    out.append(dict(tag="r20acc", dims=[20] * 6, perm="654321", factor=1.0, beta=1.0,
                    input=u(13), y0=u(14)))
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libold_sort.so is missing: run `make -C oracle` "
                 "where the reference checkout exists")
    t0 = time.time()
    records, small = [], {}
    for i, c in enumerate(cases()):
        y = run_case(c)
        rec = {k: v for k, v in c.items() if k != "small"}
        rec["sha256"] = digest(y)
        records.append(rec)
        if c.get("small"):
            small[f"case{i}"] = y
    with open(os.path.join(HERE, "golden_digests.json"), "w") as f:
        json.dump({"source": "trunk/old_sort.f compiled with amdflang -O2 -ffp-contract=off "
                             "(oracle/Makefile), called by tests/golden/make_golden.py",
                   "cases": records}, f, indent=0)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small)
    print(f"{len(records)} cases, {len(small)} small arrays, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
