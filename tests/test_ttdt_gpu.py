"""GPU tests of the contraction around the sort (spg_ttdt): the three sorts
on the library's kernels and the DGEMM on rocBLAS, against the oracle's
restatement of branches/old/python/macaroni/ccsd_t2_8_skeleton.F:162-183.

With integer-valued inputs every partial sum is exact in FP64, so results
are compared bit for bit, as the reference's own skeleton compares its
number_array() runs (compare_arrays, exact).  Uniform(-1,1) inputs are
compared with a stated tolerance: the DGEMM's summation order is rocBLAS's,
so |got - want| <= 8 * K * 2^-53 * sum_k |a_k b_k| bounds the difference.
"""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def _t2_8_args(d):
    d1, d2, d3, d4, d5, d6 = d
    return dict(adims=(d5, d6, d1, d2), aperm=(4, 3, 2, 1), bdims=(d3, d4, d5, d6),
                bperm=(2, 1, 4, 3), cdims=(d4, d3, d2, d1), cperm=(2, 1, 4, 3),
                m=d1 * d2, n=d3 * d4, k=d5 * d6)


@pytest.mark.parametrize("d", [(3, 4, 5, 2, 6, 3), (30, 30, 30, 30, 30, 30), (16, 20, 12, 24, 14, 10)])
def test_ccsd_t2_8_integer_data_exact(dev, d):
    log("case: ccsd_t2_8 integer", d)
    a = _t2_8_args(d)
    rng = np.random.default_rng(sum(d))
    nt, nv, nr = a["m"] * a["k"], a["k"] * a["n"], a["m"] * a["n"]
    t = rng.integers(-8, 9, size=nt).astype(np.float64)
    v = rng.integers(-8, 9, size=nv).astype(np.float64)
    r0 = rng.integers(-8, 9, size=nr).astype(np.float64)
    for beta, alpha in ((0.0, 1.0), (1.0, 1.0), (1.0, -2.0)):
        rd = to_dev(r0, dev)
        S.ttdt(to_dev(t, dev), a["adims"], a["aperm"], to_dev(v, dev), a["bdims"], a["bperm"],
               rd, a["cdims"], a["cperm"], a["m"], a["n"], a["k"], alpha, beta)
        want = O.ttdt(t, a["adims"], a["aperm"], v, a["bdims"], a["bperm"], a["cdims"], a["cperm"],
                      a["m"], a["n"], a["k"], alpha, beta, r0)
        assert bits_equal(rd.cpu().numpy(), want), (d, alpha, beta)
    # the reference-shaped entry: its last tce_sort_4 overwrites r
    # (ccsd_t2_8_skeleton.F:178); accumulate=True adds into r instead
    for acc in (False, True):
        rd = to_dev(r0, dev)
        S.ccsd_t2_8(to_dev(t, dev), to_dev(v, dev), rd, *d, accumulate=acc)
        want = O.ttdt(t, a["adims"], a["aperm"], v, a["bdims"], a["bperm"], a["cdims"], a["cperm"],
                      a["m"], a["n"], a["k"], 1.0, 1.0 if acc else 0.0, r0)
        assert bits_equal(rd.cpu().numpy(), want), acc


def test_ccsd_t2_8_uniform_data_within_bound(dev):
    d = (20, 20, 20, 20, 20, 20)
    log("case: ccsd_t2_8 uniform", d)
    a = _t2_8_args(d)
    t = O.uniform(a["m"] * a["k"], 1)
    v = O.uniform(a["k"] * a["n"], 2)
    r0 = O.uniform(a["m"] * a["n"], 3)
    rd = to_dev(r0, dev)
    S.ccsd_t2_8(to_dev(t, dev), to_dev(v, dev), rd, *d, accumulate=True)
    want = O.ttdt(t, a["adims"], a["aperm"], v, a["bdims"], a["bperm"], a["cdims"], a["cperm"],
                  a["m"], a["n"], a["k"], 1.0, 1.0, r0)
    bound = O.ttdt(np.abs(t), a["adims"], a["aperm"], np.abs(v), a["bdims"], a["bperm"],
                   a["cdims"], a["cperm"], a["m"], a["n"], a["k"])
    err = np.abs(rd.cpu().numpy() - want)
    tol = 8 * a["k"] * 2.0 ** -53 * bound + 4 * 2.0 ** -53 * np.abs(want)
    assert np.all(err <= tol), float(np.max(err / np.maximum(tol, 1e-300)))


def test_ttdt_edges_and_errors(dev):
    import torch
    log("case: ttdt edges")
    # empty contraction index: C' = 0, so C = beta*C
    c0 = O.uniform(12, 4)
    cd = to_dev(c0, dev)
    z = torch.zeros(1, dtype=torch.float64, device=dev)
    S.ttdt(z, (3, 0), (2, 1), z, (0, 4), (2, 1), cd, (4, 3), (2, 1), 3, 4, 0, 1.0, 1.0)
    assert bits_equal(cd.cpu().numpy(), c0 + 0.0)
    # dims that do not multiply to the matrix sizes
    x = torch.zeros(64, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        S.ttdt(x, (4, 4), (2, 1), x[:16], (4, 4), (2, 1), x[32:48], (4, 4), (2, 1), 4, 4, 5)
    # a bad permutation
    with pytest.raises(ValueError):
        S.ttdt(x[:16], (4, 4), (2, 2), x[16:32], (4, 4), (2, 1), x[32:48], (4, 4), (2, 1), 4, 4, 4)


def test_ttdt_calls_on_two_streams_share_the_work_buffer(dev):
    # two contractions enqueued back to back on different streams use the
    # library's one work buffer; the second must not overwrite A', B' or C'
    # while the first still reads them
    import torch
    d = (30, 30, 30, 30, 30, 30)
    log("case: ttdt on two streams", d)
    a = _t2_8_args(d)
    rng = np.random.default_rng(7)
    probs = []
    for _ in range(2):
        t = rng.integers(-8, 9, size=a["m"] * a["k"]).astype(np.float64)
        v = rng.integers(-8, 9, size=a["k"] * a["n"]).astype(np.float64)
        r0 = rng.integers(-8, 9, size=a["m"] * a["n"]).astype(np.float64)
        probs.append((t, v, r0, to_dev(t, dev), to_dev(v, dev), to_dev(r0, dev)))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev) for _ in range(2)]
    for (t, v, r0, td, vd, rd), st in zip(probs, streams):
        with torch.cuda.stream(st):
            S.ccsd_t2_8(td, vd, rd, *d, accumulate=True)
    for st in streams:
        st.synchronize()
    for t, v, r0, td, vd, rd in probs:
        want = O.ttdt(t, a["adims"], a["aperm"], v, a["bdims"], a["bperm"], a["cdims"], a["cperm"],
                      a["m"], a["n"], a["k"], 1.0, 1.0, r0)
        assert bits_equal(rd.cpu().numpy(), want)
