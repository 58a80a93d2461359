"""16-byte pairs over odd extents (plans with pair8) on the GPU.

With odd extents a pair of neighbouring elements may start at an odd
element offset (its 16-byte access is then 8-byte aligned), and a partial
tile may hold one element of a pair only; such tiles move the two elements
separately.  Every case below forces tile shapes with even tile extents
(Tx, Ty) over odd array extents, so the plans take pairs (asserted), and
compares the result bit for bit with the oracle (old_sort.f:33-95, 97-234,
512-574, 576-713 restated): sort, accumulate and the five-case form, the
kernel's access variants, tiles above 4,096 elements and views that start
at an odd element.
"""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import TileParams
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PIPE, NT, NT_LOAD, LDS1 = 2, 1, 4, 32

# (dims slowest -> fastest, permutation, forced tile shape)
CASES = [
    ((15, 17, 19, 21), (4, 3, 2, 1), dict(nx=1, tx=16, ny=1, ty=6)),
    ((15, 17, 19, 21), (4, 3, 2, 1), dict(nx=2, tx=2, ny=2, ty=2)),
    ((15, 13, 11, 9), (3, 4, 1, 2), dict(nx=1, tx=6, ny=1, ty=10)),
    ((9, 7, 15, 13, 11, 5), (4, 6, 2, 5, 1, 3), dict(nx=2, tx=2, ny=1, ty=8)),
    ((7, 9, 11, 13, 15, 17), (6, 5, 4, 3, 2, 1), dict(nx=2, tx=4, ny=2, ty=2)),
    ((1001, 999), (2, 1), dict(nx=1, tx=64, ny=1, ty=64)),
]


def log(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def oracle_sort(x, y0, dims, perm, f, beta):
    mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
    return O.sort_n(x, y0.copy(), dims, perm, f, beta, mode)


def run(x, y0, dims, perm, f, beta, dev, params, shift=0):
    import torch
    n = x.size
    xd = torch.from_numpy(np.concatenate([np.zeros(shift), x])).to(dev)[shift:]
    yd = torch.from_numpy(np.concatenate([np.zeros(shift), y0])).to(dev)[shift:]
    assert xd.numel() == n
    S.sort(xd, yd, dims, perm, f, beta, params)
    return yd.cpu().numpy()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_odd_extent_pairs_exact(dev, case):
    dims, perm, shape = CASES[case]
    n = int(np.prod(dims))
    x, y0 = O.uniform(n, 300 + case), O.uniform(n, 400 + case)
    for beta in (0.0, 1.0, -2.5):
        want = oracle_sort(x, y0, dims, perm, -0.5, beta)
        for fl in (0, PIPE, NT | PIPE, NT_LOAD, NT_LOAD | PIPE, LDS1, LDS1 | PIPE):
            for block in (0, 64):
                p = TileParams.default(kind=0, flags=fl, block=block, **shape)
                d = S.plan(dims, perm, beta, p)
                assert d["vec"] == 2 and d["pair8"] == 1, d
                log("case:", dims, perm, shape, "beta", beta, "flags", fl, "block", d["block"],
                    "slots", d["slots"])
                assert bits_equal(run(x, y0, dims, perm, -0.5, beta, dev, p), want)
        # views starting at an odd element: pairs are 8-byte aligned either way
        p = TileParams.default(kind=0, flags=PIPE, **shape)
        log("case:", dims, perm, shape, "beta", beta, "shifted views")
        assert bits_equal(run(x, y0, dims, perm, -0.5, beta, dev, p, shift=1), want)


def test_odd_extent_pairs_big_tiles(dev):
    # tiles above 4,096 elements (8 pairs per thread, up to 768 threads)
    dims, perm = (45, 47, 49, 51), (4, 3, 2, 1)
    n = int(np.prod(dims))
    x, y0 = O.uniform(n, 500), O.uniform(n, 501)
    for beta in (0.0, 1.0):
        want = oracle_sort(x, y0, dims, perm, 0.5, beta)
        # a pipelined accumulate holds tiles of up to 8,192 elements
        wide = dict(nx=2, tx=2, ny=2, ty=2) if beta == 0.0 else dict(nx=2, tx=4, ny=1, ty=38)
        for fl, shape in ((0, dict(nx=2, tx=2, ny=1, ty=44)), (PIPE, wide),
                          (LDS1 | PIPE, dict(nx=1, tx=50, ny=2, ty=2)), (NT | PIPE, dict()),
                          (0, dict())):
            p = TileParams.default(kind=0, flags=fl, max_tile=12288, walk_mode=2, walk_x=4, **shape)
            d = S.plan(dims, perm, beta, p)
            log("case: big", dims, perm, beta, fl, shape, "V", d["V"], "block", d["block"], "pair8",
                d["pair8"])
            if shape:
                assert d["vec"] == 2 and d["pair8"] == 1 and d["V"] > 4096, d
            assert bits_equal(run(x, y0, dims, perm, 0.5, beta, dev, p), want)
        # a tile above 4,096 elements over odd extents on views that start at
        # an odd element keeps its pairs (no re-plan to single elements)
        p = TileParams.default(kind=0, flags=PIPE, max_tile=12288, nx=2, tx=2, ny=1, ty=44)
        log("case: big tile, shifted views", dims, perm, beta)
        assert bits_equal(run(x, y0, dims, perm, 0.5, beta, dev, p, shift=1), want)


def test_odd_extent_default_plans(dev):
    # the library's own plans on odd-extent problems, tuned or not
    rng = np.random.default_rng(12)
    for i in range(40):
        r = int(rng.choice([4, 6]))
        lo, hi = (9, 41) if r == 4 else (5, 16)
        dims = tuple(int(v) for v in rng.integers(lo, hi, size=r))
        perm = tuple(int(v) + 1 for v in rng.permutation(r))
        beta = float(rng.choice([0.0, 1.0, 0.25]))
        n = int(np.prod(dims))
        x, y0 = O.uniform(n, 600 + i), O.uniform(n, 700 + i)
        d = S.plan(dims, perm, beta)
        log("case: default", dims, perm, beta, "vec", d.get("vec"), "pair8", d.get("pair8"))
        want = oracle_sort(x, y0, dims, perm, 1.5, beta)
        assert bits_equal(run(x, y0, dims, perm, 1.5, beta, dev, None), want)


def test_odd_extent_pairs_random_parameters(dev):
    # random ranks, odd-heavy extents and forced tile parameters (walk
    # orders, chunked walks, one or two LDS buffers, pipelining): every plan
    # that takes pairs over odd extents, against the oracle
    rng = np.random.default_rng(2718)
    done = 0
    while done < 60:
        r = int(rng.integers(2, 9))
        dims = tuple(int(v) for v in rng.integers(1, 10 if r > 5 else 24, size=r))
        n = int(np.prod(dims))
        if n > 300000:
            continue
        perm = tuple(int(v) + 1 for v in rng.permutation(r))
        beta = float(rng.choice([0.0, 1.0, -1.5]))
        kw = dict(kind=0, nx=int(rng.integers(1, 4)), tx=int(rng.integers(1, 9)), ny=int(rng.integers(1, 4)),
                  ty=int(rng.integers(1, 9)), flags=int(rng.choice([0, 2, 32, 34, 6])),
                  walk_mode=int(rng.integers(0, 4)), walk_x=int(rng.choice([0, 2, 3])),
                  walk_chunk=int(rng.choice([0, 2, 3])))
        p = TileParams.default(**kw)
        try:
            d = S.plan(dims, perm, beta, p)
        except ValueError:
            continue
        if not d.get("pair8"):
            continue
        x, y0 = O.uniform(n, 800 + done), O.uniform(n, 900 + done)
        log("case: random", dims, perm, beta, kw, "slots", d["slots"])
        want = oracle_sort(x, y0, dims, perm, 0.75, beta)
        assert bits_equal(run(x, y0, dims, perm, 0.75, beta, dev, p, shift=done % 2), want)
        done += 1
