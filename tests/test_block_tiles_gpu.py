"""Tiles of several blocks (plans with tw > 1) on the GPU.

Where the two fastest axes of the input are the two fastest of the output,
swapped (432165 among the (T) permutations, 2143-like 4D ones), each tile of
those two whole axes is one contiguous block on both sides, and a tile may
span several blocks along one more axis (the input's next axis, or the
output's).  Every case compares the GPU result bit for bit with the oracle
(old_sort.f:33-95, 97-234, 512-574, 576-713 restated): sort, accumulate and
the five-case form; whole and partial W chunks; every kernel form; and the
(T) permutation at its 15^6, 16^6 and 17^6 sizes.  A batch of such
problems (which plans one block per tile) is checked too.
"""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import TileParams
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NT, PIPE, NT_LOAD, CLASSIC, SCALAR, LDS1 = 1, 2, 4, 8, 16, 32
FLAGS = (0, PIPE, NT | PIPE, NT_LOAD, LDS1, LDS1 | PIPE, CLASSIC, SCALAR)

# (dims slowest -> fastest, permutation, W chunks: tw > 1 along the input's
# next axis, tw < -1 along the output's)
CASES = [
    ((5, 6, 7, 4, 3, 9), (4, 3, 2, 1, 6, 5), (2, 3, 4, -2, -5)),
    ((7, 5, 9, 6, 8, 4), (4, 3, 2, 1, 6, 5), (2, 4, 6, -3, -7)),
    ((6, 8, 11, 10), (1, 2, 4, 3), (2, 4, 11, -3)),
    ((5, 13, 12, 4), (2, 1, 4, 3), (3, -2, -5)),
    ((3, 5, 4, 6, 7), (2, 1, 4, 3, 5), (2, -2)),
    ((3, 5, 4, 6, 8), (1, 2, 4, 3, 5), (2, 3)),
    ((3, 2, 3, 4, 2, 3, 3, 4), (1, 3, 4, 2, 5, 6, 8, 7), (2, 3, -2, -3)),  # rank 8
]


def log(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def oracle_sort(x, y0, dims, perm, f, beta):
    mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
    return O.sort_n(x, y0.copy(), dims, perm, f, beta, mode)


def run(x, y0, dims, perm, f, beta, dev, params):
    import torch
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y0).to(dev)
    S.sort(xd, yd, dims, perm, f, beta, params)
    return yd.cpu().numpy()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_block_tiles_exact(dev, case):
    dims, perm, tws = CASES[case]
    n = int(np.prod(dims))
    x, y0 = O.uniform(n, 500 + case), O.uniform(n, 600 + case)
    for beta in (0.0, 1.0, -2.5):
        want = oracle_sort(x, y0, dims, perm, 0.75, beta)
        for tw in tws:
            for fl in FLAGS:
                p = TileParams.default(kind=0, flags=fl, tw=tw)
                d = S.plan(dims, perm, beta, p)
                assert d["tw"] == abs(tw), d
                log("case:", dims, perm, "tw", tw, "beta", beta, "flags", fl, "V", d["V"], "vec", d["vec"],
                    "tk", d["tk"], "block", d["block"])
                assert bits_equal(run(x, y0, dims, perm, 0.75, beta, dev, p), want)
        # the cost model's own choice
        p = TileParams.default()
        log("case:", dims, perm, "automatic", S.plan(dims, perm, beta, p)["tw"], "beta", beta)
        assert bits_equal(run(x, y0, dims, perm, 0.75, beta, dev, p), want)


@pytest.mark.parametrize("e", [15, 16, 17])
def test_triples_432165_full_size(dev, e):
    # the (T) permutation whose two fastest axes swap, at tile sizes NWChem
    # runs: tiles of many blocks, partial W chunks where e is odd
    dims, perm = (e,) * 6, (4, 3, 2, 1, 6, 5)
    n = e ** 6
    x, y0 = O.uniform(n, 700 + e), O.uniform(n, 800 + e)
    for beta in (0.0, 1.0):
        want = oracle_sort(x, y0, dims, perm, -0.5, beta)
        for tw, fl in ((0, 0), (0, PIPE), (e, 0), (-e, NT_LOAD), (7, PIPE), (-9, LDS1), (4, CLASSIC)):
            p = TileParams.default(flags=fl, tw=tw)
            try:
                d = S.plan(dims, perm, beta, p)
            except ValueError:
                # more than 4,096 elements per tile need pairs (even extents)
                assert e % 2 and e * e * abs(tw) > 4096, (e, tw, fl)
                continue
            log("case:", dims, perm, "tw", tw, "->", d["tw"], "beta", beta, "flags", fl, "V", d["V"])
            assert d["tw"] > 1
            assert bits_equal(run(x, y0, dims, perm, -0.5, beta, dev, p), want)


def test_swapped_pairs_batched(dev):
    # a batch of swapped-pair tasks and of other permutations, against the
    # oracle; the batch plans one block per tile for them (tiles of several
    # blocks measured slower there)
    import torch
    rng = np.random.default_rng(11)
    tiles = []
    for i in range(120):
        if i % 3 == 2:
            dims = tuple(int(v) for v in rng.integers(3, 21, size=4))
            perm = tuple(int(v) + 1 for v in rng.permutation(4))
        elif i % 3 == 1:
            dims = tuple(int(v) for v in rng.integers(4, 12, size=6))
            perm = (4, 3, 2, 1, 6, 5)
        else:
            dims = tuple(int(v) for v in rng.integers(5, 33, size=4))
            perm = ((1, 2, 4, 3), (2, 1, 4, 3), (3, 4, 2, 1))[i % 2]
        f = float(rng.choice([1.0, -1.0, 0.5]))
        beta = (0.0, 1.0, -2.0)[i % 3]
        tiles.append((dims, perm, f, beta))
    batch_params = TileParams.default(max_tile=2048, block=256, flags=0, tw=1)
    swapped = sum(S.plan(d, p, b, TileParams.default(kind=0, tw=1))["out_axis"][:2] == [1, 0]
                  for d, p, _, b in tiles)
    assert swapped >= 60, swapped
    assert all(S.plan(d, p, b, batch_params)["tw"] == 1 for d, p, _, b in tiles)
    total = sum(int(np.prod(d)) for d, _, _, _ in tiles)
    xs, ys = O.uniform(total, 901), O.uniform(total, 902)
    xd = torch.from_numpy(xs).to(dev)
    yb = torch.from_numpy(ys).to(dev)
    offs, off = [], 0
    for d, _, _, _ in tiles:
        offs.append((off, int(np.prod(d))))
        off += int(np.prod(d))
    for flags in (None, PIPE, LDS1 | PIPE, 64):
        log("case: batched swapped-pair mix,", len(tiles), "tiles,", swapped, "swapped pairs, flags", flags)
        yb.copy_(torch.from_numpy(ys).to(dev))
        b = S.Batch()
        for (d, p, f, beta), (o, n) in zip(tiles, offs):
            b.add(xd[o:o + n], yb[o:o + n], d, p, f, beta)
        if flags is not None:
            b.set_flags(flags)
        b.commit()
        b.launch()
        torch.cuda.synchronize()
        got = yb.cpu().numpy()
        b.close()
        for (d, p, f, beta), (o, n) in zip(tiles, offs):
            want = oracle_sort(xs[o:o + n], ys[o:o + n], d, p, f, beta)
            assert bits_equal(got[o:o + n], want), (d, p, f, beta, flags)
