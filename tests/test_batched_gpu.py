"""GPU parity of the batched form: many uneven 4D/6D sorts in one persistent
launch (and as a replayed hipGraph) against the oracle and against one
per-call launch per tile, bit for bit.

The tiles follow the coupled-cluster mix of the batched workload (random
permutation per tile, factor from {+-1, +-0.5}, accumulate on half of them)
at extents the oracle checks in well under a second.  A batch has no
reductions, so every tile's result must equal its per-call result exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def small_mix(count, seed):
    """Uneven tiles: 70% 4D with extents in [3,20], 30% 6D with extents in
    [2,7]; random permutation, factor in {+-1, +-0.5, 37}, accumulate on
    half; a few five-case tiles with g = -2."""
    rng = np.random.default_rng(seed)
    tiles = []
    for i in range(count):
        if rng.random() < 0.7:
            dims = tuple(int(v) for v in rng.integers(3, 21, size=4))
        else:
            dims = tuple(int(v) for v in rng.integers(2, 8, size=6))
        perm = tuple(int(v) + 1 for v in rng.permutation(len(dims)))
        f = float(rng.choice([1.0, -1.0, 0.5, -0.5, 37.0]))
        beta = 1.0 if rng.random() < 0.5 else 0.0
        if i % 17 == 5:
            beta = -2.0
        tiles.append((dims, perm, f, beta))
    return tiles


def oracle_out(x, y0, dims, perm, f, beta):
    mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
    return O.sort_n(x, y0.copy(), dims, perm, f, beta, mode)


def _setup(torch, dev, tiles, seed):
    total = sum(int(np.prod(d)) for d, _, _, _ in tiles)
    xs = O.uniform(total, seed)
    ys = O.uniform(total, seed + 1)
    xd = torch.from_numpy(xs).to(dev)
    y0d = torch.from_numpy(ys).to(dev)
    offs = []
    off = 0
    for d, _, _, _ in tiles:
        n = int(np.prod(d))
        offs.append((off, n))
        off += n
    return xs, ys, xd, y0d, offs


def _check_all(got, xs, ys, tiles, offs, what):
    for (d, p, f, b), (o, n) in zip(tiles, offs):
        want = oracle_out(xs[o:o + n], ys[o:o + n], d, p, f, b)
        assert bits_equal(got[o:o + n], want), f"{what}: dims {d} perm {p} f {f} beta {b}"


@pytest.mark.parametrize("flags", [None, 0, 1, 2, 3, 32, 34, 64, 66, 67, 96, 98])
def test_batched_equals_oracle_and_per_call(dev, flags):
    import torch
    log(f"case: batched small mix, 300 tiles, flags {flags}")
    tiles = small_mix(300, seed=21)
    xs, ys, xd, y0d, offs = _setup(torch, dev, tiles, seed=40)
    yb = y0d.clone()
    b = S.Batch()
    for (d, p, f, beta), (o, n) in zip(tiles, offs):
        b.add(xd[o:o + n], yb[o:o + n], d, p, f, beta)
    if flags is not None:
        b.set_flags(flags)
    b.commit()
    b.launch()
    torch.cuda.synchronize()
    got = yb.cpu().numpy()
    _check_all(got, xs, ys, tiles, offs, "batched")
    # one launch per tile gives the same bits
    yc = y0d.clone()
    for (d, p, f, beta), (o, n) in zip(tiles, offs):
        S.sort(xd[o:o + n], yc[o:o + n], d, p, f, beta)
    torch.cuda.synchronize()
    assert bits_equal(yc.cpu().numpy(), got)
    b.close()


def test_batched_graph_replays(dev):
    import torch
    log("case: batched graph, 200 tiles, 4 replays")
    tiles = small_mix(200, seed=5)
    xs, ys, xd, y0d, offs = _setup(torch, dev, tiles, seed=60)
    yb = y0d.clone()
    b = S.Batch()
    for (d, p, f, beta), (o, n) in zip(tiles, offs):
        b.add(xd[o:o + n], yb[o:o + n], d, p, f, beta)
    b.commit()
    stream = torch.cuda.Stream(device=dev)
    for rep in range(4):
        # eager work between replays: restore the outputs, then replay
        yb.copy_(y0d)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            b.graph_launch(stream)
        stream.synchronize()
        got = yb.cpu().numpy()
        _check_all(got, xs, ys, tiles, offs, f"graph replay {rep}")
    b.close()


def test_batched_cc_tiles_4d_full_size(dev):
    # 64 tiles of the coupled-cluster 4D distribution (occupied U[8,20],
    # virtual U[16,40]) at their real sizes
    import torch
    from spaghetty_amd.workloads import cc_tile_mix
    log("case: batched cc 4D tiles at full size")
    tiles = [(t.dims, t.perm, t.factor, t.beta)
             for t in cc_tile_mix(200, seed=3) if len(t.dims) == 4][:64]
    xs, ys, xd, y0d, offs = _setup(torch, dev, tiles, seed=80)
    yb = y0d.clone()
    b = S.Batch()
    for (d, p, f, beta), (o, n) in zip(tiles, offs):
        b.add(xd[o:o + n], yb[o:o + n], d, p, f, beta)
    b.commit()
    b.launch()
    torch.cuda.synchronize()
    _check_all(yb.cpu().numpy(), xs, ys, tiles, offs, "batched cc")
    b.close()


def test_batched_rejects_overlapping_outputs(dev):
    import torch
    log("case: batched alias check")
    x = torch.zeros(1000, dtype=torch.float64, device=dev)
    y = torch.zeros(1000, dtype=torch.float64, device=dev)
    b = S.Batch()
    b.add(x[:500], y[:500], (5, 10, 10), (3, 2, 1), 1.0, 0.0)
    b.add(x[500:], y[250:750], (5, 10, 10), (3, 2, 1), 1.0, 0.0)
    with pytest.raises(ValueError):
        b.commit()
    b.close()
    b = S.Batch()
    b.add(x[:500], y[:500], (5, 10, 10), (3, 2, 1), 1.0, 0.0)
    b.add(y[:500], x[500:], (5, 10, 10), (3, 2, 1), 1.0, 0.0)
    with pytest.raises(ValueError):
        b.commit()
    b.close()
