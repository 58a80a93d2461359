"""The batched form without device allocation (spg_batch_run_noalloc) gives
bit for bit the results of the allocating batched entry points
(spg_batch_create / add / commit / launch) on the same tasks: two calls in a
row, the second with fewer and smaller tasks, so the second reuses the
scratch array the first filled.  Also: a work list larger than the scratch
array is refused on the host, before any launch."""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import _lib
from spaghetty_amd.batched import run_noalloc

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def mix(count, seed, lo4=3, hi4=21, lo6=2, hi6=8):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if rng.random() < 0.7:
            dims = tuple(int(v) for v in rng.integers(lo4, hi4, size=4))
        else:
            dims = tuple(int(v) for v in rng.integers(lo6, hi6, size=6))
        perm = tuple(int(v) + 1 for v in rng.permutation(len(dims)))
        f = float(rng.choice([1.0, -1.0, 0.5, -0.5, 37.0]))
        beta = 1.0 if rng.random() < 0.5 else 0.0
        if i % 13 == 4:
            beta = 0.7
        out.append((dims, perm, f, beta))
    return out


def run_both(torch, dev, tiles, seed, flags):
    sizes = [int(np.prod(d)) for d, _, _, _ in tiles]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    x = torch.rand(int(offs[-1]), dtype=torch.float64, device=dev, generator=g) * 2 - 1
    y0 = torch.rand(int(offs[-1]), dtype=torch.float64, device=dev, generator=g) * 2 - 1
    ya, yb = y0.clone(), y0.clone()
    b = S.Batch()
    for (d, p, f, beta), a, e in zip(tiles, offs[:-1], offs[1:]):
        b.add(x[a:e], ya[a:e], d, p, f, beta)
    b.set_flags(flags)
    b.commit()
    b.launch()
    torch.cuda.synchronize()
    b.close()
    run_noalloc([(x[a:e], yb[a:e], d, p, f, beta)
                 for (d, p, f, beta), a, e in zip(tiles, offs[:-1], offs[1:])], flags=flags)
    torch.cuda.synchronize()
    assert not torch.equal(ya.view(torch.int64), y0.view(torch.int64))
    return ya, yb


@pytest.mark.parametrize("flags", [0, 2, 64])
def test_noalloc_equals_batched_two_calls(dev, flags):
    import torch
    log(f"case: noalloc vs batched, flags {flags}, call 1: 400 tiles")
    ya, yb = run_both(torch, dev, mix(400, seed=3), seed=11, flags=flags)
    assert torch.equal(ya.view(torch.int64), yb.view(torch.int64))
    log(f"case: noalloc vs batched, flags {flags}, call 2: 60 smaller tiles")
    ya, yb = run_both(torch, dev, mix(60, seed=4, hi4=9, hi6=4), seed=12, flags=flags)
    assert torch.equal(ya.view(torch.int64), yb.view(torch.int64))


def test_noalloc_full_size_cc_tiles(dev):
    # 4D tiles of the coupled-cluster mix at their stated sizes
    import torch
    from spaghetty_amd.workloads import cc_tile_mix
    log("case: noalloc vs batched, 48 full-size 4D cc tiles")
    tiles = [(t.dims, t.perm, t.factor, t.beta) for t in cc_tile_mix(100, seed=9) if len(t.dims) == 4][:48]
    ya, yb = run_both(torch, dev, tiles, seed=13, flags=0)
    assert torch.equal(ya.view(torch.int64), yb.view(torch.int64))


def test_noalloc_rejects_work_list_beyond_scratch(dev):
    import torch
    log("case: noalloc oversized work list")
    n = _lib.SPG_NOALLOC_SCRATCH_BYTES // 1000 + 16  # more task records than the scratch holds
    x = torch.zeros(4 * n, dtype=torch.float64, device=dev)
    y = torch.zeros(4 * n, dtype=torch.float64, device=dev)
    tasks = [(x[4 * i:4 * i + 4], y[4 * i:4 * i + 4], (2, 2), (2, 1), 1.0, 0.0) for i in range(n)]
    with pytest.raises(ValueError, match="scratch"):
        run_noalloc(tasks)
    torch.cuda.synchronize()
    assert torch.count_nonzero(y).item() == 0
