"""The batched form at the coupled-cluster configuration's stated sizes, and
its waves.

* Waves run in order through the same buffers: later waves read and
  accumulate into what earlier waves wrote (checked against the oracle
  applied wave after wave).
* Stated-size 6D tiles (occupied extents in [8,20], virtual in [16,40]; up to
  20^3 x 40^3 = 512M elements, 4 GB per array) in one batch, with their own
  per-call launches (the default) and all inside the persistent launch
  (SPG_BATCH_INLINE).  One tile's plans have tile offsets beyond a buffer
  resource's 2 GB range and take the classic kernel on both paths.  Arrays
  this large are checked bit for bit against the oracle's closed form on
  2^20 seeded output positions plus the first and last 1024, with the
  inputs gathered on the device.
* The serving form with full-size 6D requests: each result equals the
  request run alone.
"""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import _lib
from spaghetty_amd.workloads import Tile, cc_tile_mix
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def test_waves_run_in_order_through_shared_buffers(dev):
    import torch
    log("case: 3 waves over one input and one output region")
    rng = np.random.default_rng(4)
    n_region = 60_000
    x = O.uniform(n_region, 70)
    y0 = O.uniform(n_region, 71)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y0.copy()).to(dev)
    want = y0.copy()
    b = S.Batch()
    for w in range(3):
        if w:
            b.next_wave()
        off = 0
        while True:
            dims = tuple(int(v) for v in rng.integers(3, 12, size=4))
            n = int(np.prod(dims))
            if off + n > n_region:
                break
            perm = tuple(int(v) + 1 for v in rng.permutation(4))
            f = float(rng.choice([1.0, -0.5, 37.0]))
            beta = float(rng.choice([0.0, 1.0, 0.7]))
            b.add(xd[off:off + n], yd[off:off + n], dims, perm, f, beta)
            mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
            want[off:off + n] = O.sort_n(x[off:off + n], want[off:off + n].copy(), dims, perm, f, beta, mode)
            off += n
    b.commit()
    assert b.info()["waves"] == 3
    b.launch()
    torch.cuda.synchronize()
    assert bits_equal(yd.cpu().numpy(), want)
    b.close()


def _stated_6d_tiles():
    # the first eight 6D tiles of the coupled-cluster mix, plus two of the
    # largest stated shape whose plans take the classic kernel
    tiles = [t for t in cc_tile_mix(4096, seed=7) if len(t.dims) == 6][:8]
    tiles.append(Tile((20, 20, 20, 40, 40, 40), (2, 3, 4, 5, 6, 1), -0.5, 0.0))
    tiles.append(Tile((20, 20, 20, 40, 40, 40), (2, 3, 4, 5, 6, 1), 1.0, 1.0))
    return tiles


def _sample_positions(n, seed, k=1 << 20):
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([np.arange(min(1024, n)), np.arange(max(0, n - 1024), n),
                                     rng.integers(0, n, size=k)]))


@pytest.mark.parametrize("flags", [0, _lib.SPG_BATCH_INLINE])
def test_stated_size_6d_tiles_batched(dev, flags):
    import torch
    tiles = _stated_6d_tiles()
    for t in tiles[-2:]:
        assert S.plan(t.dims, t.perm, t.beta)["tk"] == 0  # classic per-call plan
    sizes = [t.nelem for t in tiles]
    offs = np.concatenate([[0], np.cumsum([s + (-s) % 32 for s in sizes])]).astype(np.int64)
    total = int(offs[-1])
    log(f"case: {len(tiles)} stated-size 6D tiles, {total * 8 / 1e9:.1f} GB per side, flags {flags}")
    g = torch.Generator(device=dev)
    g.manual_seed(5 + flags)
    xd = torch.rand(total, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    yd = torch.rand(total, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    samples = []
    for i, t in enumerate(tiles):
        pos = _sample_positions(t.nelem, 100 + i)
        src = O.source_positions(t.dims, t.perm, pos)
        o = int(offs[i])
        pos_d = torch.from_numpy(pos + o).to(dev)
        x_at = xd.index_select(0, torch.from_numpy(src + o).to(dev)).cpu().numpy()
        y_at = yd.index_select(0, pos_d).cpu().numpy()
        samples.append((pos_d, x_at, y_at))
    b = S.Batch()
    for t, o in zip(tiles, offs[:-1]):
        b.add(xd[o:o + t.nelem], yd[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta)
    b.set_flags(flags)
    b.commit()
    info = b.info()
    log("batch", info)
    assert info["big_tasks"] == (0 if flags else len(tiles))
    b.launch()
    torch.cuda.synchronize()
    for t, (pos_d, x_at, y_at) in zip(tiles, samples):
        got = yd.index_select(0, pos_d).cpu().numpy()
        want = O.sampled_values(x_at, t.factor, y_at if t.beta else None)
        assert bits_equal(got, want), (t.dims, t.perm, t.factor, t.beta, flags)
    b.close()
    del xd, yd
    torch.cuda.empty_cache()


def test_serving_full_size_6d_requests_equal_alone(dev):
    import torch
    from spaghetty_amd.serving import Server, poisson_arrivals
    mix = cc_tile_mix(64, seed=23)
    reqs = [t for t in mix if len(t.dims) == 6][:5] + [t for t in mix if len(t.dims) == 4][:40]
    sizes = [t.nelem for t in reqs]
    offs = np.concatenate([[0], np.cumsum([s + (-s) % 32 for s in sizes])]).astype(np.int64)
    total = int(offs[-1])
    log(f"case: serving, {sum(len(t.dims) == 6 for t in reqs)} full-size 6D + 40 4D requests, "
        f"{total * 8 / 1e9:.1f} GB per side")
    g = torch.Generator(device=dev)
    g.manual_seed(17)
    xd = torch.rand(total, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    y0 = torch.rand(total, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    alone = y0.clone()
    for t, o in zip(reqs, offs[:-1]):
        S.sort(xd[o:o + t.nelem], alone[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta)
        torch.cuda.synchronize()
    got = y0.clone()
    del y0
    srv = Server(64 << 30, slots=4, max_batch=16)
    arr = poisson_arrivals(len(reqs), rate_per_us=0.01, seed=2)
    for t, o, a in zip(reqs, offs[:-1], arr):
        srv.submit(xd[o:o + t.nelem], got[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta, arrival_us=float(a))
    srv.start()
    srv.wait(300)
    st = srv.stats()
    srv.close()
    log("serving stats", st)
    assert st["completed"] == len(reqs) and st["error"] == 0
    assert torch.equal(got.view(torch.int64), alone.view(torch.int64))
