"""Serving form on the GPU: every request's result in the continuous-batching
mix is bit-identical to the same request run alone, and accumulations into
one tensor are applied in arrival order."""
from __future__ import annotations

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd.serving import Server, poisson_arrivals
from spaghetty_amd.workloads import cc_tile_mix

pytestmark = pytest.mark.gpu


def small_mix(n, seed):
    # the coupled-cluster mix, with 6D tiles shrunk to keep the test small
    out = []
    for t in cc_tile_mix(n, seed=seed):
        dims = t.dims if len(t.dims) == 4 else tuple(max(2, d // 3) for d in t.dims)
        out.append((dims, t.perm, t.factor, t.beta))
    return out


def test_results_alone_equal_results_in_mix(dev):
    import torch
    reqs = small_mix(300, seed=11)
    sizes = [int(np.prod(d)) for d, _, _, _ in reqs]
    total = sum(sizes)
    inp = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    init = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    mix = init.clone()
    alone = init.clone()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    # alone: one call at a time
    for i, (dims, perm, f, beta) in enumerate(reqs):
        a, b = offs[i], offs[i + 1]
        S.sort(inp[a:b], alone[a:b], dims, perm, f, beta)
        torch.cuda.synchronize()
    srv = Server(1 << 30, slots=3, max_batch=64)
    t = poisson_arrivals(len(reqs), rate_per_us=0.2, seed=5)
    for i, (dims, perm, f, beta) in enumerate(reqs):
        a, b = offs[i], offs[i + 1]
        srv.submit(inp[a:b], mix[a:b], dims, perm, f, beta, arrival_us=float(t[i]))
    srv.start()
    srv.wait(120)
    st = srv.stats()
    srv.close()
    print("serving stats", st, flush=True)
    assert st["completed"] == len(reqs) and st["error"] == 0
    assert st["batches"] > 1
    assert torch.equal(mix.view(torch.int64), alone.view(torch.int64))


def test_accumulations_keep_arrival_order(dev):
    import torch
    dims, perm = (12, 10, 14, 9), (3, 1, 4, 2)
    n = int(np.prod(dims))
    xs = [torch.rand(n, dtype=torch.float64, device=dev) for _ in range(8)]
    y0 = torch.rand(n, dtype=torch.float64, device=dev)
    want = y0.clone()
    for i, x in enumerate(xs):
        S.sort(x, want, dims, perm, 0.3 + i, 1.0)
    got = y0.clone()
    srv = Server(1 << 30, slots=4, max_batch=16)
    for i, x in enumerate(xs):
        srv.submit(x, got, dims, perm, 0.3 + i, 1.0, arrival_us=0.0)
    srv.start()
    srv.wait(60)
    r = srv.results()
    srv.close()
    assert len(set(r["batch"].tolist())) == len(xs)
    assert torch.equal(got.view(torch.int64), want.view(torch.int64))


def test_server_allocated_outputs_and_budget(dev):
    import torch
    x = torch.rand(1 << 20, dtype=torch.float64, device=dev)
    srv = Server(64 << 20, slots=2, max_batch=8)
    for i in range(64):
        srv.submit(x, None, (64, 64, 16, 16), (4, 3, 2, 1), 1.0, 0.0, arrival_us=float(i))
    srv.start()
    srv.wait(60)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 64 and st["peak_in_flight"] <= 64 << 20
