"""GPU parity: the HIP kernels against the oracle and the reference's golden
digests, bit for bit (uint64 views), through the C ABI.

Sizes the oracle finishes in seconds are compared element by element;
full-size configurations are compared with digests made from the
reference's own old_sort.f (tests/golden) and with size-independent
properties (inverse round trip, linearity in the factor, accumulate of zero
equals sort).
"""
from __future__ import annotations

import ctypes
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import TileParams, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden_digests.json")
PERMS4 = list(itertools.permutations([1, 2, 3, 4]))


def torch_mod():
    import torch
    return torch


def to_dev(a, dev):
    t = torch_mod().from_numpy(np.ascontiguousarray(a)).to(dev)
    return t


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def gpu_sort(x, y0, dims, perm, f, beta, dev, params=None):
    xd = to_dev(x, dev)
    yd = to_dev(y0, dev)
    S.sort(xd, yd, dims, perm, f, beta, params)
    return yd.cpu().numpy()


def oracle_sort(x, y0, dims, perm, f, beta):
    mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
    return O.sort_n(x, y0.copy(), dims, perm, f, beta, mode)


def log(*a):
    print(*a, flush=True)


def test_smoke_config1_dropin_host(dev):
    # configs[0]/[1]: tce_sort_4 30^4, b(i,j,k,l) = a(l,k,j,i), factor 1.0,
    # through the NWChem drop-in with host arrays
    log("case: tce_sort_4 host 30^4 4321")
    n = 30 ** 4
    x = O.uniform(n, 1)
    s = np.zeros(n)
    S.tce_sort_4(x, s, 30, 30, 30, 30, 4, 3, 2, 1, 1.0)
    ref = O.old_sort(x, np.zeros(n), (30,) * 4, (4, 3, 2, 1), 1.0, False)
    assert bits_equal(s, ref)


def test_fortran_abi_entry_points(dev):
    # call the Fortran-mangled symbols directly, every argument by reference
    log("case: tce_sort_4_/tce_sortacc_6_ by reference")
    L = _lib.lib()
    dims = (5, 6, 7, 8)
    n = int(np.prod(dims))
    x = O.uniform(n, 2)
    s = O.uniform(n, 3)
    ints = [ctypes.byref(ctypes.c_int(v)) for v in (*dims, 2, 4, 1, 3)]
    want = O.old_sort(x, s.copy(), dims, (2, 4, 1, 3), -0.5, True)
    L.tce_sortacc_4_(x.ctypes.data, s.ctypes.data, *ints, ctypes.byref(ctypes.c_double(-0.5)))
    assert bits_equal(s, want)
    d6 = (3, 4, 2, 5, 3, 2)
    n6 = int(np.prod(d6))
    x6, s6 = O.uniform(n6, 4), O.uniform(n6, 5)
    want6 = O.old_sort(x6, s6.copy(), d6, (6, 5, 4, 3, 2, 1), 37.0, True)
    ints6 = [ctypes.byref(ctypes.c_int(v)) for v in (*d6, 6, 5, 4, 3, 2, 1)]
    L.tce_sortacc_6(x6.ctypes.data, s6.ctypes.data, *ints6, ctypes.byref(ctypes.c_double(37.0)))
    assert bits_equal(s6, want6)


@pytest.mark.parametrize("dims", [(20,) * 4, (30,) * 4, (10, 30, 20, 40)])
def test_4d_sweep_all_perms(dev, dims):
    # 24 perms x factors {1,-1,0.5} x {sort, accumulate}
    n = int(np.prod(dims))
    x = O.uniform(n, 11)
    y0 = O.uniform(n, 12)
    xd = to_dev(x, dev)
    for perm in PERMS4:
        log("case: 4d", dims, perm)
        for f in (1.0, -1.0, 0.5):
            for beta in (0.0, 1.0):
                yd = to_dev(y0, dev)
                S.sort(xd, yd, dims, perm, f, beta)
                assert bits_equal(yd.cpu().numpy(), oracle_sort(x, y0, dims, perm, f, beta)), \
                    (dims, perm, f, beta)


def _sampled_check(dims, perms, seed, dev, nsample=1 << 20):
    """Every permutation of `perms`, sort and accumulate, with its tuned plan:
    the GPU output at 2^20 random positions plus the first and last 4096
    against the oracle's sampled closed form, bit for bit."""
    torch = torch_mod()
    n = int(np.prod(dims))
    x = O.uniform(n, seed)
    y0 = O.uniform(n, seed + 1)
    xd = to_dev(x, dev)
    y0d = to_dev(y0, dev)
    yd = torch.empty_like(y0d)
    rng = np.random.default_rng(seed)
    pos = np.concatenate([np.arange(4096), np.arange(n - 4096, n), rng.integers(0, n, size=nsample)])
    posd = torch.from_numpy(pos).to(dev)
    for perm in perms:
        for beta, f in ((0.0, 1.0), (1.0, -0.5)):
            log("case: sampled", dims, perm, beta)
            yd.copy_(y0d)
            S.sort(xd, yd, dims, perm, f, beta)
            got = yd.index_select(0, posd).cpu().numpy()
            want = O.closed_form_at(x, dims, perm, f, pos, y0 if beta else None)
            assert bits_equal(got, want), (dims, perm, beta)


@pytest.mark.parametrize("dims", [(12,) * 6, (14,) * 6, (15,) * 6, (16,) * 6, (17,) * 6, (18,) * 6,
                                  (19,) * 6, (56,) * 4, (64,) * 4, (72,) * 4, (80,) * 4])
def test_streamed_output_plans_sampled(dev, dims):
    # the plans for outputs written cold (dispatch_gfx950_streamed.jsonl,
    # taken after set_output_reuse(False)): every permutation of the sizes it
    # holds ((T) ones in 6D, plus the seeded random ones at 16^6), sort and
    # accumulate, against the oracle
    from spaghetty_amd.autotune import TRIPLES_6D, random_perms_6d
    if len(dims) == 4:
        perms = PERMS4
    else:
        extra = random_perms_6d() if dims[0] == 16 else []
        perms = sorted({tuple(int(c) for c in p) for p in TRIPLES_6D + extra})
    prev = S.set_output_reuse(False)
    try:
        _sampled_check(dims, perms, 131 + dims[0], dev, nsample=1 << 18)
    finally:
        S.set_output_reuse(prev)


def test_100x4_every_tuned_plan_sampled(dev):
    # configs: tce_sort_4 sweep at 100^4 -- all 24 permutations, sort and
    # accumulate, each on its tuned plan (tile shape, walk, kernel)
    _sampled_check((100,) * 4, PERMS4, 71, dev)


def test_20x6_random_perms_sampled(dev):
    # configs: tce_sort_6 at 20^6 -- all 64 seeded random permutations, sort
    # and accumulate, each on the plan the library picks for it
    from spaghetty_amd.autotune import random_perms_6d
    perms = [tuple(int(c) for c in p) for p in random_perms_6d()]
    _sampled_check((20,) * 6, perms, 73, dev)


def test_16x6_random_perms_sampled(dev):
    # configs: tce_sort_6 at 16^6 -- all 64 seeded random permutations
    # against the oracle's sampled closed form (beside the property checks
    # of test_6d_random_perms_16_properties)
    from spaghetty_amd.autotune import random_perms_6d
    perms = [tuple(int(c) for c in p) for p in random_perms_6d()]
    _sampled_check((16,) * 6, perms, 77, dev, nsample=1 << 18)


@pytest.mark.parametrize("e", [24, 32, 40, 48, 56, 64, 72, 80])
def test_nwchem_4d_sizes_every_tuned_plan_sampled(dev, e):
    # the other 4D tile sizes the table is tuned for (README.md:30-34):
    # all 24 permutations, sort and accumulate, each on its tuned plan
    _sampled_check((e,) * 4, PERMS4, 80 + e, dev, nsample=1 << 18)


@pytest.mark.parametrize("e", [12, 14, 15, 17, 18, 19])
def test_nwchem_6d_triples_sampled(dev, e):
    # every (T) permutation at the other 6D tile sizes the table is tuned for
    from spaghetty_amd.autotune import TRIPLES_6D
    perms = [tuple(int(c) for c in p) for p in TRIPLES_6D]
    _sampled_check((e,) * 6, perms, 90 + e, dev, nsample=1 << 18)


def test_rank_2_and_8_dropins(dev):
    # tce_sort_2/8 and tce_sortacc_2/8 (old_sort.f:10-31, 236-478, 489-510,
    # 715-957): the Fortran-mangled symbols by reference on host arrays, and
    # the Python mirror on device tensors, against the oracle (itself pinned
    # to the compiled reference old_sort_2/8 on the CPU)
    L = _lib.lib()
    rng = np.random.default_rng(12)
    log("case: tce_sort_2/tce_sortacc_2 300x500")
    dims = (300, 500)
    n = 300 * 500
    x = O.uniform(n, 31)
    for perm, acc in (((2, 1), False), ((2, 1), True), ((1, 2), True)):
        s0 = O.uniform(n, 32)
        want = O.sort_n(x, s0.copy(), dims, perm, -0.5, 1.0 if acc else 0.0, int(acc))
        s = s0.copy()
        fn = L.tce_sortacc_2_ if acc else L.tce_sort_2_
        fn(x.ctypes.data, s.ctypes.data, *[ctypes.byref(ctypes.c_int(v)) for v in (*dims, *perm)],
           ctypes.byref(ctypes.c_double(-0.5)))
        assert bits_equal(s, want), (perm, acc)
    log("case: tce_sort_8 random perms 3x4x2x5x3x2x4x3")
    dims8 = (3, 4, 2, 5, 3, 2, 4, 3)
    n8 = int(np.prod(dims8))
    x8 = O.uniform(n8, 33)
    for t in range(40):
        perm = tuple(int(v) + 1 for v in rng.permutation(8))
        acc = bool(t % 2)
        s0 = O.uniform(n8, 34 + t)
        want = O.sort_n(x8, s0.copy(), dims8, perm, 37.0, 1.0 if acc else 0.0, int(acc))
        s = s0.copy()
        fn = L.tce_sortacc_8 if acc else L.tce_sort_8
        fn(x8.ctypes.data, s.ctypes.data, *[ctypes.byref(ctypes.c_int(v)) for v in (*dims8, *perm)],
           ctypes.byref(ctypes.c_double(37.0)))
        assert bits_equal(s, want), (perm, acc)
    log("case: tce_sort_8 8^8 on device tensors")
    dims8 = (8,) * 8
    n8 = 8 ** 8
    x8 = O.uniform(n8, 35)
    for perm, acc in (((8, 7, 6, 5, 4, 3, 2, 1), False), ((2, 4, 6, 8, 1, 3, 5, 7), True)):
        s0 = O.uniform(n8, 36)
        want = O.sort_n(x8, s0.copy(), dims8, perm, 0.5, 1.0 if acc else 0.0, int(acc))
        xd, sd = to_dev(x8, dev), to_dev(s0, dev)
        (S.tce_sortacc_8 if acc else S.tce_sort_8)(xd, sd, *dims8, *perm, 0.5)
        assert bits_equal(sd.cpu().numpy(), want), (perm, acc)


def test_4d_dropin_names_host(dev):
    dims = (10, 30, 20, 40)
    n = int(np.prod(dims))
    x = O.uniform(n, 13)
    for perm in PERMS4[::5]:
        s = O.uniform(n, 14)
        want = O.old_sort(x, s.copy(), dims, perm, 0.5, True)
        S.tce_sortacc_4(x, s, *dims, *perm, 0.5)
        assert bits_equal(s, want)
        s = np.zeros(n)
        S.tce_sort_4(x, s, *dims, *perm, -1.0)
        assert bits_equal(s, O.old_sort(x, np.zeros(n), dims, perm, -1.0, False))


def test_6d_all_720_perms_tilesize_10(dev):
    # all 720 perms at 10^6, factor and mode rotate
    dims = (10,) * 6
    n = 10 ** 6
    x = O.uniform(n, 21)
    y0 = O.uniform(n, 22)
    xd = to_dev(x, dev)
    facs = (1.0, -1.0, 0.5)
    for i, perm in enumerate(itertools.permutations(range(1, 7))):
        if i % 60 == 0:
            log("case: 6d 10^6 perm #", i, perm)
        f = facs[i % 3]
        beta = float(i % 2)
        yd = to_dev(y0, dev)
        S.sort(xd, yd, dims, perm, f, beta)
        assert bits_equal(yd.cpu().numpy(), oracle_sort(x, y0, dims, perm, f, beta)), (perm, f, beta)


def _golden(tags):
    with open(GOLD) as f:
        return [c for c in json.load(f)["cases"] if c["tag"] in tags]


def _gen(spec, n, dims):
    if spec["kind"] == "uniform":
        return O.uniform(n, spec["seed"])
    if spec["kind"] == "kat":
        return O.kat(list(dims))
    return np.zeros(n)


def _digest_run(c, dev):
    dims = c["dims"]
    n = int(np.prod(dims))
    x = _gen(c["input"], n, dims)
    y0 = _gen(c["y0"], n, dims) if c["beta"] != 0.0 else np.zeros(n)
    xd, yd = to_dev(x, dev), to_dev(y0, dev)
    del x, y0
    S.sort(xd, yd, dims, [int(t) for t in c["perm"]], c["factor"], c["beta"])
    out = yd.cpu().numpy()
    return hashlib.sha256(np.ascontiguousarray(out, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("tags", [("kat4", "kat4acc", "kat6", "uni6acc", "uni4"),
                                  ("c30", "c30acc", "c20", "cnc", "cncacc"),
                                  ("t10", "t10acc"),
                                  ("c100", "c100acc"),
                                  ("r16", "r16acc", "r20", "r20acc")])
def test_reference_golden_digests(dev, tags):
    # outputs of the reference's own old_sort.f (compiled where the
    # reference exists) at KAT and full BASELINE sizes up to 100^4 / 20^6
    cases = _golden(tags)
    assert cases
    for c in cases:
        log("case: golden", c["tag"], c["dims"], c["perm"], c["factor"], c["beta"])
        assert _digest_run(c, dev) == c["sha256"], c


def test_6d_random_perms_16_properties(dev):
    # 64 seeded random perms at 16^6: inverse round trip is
    # the identity and accumulate onto zero equals sort
    from tests.golden.make_golden import random_perms_6d
    torch = torch_mod()
    dims = (16,) * 6
    n = 16 ** 6
    x = to_dev(O.uniform(n, 31), dev)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    for i, p in enumerate(random_perms_6d()):
        perm = [int(c) for c in p]
        if i % 16 == 0:
            log("case: 16^6 round trip", perm)
        inv = [0] * 6
        for q, a in enumerate(perm):
            inv[a - 1] = q + 1
        S.sort(x, y, dims, perm, 1.0, 0.0)
        S.sort(y, z, dims, inv, 1.0, 0.0)
        assert torch.equal(z.view(torch.int64), x.view(torch.int64)), perm
        z.zero_()
        S.sort(x, z, dims, perm, -0.5, 1.0)
        S.sort(x, y, dims, perm, -0.5, 0.0)
        assert torch.equal(z.view(torch.int64), y.view(torch.int64)), perm


def test_20_6_oracle_sample(dev):
    # 20^6 (512 MB per array): a few random perms against the OpenMP port
    # of the generated kernel (pinned to old_sort by the CPU tests)
    from tests.golden.make_golden import random_perms_6d
    dims = (20,) * 6
    n = 20 ** 6
    x = O.uniform(n, 41)
    y0 = O.uniform(n, 42)
    xd = to_dev(x, dev)
    for p in random_perms_6d()[60:62]:
        perm = [int(c) for c in p]
        log("case: 20^6 vs oracle", perm)
        yd = to_dev(y0, dev)
        S.sort(xd, yd, dims, perm, 0.5, 1.0)
        want = O.trans(x, y0.copy(), dims, perm, list(range(1, 7)), 0.5, 0.0, 1)
        assert bits_equal(yd.cpu().numpy(), want)


def test_special_values_and_contraction(dev):
    # +-0, subnormals, +-inf, huge/tiny products; factor 37 and random
    # factors make a fused multiply-add visible
    vals = np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, np.inf, -np.inf,
                     1e308, -1e308, 1.0 + 2.0 ** -52, 3.0, -7.5])
    rng = np.random.default_rng(5)
    dims = (3, 4, 5, 6)
    n = int(np.prod(dims))
    x = rng.choice(vals, n)
    y0 = rng.choice(vals, n)
    x[:n // 2] = O.uniform(n // 2, 51)
    for f in (37.0, float(rng.uniform(-2, 2)), 1.0 + 2.0 ** -30, -0.0):
        for beta in (0.0, 1.0, -2.5):
            for perm in PERMS4[::3]:
                log("case: special", f, beta, perm)
                got = gpu_sort(x, y0, dims, perm, f, beta, dev)
                want = oracle_sort(x, y0, dims, perm, f, beta)
                nan = np.isnan(want)
                assert np.array_equal(np.isnan(got), nan)
                assert bits_equal(got[~nan], want[~nan]), (f, beta, perm)


def test_nan_propagates_as_nan(dev):
    dims = (4, 5, 6, 7)
    n = int(np.prod(dims))
    x = O.uniform(n, 61)
    x[::7] = np.nan
    got = gpu_sort(x, np.zeros(n), dims, (4, 2, 3, 1), 2.0, 0.0, dev)
    want = oracle_sort(x, np.zeros(n), dims, (4, 2, 3, 1), 2.0, 0.0)
    assert np.array_equal(np.isnan(got), np.isnan(want))


def test_edge_shapes(dev):
    cases = [((1, 1, 1, 1), (4, 3, 2, 1)), ((1, 30, 1, 30), (4, 3, 2, 1)),
             ((2, 3), (2, 1)), ((7,), (1,)), ((), ()), ((3, 1, 4, 1, 5, 9), (6, 5, 4, 3, 2, 1)),
             ((13, 17, 19, 23), (3, 1, 4, 2)), ((1, 2, 1, 2, 1, 2, 1, 2), (8, 7, 6, 5, 4, 3, 2, 1)),
             ((4097, 3), (2, 1)), ((3, 5000), (2, 1)), ((64, 64, 64), (3, 2, 1))]
    for dims, perm in cases:
        log("case: edge", dims, perm)
        n = int(np.prod(dims)) if dims else 1
        x, y0 = O.uniform(n, 71), O.uniform(n, 72)
        for beta in (0.0, 1.0):
            assert bits_equal(gpu_sort(x, y0, dims, perm, -1.5, beta, dev),
                              oracle_sort(x, y0, dims, perm, -1.5, beta)), (dims, perm, beta)


def test_random_untuned_problems(dev):
    # problems the dispatch table does not hold run on the library's default
    # plan (nontemporal access and small tiles up to 2^23 elements): random
    # ranks 2..8, extents, permutations, factors and g, against the oracle
    rng = np.random.default_rng(2026)
    for i in range(160):
        R = int(rng.integers(2, 9))
        hi = {2: 700, 3: 90, 4: 40, 5: 16, 6: 10, 7: 7, 8: 5}[R]
        dims = tuple(int(v) for v in rng.integers(1, hi + 1, size=R))
        perm = tuple(int(v) + 1 for v in rng.permutation(R))
        f = float(rng.choice([1.0, -1.0, 0.5, 37.0]))
        beta = float(rng.choice([0.0, 1.0, -2.0]))
        n = int(np.prod(dims))
        log("case: random", i, dims, perm, f, beta)
        x, y0 = O.uniform(n, 200 + i), O.uniform(n, 400 + i)
        assert bits_equal(gpu_sort(x, y0, dims, perm, f, beta, dev),
                          oracle_sort(x, y0, dims, perm, f, beta)), (dims, perm, f, beta)


def test_empty_and_errors(dev):
    torch = torch_mod()
    x = torch.zeros(10, dtype=torch.float64, device=dev)
    S.sort(x, torch.zeros_like(x), (0, 5, 2), (3, 2, 1))
    with pytest.raises(ValueError):
        S.sort(x, x, (2, 5), (2, 1))  # aliasing
    with pytest.raises(ValueError):
        S.sort(x, torch.zeros_like(x), (2, 5), (2, 2))
    # a bad permutation through the drop-in leaves sorted untouched
    s = torch.full((10,), 3.0, dtype=torch.float64, device=dev)
    S.tce_sort_4(x, s, 1, 2, 5, 1, 1, 1, 2, 3, 1.0)
    assert torch.all(s == 3.0)


def test_forced_tile_parameters_on_device(dev):
    # every tile shape the autotuner can choose gives identical results
    dims, perm = (30, 30, 30, 30), (4, 3, 2, 1)
    n = 30 ** 4
    x, y0 = O.uniform(n, 81), O.uniform(n, 82)
    want = oracle_sort(x, y0, dims, perm, 0.5, 1.0)
    for nx, tx, ny, ty, block, g, fl in ((1, 30, 1, 30, 256, 0, 0), (2, 2, 2, 2, 256, 0, 1),
                                         (1, 16, 2, 3, 128, 2, 2), (2, 1, 1, 30, 512, 1, 3),
                                         (1, 30, 2, 4, 256, 4, 4), (1, 7, 1, 11, 64, 1, 6)):
        log("case: forced", nx, tx, ny, ty, block, g, fl)
        p = TileParams.default(kind=0, nx=nx, tx=tx, ny=ny, ty=ty, block=block, grid_per_cu=g,
                               flags=fl)
        assert bits_equal(gpu_sort(x, y0, dims, perm, 0.5, 1.0, dev, p), want)
    for kind in (0, 1):
        for fl in (-1, 0, 1):
            p = TileParams.default(kind=kind, flags=fl)
            w2 = oracle_sort(x, y0, dims, (2, 1, 3, 4), 0.5, 1.0)
            assert bits_equal(gpu_sort(x, y0, dims, (2, 1, 3, 4), 0.5, 1.0, dev, p), w2)


def test_kernel_variants_on_device(dev):
    # classic vs branch-free kernel, pairs vs single elements, one or two LDS
    # buffers, grouped walk orders, nontemporal / pipelined access: all exact,
    # on 4D and 6D shapes with even extents (pairs) and odd ones (singles)
    CLASSIC, SCALAR, LDS1 = 8, 16, 32
    cases = (((30, 30, 30, 30), (4, 3, 2, 1)), ((16, 16, 16, 16, 16, 16), (6, 5, 4, 3, 2, 1)),
             ((10, 30, 20, 40), (2, 4, 1, 3)), ((15, 17, 19, 21), (3, 4, 1, 2)),
             ((8, 12, 10, 6, 14, 4), (4, 6, 2, 5, 1, 3)), ((40, 50, 60, 70), (4, 3, 2, 1)))
    for dims, perm in cases:
        n = int(np.prod(dims))
        x, y0 = O.uniform(n, 83), O.uniform(n, 84)
        for beta in (0.0, 1.0, -2.0):
            want = oracle_sort(x, y0, dims, perm, -0.5, beta)
            for fl in (0, 2, 6, 3, CLASSIC, SCALAR | 2, LDS1, LDS1 | 6, SCALAR | LDS1):
                for wx in (0, 3):
                    log("case: variant", dims, perm, beta, fl, wx)
                    p = TileParams.default(flags=fl, walk_x=wx)
                    assert bits_equal(gpu_sort(x, y0, dims, perm, -0.5, beta, dev, p), want)
            # walk orders, chunked walks and tiles above 4096 elements (pair kernel)
            for mode in (1, 2, 3):
                for mt, fl, ch in ((0, 2, 0), (12288, 0, 3), (12288, 2, 0), (12288, LDS1 | 6, 5)):
                    log("case: walk/big", dims, perm, beta, mode, mt, fl, ch)
                    p = TileParams.default(flags=fl, walk_mode=mode, walk_x=4, max_tile=mt,
                                           walk_chunk=ch)
                    assert bits_equal(gpu_sort(x, y0, dims, perm, -0.5, beta, dev, p), want)


def test_misaligned_device_pointers(dev):
    # views that start at an odd element (8-byte, not 16-byte aligned)
    torch = torch_mod()
    dims = (1, 1, 6, 1000)
    n = 6000
    x = O.uniform(n + 1, 91)
    xd = to_dev(x, dev)[1:]
    yd = torch.zeros(n + 1, dtype=torch.float64, device=dev)[1:]
    S.sort(xd, yd, dims, (1, 2, 3, 4), 2.0, 0.0)
    assert bits_equal(yd.cpu().numpy(), 2.0 * x[1:])
    # a transposing plan whose tiles would take 16-byte pairs falls back to
    # single elements on 8-byte aligned views
    dims, perm = (16, 16, 16, 16), (4, 3, 2, 1)
    n = 16 ** 4
    x = O.uniform(n + 1, 92)
    y0 = O.uniform(n + 1, 93)
    xd = to_dev(x, dev)[1:]
    yd = to_dev(y0, dev)[1:]
    S.sort(xd, yd, dims, perm, 0.5, 1.0)
    want = oracle_sort(x[1:].copy(), y0[1:].copy(), dims, perm, 0.5, 1.0)
    assert bits_equal(yd.cpu().numpy(), want)
