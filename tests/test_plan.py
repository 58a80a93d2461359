"""Plan logic on the CPU (no GPU): validation, fusion, tile choice, and the
kernels' exact index math walked on the host (spg_plan_check), compared
bit-for-bit with the oracle.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import spaghetty_amd as S
from spaghetty_amd import TileParams
from oracle import oracle as O


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def check_walk(dims, perm, f, beta, params=None, seed=1):
    n = int(np.prod(dims))
    x = O.uniform(n, seed)
    y0 = O.uniform(n, seed + 1)
    got = y0.copy()
    S.plan_check(x, got, dims, perm, f, beta, params)
    mode = 0 if beta == 0.0 else (1 if beta == 1.0 else 2)
    want = O.sort_n(x, y0.copy(), dims, perm, f, beta, mode)
    assert same(got, want), (dims, perm, f, beta, params)


@pytest.mark.parametrize("dims", [(3, 4, 5, 2), (7, 1, 9, 6), (10, 30, 20, 40), (20, 20, 20, 20)])
def test_all_24_perms_4d(dims):
    for perm in itertools.permutations([1, 2, 3, 4]):
        for f, beta in ((1.0, 0.0), (-1.0, 1.0), (0.5, 0.7)):
            check_walk(dims, perm, f, beta)


@pytest.mark.parametrize("dims", [(2, 3, 4, 3, 2, 3), (5, 1, 3, 2, 4, 3)])
def test_all_720_perms_6d(dims):
    for perm in itertools.permutations(range(1, 7)):
        check_walk(dims, perm, 37.0, 1.0)


def test_ranks_0_to_8():
    check_walk((), (), 2.0, 0.0)
    rng = np.random.default_rng(7)
    for R in (1, 2, 3, 5, 7, 8):
        for _ in range(30):
            dims = [int(v) for v in rng.integers(1, 5, size=R)]
            perm = [int(v) + 1 for v in rng.permutation(R)]
            check_walk(dims, perm, -0.25, float(rng.choice([0.0, 1.0, 3.0])))


def test_forced_tile_parameters_cover_every_element():
    # every (nx, ny, tx, ty) the autotuner may ask for must still be exact
    dims, perm = (6, 7, 5, 9), (4, 3, 2, 1)
    for nx in (1, 2):
        for ny in (1, 2):
            for tx in (1, 2, 3, 5):
                for ty in (1, 2, 4):
                    p = TileParams.default(kind=0, nx=nx, ny=ny, tx=tx, ty=ty)
                    try:
                        plan = S.plan(dims, perm, 0.0, p)
                    except ValueError:
                        continue
                    assert plan["nx"] == nx and plan["ny"] == ny
                    check_walk(dims, perm, 0.5, 1.0, p)
    # shared contiguous axis: rows form and tiled form
    for kind in (0, 1):
        check_walk((5, 6, 7, 8), (2, 1, 3, 4), 1.0, 0.0, TileParams.default(kind=kind))


def test_walk_orders_cover_every_element():
    # every walk mode, with tiles walked in groups of g values of one grid
    # digit (narrow last group included), still produces every element
    # once, with the kernels' own index math
    for dims, perm in (((6, 7, 5, 9), (4, 3, 2, 1)), ((9, 10, 11), (3, 1, 2)),
                       ((4, 6, 8, 10), (2, 4, 1, 3)), ((3, 4, 2, 5, 2, 3), (6, 5, 4, 3, 2, 1)),
                       ((5, 1, 6, 7), (1, 4, 2, 3))):
        for mode in (0, 1, 2, 3):
            for g in (0, 1, 2, 3, 5):
                for tx, ty in ((1, 1), (2, 1), (1, 2)):
                    p = TileParams.default(kind=0, nx=1, ny=1, tx=tx, ty=ty, walk_x=g,
                                           walk_mode=mode)
                    try:
                        S.plan(dims, perm, 1.0, p)
                    except ValueError:
                        continue
                    check_walk(dims, perm, -0.5, 1.0, p)
                    check_walk(dims, perm, 2.0, 0.0, p)
    # the grouped order is really taken where the digits are wide enough:
    # x chunks against y chunks (mode 0), the input-nearest Z axis against
    # y chunks (mode 1)
    pl = S.plan((9, 10, 11), (3, 1, 2), 0.0, TileParams.default(kind=0, nx=1, ny=1, tx=1, ty=1,
                                                               walk_x=3))
    assert pl["gw"] == 3 and pl["gp"] == 0
    pl = S.plan((8, 9, 10, 11), (4, 3, 2, 1), 0.0,
                TileParams.default(kind=0, nx=1, ny=1, tx=2, ty=2, walk_x=4, walk_mode=1))
    assert pl["gw"] == 4 and pl["gp"] == 1 and pl["ng"] == 4
    with pytest.raises(ValueError):
        S.plan((8, 9, 10, 11), (4, 3, 2, 1), 0.0, TileParams.default(walk_mode=7))


def test_big_tiles_only_with_pairs():
    # tiles above 4096 elements run only on the pair kernel, up to 12288
    # elements (768 threads), except a pipelined accumulate: 8192 (512)
    big = TileParams.default(max_tile=12288)
    dims, perm = (40, 50, 60, 70), (4, 3, 2, 1)  # X = the 70 axis, Y = 40 x chunk of 50
    PIPE = 2
    for beta in (0.0, 1.0, -2.0):
        for fl, cap in ((0, 12288), (PIPE, 12288 if beta == 0.0 else 8192)):
            for ty in (2, 4):
                p = TileParams.default(kind=0, nx=1, tx=70, ny=2, ty=ty, max_tile=12288, flags=fl)
                V = 70 * 40 * ty
                if V > cap:
                    with pytest.raises(ValueError):
                        S.plan(dims, perm, beta, p)
                    continue
                pl = S.plan(dims, perm, beta, p)
                assert pl["V"] == V > 4096
                assert pl["vec"] == 2 and pl["tk"] == 1 and pl["block"] * 16 >= pl["V"]
                assert pl["block"] <= (768 if cap == 12288 else 512)
                assert pl["lds_bytes"] <= 160 * 1024
                check_walk(dims, perm, 0.5, beta, p)
            q = TileParams.default(max_tile=12288, flags=fl)
            pl = S.plan(dims, perm, beta, q)
            assert pl["V"] <= cap
            check_walk(dims, perm, 0.5, beta, q)
        pl = S.plan(dims, perm, beta, big)
        check_walk(dims, perm, 0.5, beta, big)
    # the classic kernel and single elements stay within 4096
    for fl in (8, 16):
        pl = S.plan(dims, perm, 0.0, TileParams.default(max_tile=12288, flags=fl))
        assert pl["V"] <= 4096
    # over odd extents a tile pairs (and may exceed 4096 elements) only with
    # even tile extents: 8-byte aligned pairs, split at partial-tile edges
    p = TileParams.default(kind=0, nx=2, tx=2, ny=2, ty=2, max_tile=12288)
    pl = S.plan((41, 51, 61, 71), perm, 0.0, p)
    assert pl["V"] == 142 * 82 > 4096 and pl["vec"] == 2 and pl["pair8"] == 1
    check_walk((41, 51, 61, 71), perm, 0.5, 1.0, p)
    p = TileParams.default(kind=0, nx=1, tx=71, ny=2, ty=2, max_tile=12288)
    with pytest.raises(ValueError):  # an odd tile extent: single elements, at most 4096
        S.plan((41, 51, 61, 71), perm, 0.0, p)


@pytest.mark.parametrize("dims,perm,shape", [
    ((15, 17, 19, 21), (4, 3, 2, 1), dict(nx=1, tx=16, ny=1, ty=6)),
    ((15, 17, 19, 21), (4, 3, 2, 1), dict(nx=2, tx=2, ny=2, ty=2)),
    ((15, 13, 11, 9), (3, 4, 1, 2), dict(nx=1, tx=6, ny=1, ty=10)),
    ((9, 7, 15, 13, 11, 5), (4, 6, 2, 5, 1, 3), dict(nx=2, tx=2, ny=1, ty=8)),
    ((7, 9, 11, 13, 15, 17), (6, 5, 4, 3, 2, 1), dict(nx=2, tx=4, ny=2, ty=2)),
    ((1001, 999), (2, 1), dict(nx=1, tx=64, ny=1, ty=64)),
])
def test_odd_extent_pair_plans(dims, perm, shape):
    # even tile extents over odd array extents: pair plans (pair8) whose
    # pairs start at odd offsets and whose partial tiles cut pairs; the host
    # walk checks the pair invariants and every value (sort, accumulate,
    # five-case) against the oracle
    for beta in (0.0, 1.0, -2.5):
        for fl in (0, 2, 32):
            p = TileParams.default(kind=0, flags=fl, **shape)
            pl = S.plan(dims, perm, beta, p)
            assert pl["vec"] == 2 and pl["pair8"] == 1 and pl["Tx"] % 2 == 0 and pl["Ty"] % 2 == 0
            check_walk(dims, perm, -0.5, beta, p)
    # even extents keep 16-byte aligned pairs
    even = tuple(d + (d & 1) for d in dims)
    pl = S.plan(even, perm, 0.0, TileParams.default(kind=0, **shape))
    assert pl["pair8"] == 0


BLOCK_PROBLEMS = [
    # the two fastest axes swap places: every tile of them is one block,
    # contiguous on both sides, and a tile may span several blocks
    ((5, 6, 7, 4, 3, 9), (4, 3, 2, 1, 6, 5)),      # 6D, (T)-like 432165
    ((7, 5, 9, 6, 8, 4), (4, 3, 2, 1, 6, 5)),      # even X / Y extents: pairs
    ((6, 8, 11, 10), (1, 2, 4, 3)),                # 4D, W the same axis on both sides
    ((5, 13, 12, 4), (2, 1, 4, 3)),
    ((3, 5, 4, 6, 7), (2, 1, 4, 3, 5)),            # shared contiguous axis (C = 7)
    ((3, 5, 4, 6, 8), (1, 2, 4, 3, 5)),            # C = 8: pairs along c
    ((3, 2, 3, 4, 2, 3, 3, 4), (1, 3, 4, 2, 5, 6, 8, 7)),  # rank 8 (rank 6 after fusion)
]


@pytest.mark.parametrize("dims,perm", BLOCK_PROBLEMS)
def test_block_tile_plans(dims, perm):
    # tiles of several blocks along the input's next axis (tw > 1) or the
    # output's (tw < -1), whole and partial W chunks, on every kernel form;
    # the host walk checks every value against the oracle
    base = S.plan(dims, perm, 0.0, TileParams.default(kind=0, tw=1))
    s = 1 if base["C"] > 1 else 0
    assert base["out_axis"][s] == s + 1 and base["out_axis"][s + 1] == s
    for tw in (2, 3, -2, -3, 0):
        for fl in (0, 2, 32, 8, 16):
            p = TileParams.default(kind=0, flags=fl, tw=tw)
            pl = S.plan(dims, perm, 0.0, p)
            if tw:
                ext = base["ext"]
                blk = base["C"] * ext[s] * ext[s + 1]
                assert pl["tw"] == abs(tw) and pl["V"] == blk * abs(tw)
                assert pl["ntiles"] == -(-pl["w_top"] // abs(tw)) * (base["nelem"] // (blk * pl["w_top"]))
            for beta in (0.0, 1.0, -2.5):
                check_walk(dims, perm, -0.5, beta, p)


def test_block_tiles_rules():
    dims, perm = (16,) * 6, (4, 3, 2, 1, 6, 5)
    # the cost model's own choice spans several blocks, as large as fits
    pl = S.plan(dims, perm, 0.0, TileParams.default())
    assert pl["tw"] > 1 and pl["V"] == 256 * pl["tw"] and pl["vec"] == 2
    # the two candidate axes: input axis 2, or the output's next (input axis 5)
    assert S.plan(dims, perm, 0.0, TileParams.default(tw=4))["w_axis"] == 2
    assert S.plan(dims, perm, 0.0, TileParams.default(tw=-4))["w_axis"] == 5
    # where one axis is next on both sides, either sign names it
    for tw in (3, -3):
        assert S.plan((6, 8, 11, 10), (1, 2, 4, 3), 0.0, TileParams.default(tw=tw))["w_axis"] == 2
    # W spans whole X and Y axes only, and only where the two fastest axes swap
    with pytest.raises(ValueError):
        S.plan(dims, perm, 0.0, TileParams.default(tw=4, nx=1, tx=8, ny=1, ty=16))
    with pytest.raises(ValueError):
        S.plan((16,) * 6, (6, 5, 4, 3, 2, 1), 0.0, TileParams.default(tw=4))
    with pytest.raises(ValueError):
        S.plan((16,) * 4, (1, 2, 3, 4), 0.0, TileParams.default(tw=2))
    with pytest.raises(ValueError):
        S.plan(dims, perm, 0.0, TileParams.default(tw=-1))
    # tw = 1 keeps one block per tile; records of the dispatch table without
    # a "tw" field read the same way
    assert S.plan(dims, perm, 0.0, TileParams.default(tw=1))["tw"] == 1
    assert TileParams.from_record({"kind": 0, "nx": 1}).tw == 1


def test_fusion():
    # identity fuses to one contiguous stream
    p = S.plan((30, 30, 30, 30), (1, 2, 3, 4))
    assert p["rank"] == 1 and p["kind"] == 1
    # 3412 is a 2D transpose of (12)x(34)
    p = S.plan((4, 5, 6, 7), (3, 4, 1, 2))
    assert p["rank"] == 2 and p["ext"] == [42, 20]
    # extent-1 axes vanish
    p = S.plan((1, 9, 1, 8), (4, 3, 2, 1))
    assert p["rank"] == 2
    # keys are per canonical problem
    assert S.dispatch_key((30,) * 4, (4, 3, 2, 1)) == "r4:s:n30x30x30x30:p3210"
    assert S.dispatch_key((30,) * 4, (4, 3, 2, 1), 1.0).startswith("r4:a:")


def test_tiles_are_bounded_and_large_enough():
    for dims, perm in (((100,) * 4, (4, 3, 2, 1)), ((16,) * 6, (6, 5, 4, 3, 2, 1)),
                       ((30,) * 4, (2, 4, 1, 3)), ((10,) * 6, (3, 6, 2, 5, 1, 4))):
        p = S.plan(dims, perm)
        assert p["kind"] == 0
        assert 64 <= p["V"] <= 12288
        assert p["block"] * p["slots"] * p["vec"] >= p["V"]
        assert p["lds_bytes"] <= 160 * 1024


def test_untuned_small_problems_fill_the_device():
    # a problem of at most 2^23 elements that the dispatch table does not
    # hold gets nontemporal access and tiles of at most nelem/256 elements
    for dims, perm in (((26,) * 4, (4, 3, 2, 1)), ((36,) * 4, (2, 4, 1, 3)), ((11,) * 6, (6, 5, 4, 3, 2, 1))):
        assert S.dispatch_key(dims, perm) not in _table_keys()
        n = int(np.prod(dims))
        for beta in (0.0, 1.0):
            p = S.plan(dims, perm, beta)
            assert p["nt"] == 1
            assert p["V"] <= max(256, min(4096, 1 << ((n // 256).bit_length() - 1)))
            check_walk(dims, perm, 0.5, beta)
    # the cost model's own choice is still what all-automatic parameters give
    p = S.plan((26,) * 4, (4, 3, 2, 1), 0.0, TileParams.default())
    assert p["V"] > 1024
    # larger untuned problems keep the cost model's plan
    dims, perm = (21,) * 6, (6, 5, 4, 3, 2, 1)
    assert S.dispatch_key(dims, perm) not in _table_keys()
    assert S.plan(dims, perm) == S.plan(dims, perm, 0.0, TileParams.default())


def _table_keys():
    import json
    import os
    path = os.path.join(os.path.dirname(S.__file__), "dispatch_gfx950.jsonl")
    with open(path) as f:
        return {json.loads(line)["key"] for line in f if line.strip()}


def test_invalid_arguments():
    with pytest.raises(ValueError):
        S.plan((3, 4, 5, 6), (1, 2, 2, 4))
    with pytest.raises(ValueError):
        S.plan((3, 4, 5, 6), (1, 2, 3, 5))
    with pytest.raises(ValueError):
        S.plan((3, -4, 5, 6), (1, 2, 3, 4))
    with pytest.raises(ValueError):
        S.plan((2000,) * 3, (3, 2, 1))  # 8e9 elements > 2^31-1


def test_version_key_matches_reference():
    # version = 1000*i+100*j+10*k+l (trunk/makelib.py:79)
    assert S.version_of((4, 3, 2, 1)) == 4321
    assert S.version_of((1, 2, 3, 4)) == 1234


def test_fast_division_exact():
    from spaghetty_amd import _lib
    assert _lib.lib().spg_selftest_fastdiv(2_000_000) == 0


def test_every_dispatch_entry_builds_and_is_used():
    # each record of the committed table builds a plan with its parameters,
    # and a call without parameters on that problem takes exactly that plan
    import json
    import os
    path = os.path.join(os.path.dirname(S.__file__), "dispatch_gfx950.jsonl")
    with open(path) as f:
        recs = [json.loads(line) for line in f if line.strip()]
    assert len(recs) > 1000
    for r in recs:
        dims, perm, beta = tuple(r["dims"]), tuple(r["perm"]), 1.0 if r["acc"] else 0.0
        assert S.dispatch_key(dims, perm, beta) == r["key"]
        forced = S.plan(dims, perm, beta, TileParams.from_record(r))
        assert S.plan(dims, perm, beta) == forced, r["key"]


def test_streamed_output_table():
    # the second table (plans for outputs written cold) is consulted first
    # only after set_output_reuse(False); each of its records builds and is
    # then the plan of its problem, and problems it lacks keep the main table's
    import json
    import os
    here = os.path.dirname(S.__file__)
    with open(os.path.join(here, "dispatch_gfx950_streamed.jsonl")) as f:
        recs = [json.loads(line) for line in f if line.strip()]
    with open(os.path.join(here, "dispatch_gfx950.jsonl")) as f:
        main = {r["key"]: r for r in (json.loads(line) for line in f if line.strip())}
    assert len(recs) >= 100
    args = lambda r: (tuple(r["dims"]), tuple(r["perm"]), 1.0 if r["acc"] else 0.0)
    prev = S.set_output_reuse(False)
    try:
        assert prev is True
        for r in recs:
            dims, perm, beta = args(r)
            assert S.dispatch_key(dims, perm, beta) == r["key"]
            assert S.plan(dims, perm, beta) == S.plan(dims, perm, beta, TileParams.from_record(r)), r["key"]
        other = main["r4:s:n100x100x100x100:p3210"]
        assert S.plan(*args(other)) == S.plan(*args(other), TileParams.from_record(other))
    finally:
        assert S.set_output_reuse(True) is False
    differ = 0
    for r in recs:
        dims, perm, beta = args(r)
        assert S.plan(dims, perm, beta) == S.plan(dims, perm, beta, TileParams.from_record(main[r["key"]]))
        differ += S.plan(dims, perm, beta) != S.plan(dims, perm, beta, TileParams.from_record(r))
    assert differ >= len(recs) // 2


def test_streamed_outputs_from_the_environment():
    # SPG_OUTPUTS=streamed (for callers that cannot call set_output_reuse,
    # e.g. Fortran) starts the process planning for outputs written cold
    import json
    import os
    import subprocess
    import sys
    dims, perm = (16,) * 6, (4, 3, 2, 1, 6, 5)
    code = ("import json, spaghetty_amd as S; "
            f"print(json.dumps(S.plan({dims!r}, {perm!r}, 0.0)))")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, SPG_OUTPUTS="streamed", PYTHONPATH=root)
    env.pop("SPG_DISPATCH_TABLE_STREAMED", None)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True, timeout=120).stdout
    here = os.path.dirname(S.__file__)
    key = S.dispatch_key(dims, perm, 0.0)
    with open(os.path.join(here, "dispatch_gfx950_streamed.jsonl")) as f:
        rec = next(r for r in (json.loads(line) for line in f if line.strip()) if r["key"] == key)
    assert json.loads(out.strip().splitlines()[-1]) == S.plan(dims, perm, 0.0, TileParams.from_record(rec))
    assert json.loads(out.strip().splitlines()[-1]) != S.plan(dims, perm, 0.0)


def test_largest_stated_6d_tiles_plan_for_every_permutation():
    # 20^3 x 40^3 = 512M elements (4 GB per array), inside the coupled-cluster
    # tile ranges and the 2^31-element limit: every permutation must get a
    # launchable plan.  Where the cost model's favourite tile (above 4,096
    # elements, which needs the 16-byte buffer path) has offsets beyond a
    # buffer resource's 2 GB range, a tile the classic kernel can run wins.
    classic = 0
    for dims in ((20, 20, 20, 40, 40, 40),):
        for perm in itertools.permutations(range(1, 7)):
            for beta in (0.0, 1.0):
                d = S.plan(dims, perm, beta)
                if d["kind"] == 0 and d["tk"] == 0:
                    assert d["V"] <= 4096
                    classic += 1
                if d["V"] > 4096:
                    assert d["tk"] == 1 and d["vec"] == 2
    assert classic > 0
