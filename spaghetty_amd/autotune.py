"""On-device autotuner: the counterpart of the reference's brute-force loop
order search (trunk/generator.py:162-313 times every loop order of every
permutation and keeps the fastest that matches old_sort_4 exactly; the
winners are hand-copied into trunk/makelib.py:39-62).

Here the search space is the GPU plan: which input axes form the
input-contiguous group (nx), which output axes the output-contiguous group
(ny), the chunk of each group's top axis (tx, ty), threads per workgroup and
the persistent grid.  Candidates come from the library's cost model; each is
timed on the device with rotating buffers (so the 256 MB Infinity Cache does
not serve small tiles) and accepted only if its output is bit-identical to
the default plan's -- the reference's ``errors == 0`` gate
(generator.py:268-278).  Winners are written as one JSON object per line
(spaghetty_amd/dispatch_gfx950.jsonl), which the library loads at first use.

    python -m spaghetty_amd.autotune --problems baseline --out spaghetty_amd/dispatch_gfx950.jsonl
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

from . import _lib
from .api import dispatch_key, plan
from ._lib import TileParams

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "dispatch_gfx950.jsonl")
# Device memory the timed calls rotate over.  0 (the default) follows the
# table's policy: outputs that fit in the 256 MB last-level cache are written
# again two calls later (512 MB of arrays, the regime of a caller reusing its
# arrays); larger outputs rotate over at least 4 GB, so every write is cold.
ROTATE_BYTES = 0
MALL_BYTES = 256 << 20


def _chunks(m: int) -> list[int]:
    # divisors, round sizes, and (odd extents have odd divisors only) the
    # small even chunks that give tiles of even extent, which take pairs
    c = {m}
    for t in range(1, m + 1):
        if m % t == 0 or t in (2, 4, 6, 8, 10, 12, 14, 16, 24, 32, 40, 48, 64, 96, 128, 192, 256):
            c.add(t)
    return sorted(c)


CLASSIC, SCALAR, LDS1, NOSPLIT = 8, 16, 32, 128
# access flags tried: {ordinary, nontemporal, nontemporal loads only} x
# {plain, pipelined} (SPG_TILE_NT = 1, SPG_TILE_PIPE = 2, SPG_TILE_NT_LOAD = 4)
ACCESS_FLAGS = (0, 2, 4, 6, 1, 3)


def shape_candidates(dims, perm, beta, limit=12, big=4) -> list[TileParams]:
    """Distinct tile shapes (groups and chunks) the cost model ranks best,
    the library's default first, then up to `big` shapes above 4096
    elements (pair kernel only); plus the rows form where one exists, and,
    where the two fastest axes swap places, tiles of several blocks along
    either candidate axis."""
    base = plan(dims, perm, beta)
    out, seen = [], set()

    def add(p, cost=None):
        try:
            pl = plan(dims, perm, beta, p)
        except ValueError:
            return
        sig = (pl["kind"], pl["nx"], pl["ny"], pl["tx"], pl["ty"], pl["tw"], pl["w_axis"])
        if sig in seen:
            return
        seen.add(sig)
        out.append((pl["cost"] if cost is None else cost, p))

    add(TileParams.default(), cost=-1.0)
    rank = base["rank"]
    scored = []
    for nx in range(0, min(rank, 3) + 1):
        for ny in range(0, min(rank, 3) + 1):
            if nx == 0 and ny == 0:
                continue
            try:
                pl = plan(dims, perm, beta, TileParams.default(kind=0, nx=nx, ny=ny))
            except ValueError:
                continue
            xt = pl["x_ext"][-1] if pl["x_ext"] else 1
            yt = pl["y_ext"][-1] if pl["y_ext"] else 1
            for tx in (_chunks(xt) if nx else [1]):
                for ty in (_chunks(yt) if ny else [1]):
                    p = TileParams.default(kind=0, nx=nx, tx=tx if nx else 0, ny=ny,
                                           ty=ty if ny else 0, max_tile=12288)
                    try:
                        q = plan(dims, perm, beta, p)
                    except ValueError:
                        continue
                    if q["nx"] != nx or q["ny"] != ny:
                        continue
                    scored.append((q["cost"], q["V"] > 4096, q["vec"] == 2, p))
    scored.sort(key=lambda t: t[0])
    small = [(c, p) for c, b, v, p in scored if not b]
    large = [(c, p) for c, b, v, p in scored if b]
    for c, p in small[: 4 * limit]:
        if len(out) >= limit:
            break
        add(p, c)
    n0 = len(out)
    for c, p in large[: 4 * big]:
        if len(out) >= n0 + big:
            break
        add(p, c)
    # over odd extents only shapes with even tile extents move 16-byte
    # pairs, which the cost model's line count does not price: the best of
    # those join the list
    n0 = len(out)
    for c, p in [(c, p) for c, b, v, p in scored if v][: 4 * limit]:
        if len(out) >= n0 + max(2, limit // 2):
            break
        add(p, c)
    # the cost model's own choice under smaller caps: its ranking of the
    # shapes above 4096 elements is not the device's (at 100^4, 3421's best
    # shape under a 6,144 cap ran 4% faster than the tuned 12,288 one)
    for cap in (6144, 8192, 10240):
        add(TileParams.default(max_tile=cap))
    if base["C"] > 1:
        add(TileParams.default(kind=1), cost=0.0)
    # tiles of several blocks (the two fastest axes swap places): every
    # chunk of both candidate axes, the one-block tile among them
    for side in (1, -1):
        try:
            top = plan(dims, perm, beta, TileParams.default(tw=2 * side))["w_top"]
        except ValueError:
            continue
        add(TileParams.default(kind=0, tw=1, max_tile=12288))
        for t in _chunks(top):
            if t > 1:
                add(TileParams.default(kind=0, tw=side * t, max_tile=12288))
    return [p for _, p in out]


def _with(p: TileParams, **kw) -> TileParams:
    return TileParams.default(**{**p.as_dict(), **kw})


class Bench:
    """Rotating device buffers for one problem size."""

    def __init__(self, dims, dev):
        import torch
        n = 1
        for d in dims:
            n *= d
        self.n = n
        rotate = ROTATE_BYTES or ((512 << 20) if 8 * n <= MALL_BYTES else (4 << 30))
        self.count = max(1, -(-rotate // (16 * n)))
        self.x = torch.rand(self.count, n, dtype=torch.float64, device=dev) * 2 - 1
        self.y = torch.rand(self.count, n, dtype=torch.float64, device=dev) * 2 - 1
        self.i = 0

    def time(self, dims, perm, beta, params, reps=20, trials=3) -> float:
        """Device seconds per call: `reps` calls on rotating buffers are
        captured once into a graph (so host launch overhead does not hide
        the kernels' own time on small problems) and replayed `trials`
        times between HIP events; the best replay counts."""
        import torch
        from .api import sort
        for _ in range(2):
            sort(self.x[0], self.y[0], dims, perm, 1.0, beta, params)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for r in range(reps):
                k = (self.i + r) % self.count
                sort(self.x[k], self.y[k], dims, perm, 1.0, beta, params)
        self.i += reps
        g.replay()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(trials):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            g.replay()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) / 1e3 / reps)
        del g
        return best


def tune_problem(dims, perm, beta, dev, budget_s=20.0, verbose=True, incumbent=None,
                 depth=1, shapes=12, log=None) -> dict:
    """Staged search (each stage keeps the best of the previous one):
    1. tile shapes from the cost model;
    2. kernel (branch-free k_vtile or classic k_tiled) x access flags for the
       best three shapes;
    3. walk order (groups of walk_x x chunks) for the best two;
    4. threads per workgroup, persistent grid and one/two LDS buffers;
    5. consecutive tiles per workgroup turn and the workgroup-to-XCD order.
    Shapes above 4096 elements run only on the pair kernel; a classic
    variant of one is rejected by the library and skipped.
    A candidate counts only if its output is bit-identical to the default
    plan's (the reference's errors == 0 gate, generator.py:268-278).
    `incumbent` (the table's current parameters, if any) competes from the
    start, so a re-tune on another device keeps it unless something beats
    it there.  `depth` widens stages 2-4 to the best 2+depth, 1+depth and
    depth candidates of the stage before.  `log`, a list, receives every
    candidate measured as (parameters, seconds per call; inf if rejected)."""
    import torch
    from .api import sort
    t_start = time.time()
    bench = Bench(dims, dev)
    n = bench.n
    nbytes = 8 * n * (3 if beta else 2)
    reps = max(4, min(64, int(8e9 // max(nbytes, 1))))
    # reference output: the default plan (itself checked against the oracle)
    x = torch.rand(n, dtype=torch.float64, device=dev)
    y0 = torch.rand(n, dtype=torch.float64, device=dev)
    want = y0.clone()
    sort(x, want, dims, perm, 0.75, beta)
    timed = {}

    def measure(p):
        key = tuple(sorted(p.as_dict().items()))
        if key in timed:
            return timed[key][0]
        if time.time() - t_start > budget_s:
            return float("inf")
        try:
            got = y0.clone()
            sort(x, got, dims, perm, 0.75, beta, p)
        except ValueError:
            timed[key] = (float("inf"), p)
            return float("inf")
        if not torch.equal(got.view(torch.int64), want.view(torch.int64)):
            if verbose:
                print("  rejected (mismatch):", p.as_dict(), flush=True)
            timed[key] = (float("inf"), p)
            return float("inf")
        t = bench.time(dims, perm, beta, p, reps)
        timed[key] = (t, p)
        return t

    def best(k):
        r = sorted(timed.values(), key=lambda v: v[0])
        return [p for t, p in r[:k] if t < float("inf")]

    t_def = measure(TileParams.default())
    if incumbent is not None:
        measure(incumbent)
    for p in shape_candidates(dims, perm, beta, limit=shapes, big=max(4, shapes // 3)):
        measure(_with(p, flags=0))
    for p in best(2 + depth):
        for kern in (0, CLASSIC):
            for fl in ACCESS_FLAGS:
                measure(_with(p, flags=fl | kern))
    for p in best(1 + depth):
        if p.kind == 1:
            continue
        for mode in (0, 1, 2, 3):
            for w in (0, 2, 4, 8, 16, 32):
                if mode or w:
                    measure(_with(p, walk_mode=mode, walk_x=w))
    for p in best(depth):
        for blk in (128, 256, 512):
            for g in (0, 2, 3, 4):
                measure(_with(p, block=blk, grid_per_cu=g))
        if p.kind != 1 and p.flags >= 0 and not (p.flags & CLASSIC):
            for g in (0, 3, 4):
                measure(_with(p, flags=p.flags | LDS1, grid_per_cu=g))
            measure(_with(p, flags=p.flags | NOSPLIT))
    # 5. runs of consecutive tiles per workgroup turn, and the plain
    # workgroup-to-tile order instead of the XCD-aware one
    for p in best(depth):
        if p.kind == 1:
            continue
        for c in (2, 4, 8):
            measure(_with(p, walk_chunk=c))
        measure(_with(p, xcd_remap=0))
    if log is not None:
        log.extend((p.as_dict(), t) for t, p in timed.values())
    best_t, best_p = sorted(timed.values(), key=lambda v: v[0])[0]
    if best_t == float("inf"):
        best_t, best_p = t_def, TileParams.default()
    tried = sum(1 for t, _ in timed.values() if t < float("inf"))
    rec = {"key": dispatch_key(dims, perm, beta), **best_p.as_dict(),
           "gbs": round(nbytes / best_t / 1e9, 1), "default_gbs": round(nbytes / t_def / 1e9, 1),
           "dims": list(dims), "perm": list(perm), "acc": bool(beta), "tried": tried}
    if verbose:
        print(f"  {rec['key']}: best {rec['gbs']} GB/s (default {rec['default_gbs']}), "
              f"{tried} candidates, {time.time() - t_start:.1f} s", flush=True)
    return rec


# The 54 permutations NWChem's (T) code calls tce_sort_6 with
# (background/nwchem-perturbative-triples-usage:5-58).
TRIPLES_6D = [
    "261543", "265431", "326541", "362154", "362514", "362541", "421653", "431652", "432165",
    "432651", "436215", "436251", "436521", "463215", "463251", "463521", "521643", "531642",
    "532164", "532641", "541632", "542136", "542163", "542631", "543162", "543216", "543261",
    "543621", "621543", "625431", "631542", "632154", "632514", "632541", "641532", "642153",
    "642531", "643152", "643215", "643251", "643521", "651432", "652143", "652413", "652431",
    "653142", "653214", "653241", "653421", "654132", "654213", "654231", "654312", "654321",
]


def random_perms_6d(count: int = 64, seed: int = 2024) -> list[str]:
    """The seeded random 6D permutations of the 16^6 / 20^6 sweeps."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return ["".join(str(int(v) + 1) for v in rng.permutation(6)) for _ in range(count)]


def problems(which: str):
    """'4d': every 4D permutation at 20^4, 30^4, 100^4 and 10x30x20x40;
    '100': the 100^4 ones only; '6d': the (T) permutations at 10^6 and 16^6;
    '20_6': the first 8 seeded random permutations at 20^6; 'baseline':
    4d + 6d + 20_6; 'random6': all 64 seeded random permutations at 16^6
    and 20^6; 'nwchem4' / 'nwchem6': other tile sizes NWChem runs
    with (README.md:30-34: CCSD ~30, up to 100; CCSD(T) ~20; CCSDT < 20),
    every 4D permutation at 24^4, 32^4, 40^4 and 48^4 / the (T)
    permutations at 12^6, 14^6 and 18^6; 'nwchem4b' / 'nwchem6b': every
    4D permutation at 56^4, 64^4, 72^4 and 80^4 / the (T) permutations at
    15^6, 17^6 and 19^6.  Sort and accumulate each."""
    probs = []
    sizes4 = {"4d": ((20,) * 4, (30,) * 4, (100,) * 4, (10, 30, 20, 40)),
              "baseline": ((20,) * 4, (30,) * 4, (100,) * 4, (10, 30, 20, 40)),
              "100": ((100,) * 4,), "small4": ((20,) * 4, (30,) * 4, (10, 30, 20, 40)),
              "nwchem4": ((24,) * 4, (32,) * 4, (40,) * 4, (48,) * 4),
              "nwchem4b": ((56,) * 4, (64,) * 4, (72,) * 4, (80,) * 4)}
    for dims in sizes4.get(which, ()):
        for perm in itertools.permutations([1, 2, 3, 4]):
            for beta in (0.0, 1.0):
                probs.append((dims, perm, beta))
    sizes6 = {"6d": ((10,) * 6, (16,) * 6), "baseline": ((10,) * 6, (16,) * 6),
              "nwchem6": ((12,) * 6, (14,) * 6, (18,) * 6),
              "nwchem6b": ((15,) * 6, (17,) * 6, (19,) * 6)}
    if which in sizes6:
        for dims in sizes6[which]:
            for p in TRIPLES_6D:
                for beta in (0.0, 1.0):
                    probs.append((dims, tuple(int(c) for c in p), beta))
    if which in ("20_6", "baseline"):
        for p in random_perms_6d()[:8]:
            for beta in (0.0, 1.0):
                probs.append(((20,) * 6, tuple(int(c) for c in p), beta))
    if which == "random6":
        for dims in ((16,) * 6, (20,) * 6):
            for p in random_perms_6d():
                for beta in (0.0, 1.0):
                    probs.append((dims, tuple(int(c) for c in p), beta))
    return probs


def _takes_blocks(dims, perm, beta) -> bool:
    try:
        plan(dims, perm, beta, TileParams.default(tw=2))
    except ValueError:
        try:
            plan(dims, perm, beta, TileParams.default(tw=-2))
        except ValueError:
            return False
    return True


def write_table(records, path, replace=False):
    """Merges records into the table: a key keeps the faster record, or,
    with replace, the new one."""
    old = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line:
                    r = json.loads(line)
                    old[r["key"]] = r
    for r in records:
        prev = old.get(r["key"])
        if prev is None or replace or r["gbs"] >= prev.get("gbs", 0):
            old[r["key"]] = r
    with open(path, "w") as f:
        for k in sorted(old):
            f.write(json.dumps(old[k]) + "\n")


def main():
    global ROTATE_BYTES
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", default="baseline")
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--budget", type=float, default=15.0, help="seconds per problem")
    ap.add_argument("--limit", type=int, default=0, help="tune only the first N problems")
    ap.add_argument("--shard", default="0/1", help="i/n: tune every n-th problem from i")
    ap.add_argument("--replace", action="store_true",
                    help="new records replace the table's (else the faster one stays)")
    ap.add_argument("--depth", type=int, default=1,
                    help="candidates carried between stages (1: 3/2/1; 2: 4/3/2)")
    ap.add_argument("--incumbent", default="",
                    help="table whose entries compete in each problem's search")
    ap.add_argument("--shapes", type=int, default=12, help="tile shapes tried in stage 1")
    ap.add_argument("--missing", default="",
                    help="tune only the problems whose keys this table lacks")
    ap.add_argument("--rotate-mb", type=int, default=0,
                    help="device memory the timed calls rotate over (default: 512 for outputs "
                         "that fit in the 256 MB last-level cache, written again while cached as "
                         "by a loop reusing its arrays; 4096 for larger ones; 4096 for all tunes "
                         "every write cold, as for the streamed-output table)")
    ap.add_argument("--blocks", action="store_true",
                    help="tune only the problems whose two fastest axes swap places "
                         "(those that take tiles of several blocks)")
    a = ap.parse_args()
    ROTATE_BYTES = a.rotate_mb << 20
    import torch
    os.environ.setdefault("SPG_DISPATCH_TABLE", "none")
    _lib.lib()
    dev = torch.device("cuda:0")
    probs = problems(a.problems)
    i, n = (int(v) for v in a.shard.split("/"))
    probs = probs[i::n]
    if a.limit:
        probs = probs[: a.limit]
    # one key per canonical problem
    seen, todo = set(), []
    if a.missing and os.path.exists(a.missing):
        with open(a.missing) as f:
            seen = {json.loads(line)["key"] for line in f if line.strip()}
    if a.blocks:
        probs = [(d, p, b) for d, p, b in probs if _takes_blocks(d, p, b)]
    for d, p, b in probs:
        k = dispatch_key(d, p, b)
        if k not in seen:
            seen.add(k)
            todo.append((d, p, b))
    inc = {}
    if a.incumbent:
        with open(a.incumbent) as f:
            for line in f:
                line = line.strip()
                if line:
                    r = json.loads(line)
                    inc[r["key"]] = TileParams.from_record(r)
    print(f"tuning {len(todo)} problems", flush=True)
    recs = []
    for d, p, b in todo:
        recs.append(tune_problem(d, p, b, dev, a.budget,
                                 incumbent=inc.get(dispatch_key(d, p, b)), depth=a.depth,
                                 shapes=a.shapes))
        write_table(recs, a.out, a.replace)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
