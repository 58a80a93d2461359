"""Runs one sort configuration a few times, after a plain device copy of known
size, for hardware-counter collection:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR -o fetch -- \
        python -m spaghetty_amd.profile_kernel --dims 100 100 100 100 --perm 4 3 2 1
    python -m spaghetty_amd.profile_kernel --summarize DIR/fetch_counter_collection.csv \
        --dims 100 100 100 100 --beta 0

The copy moves --copy-mb MB in and out, so its FETCH_SIZE / WRITE_SIZE per
dispatch calibrate the counters (bytes the counter reports per algorithmic
byte); the sort's counters divided by that factor give its HBM traffic per
launch, to set beside the algorithmic 8*N*(2+acc) bytes.

--rotate N spreads the calls over N input/output array pairs (call k uses
pair k mod N): outputs rewritten within the 256 MB last-level cache's reach
behave differently from outputs written cold.

With --batch-mix the sorts are replaced by the batched form on the 4D tiles
of the 4,096-tile coupled-cluster mix (spaghetty_amd.workloads.cc_tile_mix,
seed 7): one committed batch launched --reps times (--batch-flags as
spg_batch_set_flags); --summarize then takes the batch's algorithmic bytes
from --alg-bytes.
"""
from __future__ import annotations

import argparse
import csv
import json
import sys


def run(a):
    import torch
    import spaghetty_amd as S
    from spaghetty_amd import _lib
    _lib.lib()
    dev = torch.device("cuda:0")
    n = 1
    for d in a.dims:
        n *= d
    ncopy = int(a.copy_mb * 1e6 // 8)
    src = torch.rand(ncopy, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    if a.batch_mix:
        run_batch_mix(a, torch, S, dev)
        return
    rot_out = a.rotate_out or a.rotate
    xs = [torch.rand(n, dtype=torch.float64, device=dev) for _ in range(a.rotate)]
    ys = [torch.rand(n, dtype=torch.float64, device=dev) for _ in range(rot_out)]
    perms = [a.perm]
    if a.all_perms:
        import itertools
        perms = [list(p) for p in itertools.permutations(range(1, len(a.dims) + 1))]
    if a.perm_list:
        perms = [[int(c) for c in p] for p in a.perm_list.split(",")]
    if a.triples:
        from .autotune import TRIPLES_6D
        perms = [[int(c) for c in p] for p in TRIPLES_6D]
    i = 0
    for perm in perms:
        params = None
        if a.params:
            # explicit tile parameters (spg_tile_params fields; "tw" 1 when absent)
            from . import TileParams
            params = TileParams.from_record(json.loads(a.params))
        elif a.flags is not None:
            # the dispatch table's plan for this problem with other access flags
            params = _table_params(a.dims, perm, a.beta)
            params.flags = a.flags
        for _ in range(a.reps):
            S.sort(xs[i % a.rotate], ys[i % rot_out], a.dims, perm, 1.0, a.beta, params)
            i += 1
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": 8 * ncopy, "sort_elems": n, "reps": a.reps,
                      "perms": len(perms), "rotate": a.rotate, "rotate_out": rot_out, "flags": a.flags}))


def _table_params(dims, perm, beta):
    from . import TileParams
    from .api import dispatch_key
    from .autotune import DEFAULT_TABLE
    key = dispatch_key(dims, perm, beta)
    with open(DEFAULT_TABLE) as f:
        for line in f:
            if line.strip():
                r = json.loads(line)
                if r["key"] == key:
                    return TileParams.from_record(r)
    return TileParams.default()


def run_batch_mix(a, torch, S, dev):
    from spaghetty_amd.workloads import cc_tile_mix, packed_offsets
    tiles = [t for t in cc_tile_mix(4096, seed=7) if len(t.dims) == 4]
    offs, total = packed_offsets(tiles)  # 256-byte aligned arrays, as the bench lays them out
    inp = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    out = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    b = S.Batch()
    for t, o in zip(tiles, offs):
        b.add(inp[o:o + t.nelem], out[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta)
    if a.batch_flags is not None:
        b.set_flags(a.batch_flags)
    b.commit()
    for _ in range(a.reps):
        b.launch()
    torch.cuda.synchronize()
    print(json.dumps({"batch": b.info(), "reps": a.reps}))
    b.close()


def summarize(a):
    n = 1
    for d in a.dims:
        n *= d
    vals = {}
    with open(a.summarize) as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name", "")
            cnt = r.get("Counter_Name", "")
            v = float(r.get("Counter_Value", 0))
            did = r.get("Dispatch_Id", "")
            ours = any(k in name for k in ("k_vtile", "k_tiled", "k_rows", "k_batch"))
            copy = "copyBuffer" in name or "copy_kernel" in name
            vals.setdefault((cnt, ours, copy), {}).setdefault(did, 0.0)
            vals[(cnt, ours, copy)][did] += v
    ncopy_bytes = int(a.copy_mb * 1e6 // 8) * 8
    out = {}
    for (cnt, ours, copy), per in vals.items():
        avg = sum(per.values()) / max(len(per), 1)
        if ours:
            out.setdefault(cnt, {})["sort_kib"] = avg
        elif copy:
            out.setdefault(cnt, {})["copy_kib"] = avg
    alg = a.alg_bytes if a.alg_bytes else 8 * n * (3 if a.beta else 2)
    for cnt, d in out.items():
        if d.get("copy_kib") and "sort_kib" in d:
            # the copy reads (or writes) ncopy_bytes: bytes per counted KiB
            factor = ncopy_bytes / (d["copy_kib"] * 1024.0)
            d["scale"] = factor
            d["sort_bytes"] = d["sort_kib"] * 1024.0 * factor
    fetch = out.get("FETCH_SIZE", {}).get("sort_bytes")
    write = out.get("WRITE_SIZE", {}).get("sort_bytes")
    res = {"counters": out, "algorithmic_bytes": alg}
    if fetch is not None:
        res["fetch_bytes"] = fetch
    if write is not None:
        res["write_bytes"] = write
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs="+", default=[100, 100, 100, 100])
    ap.add_argument("--perm", type=int, nargs="+", default=[4, 3, 2, 1])
    ap.add_argument("--beta", type=float, default=0.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mb", type=float, default=800.0)
    ap.add_argument("--all-perms", action="store_true",
                    help="every permutation of the rank (the bench's dominant-kernel set)")
    ap.add_argument("--perm-list", default="",
                    help="comma-separated permutations, e.g. 261543,265431 (in this order)")
    ap.add_argument("--triples", action="store_true",
                    help="the 54 (T) permutations (6D), in the bench's order")
    ap.add_argument("--rotate", type=int, default=1,
                    help="input/output array pairs the calls rotate over (call k uses pair k mod N)")
    ap.add_argument("--rotate-out", type=int, default=0,
                    help="output arrays the calls rotate over, if not --rotate (call k writes k mod N)")
    ap.add_argument("--flags", type=int, default=None,
                    help="the dispatch table's plans with these access flags instead")
    ap.add_argument("--params", default="",
                    help='explicit tile parameters as JSON, e.g. \'{"kind": 0, "tw": 8, "flags": 4}\'')
    ap.add_argument("--batch-mix", action="store_true",
                    help="profile the batched form on the 4D coupled-cluster tile mix")
    ap.add_argument("--batch-flags", type=int, default=None)
    ap.add_argument("--alg-bytes", type=int, default=0,
                    help="algorithmic bytes per launch for --summarize (default: from --dims)")
    ap.add_argument("--summarize", default="")
    a = ap.parse_args()
    if a.summarize:
        summarize(a)
    else:
        run(a)


if __name__ == "__main__":
    sys.exit(main())
