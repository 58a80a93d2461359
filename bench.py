#!/usr/bin/env python3
"""Benchmark of the tce_sort_4 / tce_sortacc_4 path on one MI355X.

One step = the 4D sweep of the generated library's configurations: all 24
permutations x tile sizes {20^4, 30^4, 100^4, 10x30x20x40} x factor
{1, -1, 0.5} x {sort, sort-accumulate}: 576 calls of the HIP kernels on FP64
arrays already resident in HBM.  `value` is the sweep's algorithmic HBM
traffic (8*N*(2+acc) bytes per call) divided by the step time, i.e. the
per-call aggregate GB/s, and `pct_of_peak` relates it to the MI355X's
8 TB/s.  Device arrays are rotated over a pool of >= 2 GB per size (outputs
>= 1 GB, four times the 256 MB last-level cache), so neither reads nor
writes are served by that cache.

Also reported on the same line:
  roofline      the dominant kernel (the 100^4 sorts), timed per launch with
                HIP events inside the timed region;
  per_size      GB/s per tile size and mode from the same events;
  sweep_6d      tce_sort_6 per call: the (T) permutations at 16^6 and random
                permutations at 20^6;
  ttdt          the contraction around the sort (ccsd_t2_8: sorts + DGEMM +
                sort-accumulate on the device) at tilesizes 20, 30, 60;
  batched       4,096 uneven coupled-cluster tiles (the 4D subset in one
                persistent launch vs one launch per tile; graph replay);
  serving       Poisson request streams with the same tile mix at three
                offered loads (continuous batching under a 64 GB budget):
                p50/p99 latency and GB/s;
  config1       the reference's own CPU-runnable case (30^4, 4321, factor 1,
                overwrite) per call on the GPU, beside the reference's
                old_sort_4 on one host core and the restated library kernel
                on all of them;
  cpu_baseline  the reference's library kernel trans_<P>_loop_<L>_omp_any_f
                (restated in oracle/, OpenMP, the packaged loop table) on the
                host's cores, on a bounded sample of the same sweep.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "achieved HBM GB/s (% of peak), batched vs per-call; p99 latency, bit-exact"
PEAK_GBS = 8000.0
SIZES = [(20, 20, 20, 20), (30, 30, 30, 30), (100, 100, 100, 100), (10, 30, 20, 40)]
FACTORS = (1.0, -1.0, 0.5)
PERMS = list(itertools.permutations([1, 2, 3, 4]))
POOL_BYTES = 512 << 20      # host arrays of the CPU legs
# Device arrays: outputs rotating over half of this are never still held in
# the device's 256 MB last-level cache (MALL) when they are written again
GPU_POOL_BYTES = 2 << 30
# the 6D sweep's ":warm" lines: 16^6 outputs are written again two calls
# later, while much of them is still in that cache
WARM_POOL_BYTES = 512 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--calls-log", default="", help="write the ordered launch list (JSON)")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    return ap.parse_args()


class Pool:
    """Rotating device input/output arrays of one size (footprint >= 2 GB,
    or `min_bytes`)."""

    def __init__(self, torch, dims, dev, min_bytes=None):
        n = int(np.prod(dims))
        self.n = n
        pair = 2 * 8 * n
        self.count = max(1, -(-(min_bytes or GPU_POOL_BYTES) // pair))
        self.inp = torch.rand(self.count, n, dtype=torch.float64, device=dev) * 2 - 1
        self.out = torch.rand(self.count, n, dtype=torch.float64, device=dev) * 2 - 1
        # device addresses of every array, so a call costs no tensor indexing
        self.ptrs = [(self.inp.data_ptr() + 8 * n * k, self.out.data_ptr() + 8 * n * k)
                     for k in range(self.count)]
        self.i = 0

    def next(self):
        k = self.i
        self.i = (self.i + 1) % self.count
        return self.ptrs[k]


def sweep_calls():
    calls = []
    for dims in SIZES:
        for perm in PERMS:
            for f in FACTORS:
                for beta in (0.0, 1.0):
                    calls.append((dims, perm, f, beta))
    return calls


class Caller:
    """Issues the sweep through the C ABI (spg_sort), the call a Fortran or C
    caller of the library makes, with argument arrays built once."""

    def __init__(self, torch, calls, pools):
        import ctypes
        from spaghetty_amd import _lib
        self.L = _lib.lib()
        self.stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.args = [(_lib.ints(d), _lib.ints(p), len(d), float(f), float(b), pools[d])
                     for d, p, f, b in calls]

    def run(self, events=None):
        L, st = self.L, self.stream
        for idx, (d, p, r, f, b, pool) in enumerate(self.args):
            x, y = pool.next()
            if events is not None and idx in events:
                e0, e1 = events[idx]
                e0.record()
                rc = L.spg_sort(x, y, r, d, p, f, b, st)
                e1.record()
            else:
                rc = L.spg_sort(x, y, r, d, p, f, b, st)
            if rc:
                raise RuntimeError(f"spg_sort failed with status {rc}")


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size),
    or None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "tile_kernels_tce_sort_4_100x4_counters.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def _host_model():
    try:
        return [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                if ln.startswith("model name")][0]
    except Exception:
        return "unknown"


def _loop_orders_6d(perm):
    """Loop orders (outermost -> innermost, 1-based input axes) tried for a
    6D permutation: the reference times all 720 (trunk/generator6.py:223-392);
    this bounded search keeps the 240 whose innermost loop is the input's
    unit-stride axis (6) or the output's (perm[-1])."""
    inner = sorted({6, int(perm[-1])})
    out = []
    for k in inner:
        rest = [a for a in range(1, 7) if a != k]
        out += [tuple(o) + (k,) for o in itertools.permutations(rest)]
    return out


def cpu_baseline(seconds: float) -> dict:
    """The reference's CPU library kernel, trans_<P>_loop_<L>_omp_any_f
    (trunk/generator.py:107-159; 6D: generator6.py:137-220), restated in
    oracle/ and built -O3 -fopenmp -ffp-contract=off, OpenMP over the host's
    cores (OMP_PROC_BIND=close, OMP_PLACES=cores), timed on a bounded sample
    of the GPU workload:
      4D  the sweep's four sizes x 24 permutations x {sort, accumulate},
          factor 0.5, loop orders from the packaged table (makelib.py:39-62),
          arrays rotated over >= 512 MB; `value` is this sample's aggregate
          GB/s (bytes summed over calls / time summed), as the GPU's is;
      6D  three (T) permutations at 16^6 with a bounded loop-order search
          (the best of 240 orders, as generator6's testers pick theirs), the
          best order then timed at 16^6 and 20^6;
    plus NWChem's single-thread old_sort_4 (the reference's speedup
    baseline), compiled from the reference where it was built."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    threads = O.max_threads()
    t_start = time.time()
    slices, tot_b, tot_t, calls = {}, 0, 0.0, 0
    # 4D: each size until its share of the time is spent (at least one pass)
    share = {(20,) * 4: 0.10, (30,) * 4: 0.15, (100,) * 4: 0.20, (10, 30, 20, 40): 0.10}
    for dims in SIZES:
        n = int(np.prod(dims))
        pairs = max(1, -(-POOL_BYTES // (16 * n)))
        xs = [O.uniform_c(n, 100 + i) for i in range(pairs)]
        ys = [O.uniform_c(n, 300 + i) for i in range(pairs)]
        b_sz, t_sz, c_sz, k = 0, 0.0, 0, 0
        t_end = time.time() + seconds * share[dims]
        while True:
            for perm, loop in O.KNIGHT_4D.items():
                p = [int(c) for c in perm]
                lp = [int(c) for c in loop]
                for beta in (0.0, 1.0):
                    x, y = xs[k % pairs], ys[k % pairs]
                    k += 1
                    t0 = time.perf_counter()
                    O.trans(x, y, dims, p, lp, 0.5, 0.0, int(beta), threads)
                    t_sz += time.perf_counter() - t0
                    b_sz += 8 * n * (3 if beta else 2)
                    c_sz += 1
            if time.time() >= t_end:
                break
        slices["x".join(map(str, dims))] = {"gbs": round(b_sz / t_sz / 1e9, 2), "calls": c_sz,
                                            "seconds": round(t_sz, 2)}
        tot_b, tot_t, calls = tot_b + b_sz, tot_t + t_sz, calls + c_sz
        del xs, ys
    value = tot_b / tot_t / 1e9
    # 6D: bounded loop-order search at 16^6, then the best order timed
    six = {}
    for perm in ((6, 5, 4, 3, 2, 1), (4, 5, 6, 1, 2, 3), (3, 2, 1, 6, 5, 4)):
        best = None
        for dims in ((16,) * 6, (20,) * 6):
            n = int(np.prod(dims))
            x, y = O.uniform_c(n, 7), O.uniform_c(n, 8)
            if best is None:
                tb = []
                for lp in _loop_orders_6d(perm):
                    t0 = time.perf_counter()
                    O.trans(x, y, dims, perm, lp, 1.0, 0.0, 0, threads)
                    tb.append((time.perf_counter() - t0, lp))
                best = min(tb)[1]
                searched = len(tb)
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                O.trans(x, y, dims, perm, best, 1.0, 0.0, 0, threads)
            dt = (time.perf_counter() - t0) / reps
            six[f"{'x'.join(map(str, dims))}:{''.join(map(str, perm))}"] = {
                "gbs": round(16 * n / dt / 1e9, 2), "ms_per_call": round(dt * 1e3, 2),
                "loop": "".join(map(str, best)), "orders_searched": searched}
            del x, y
    # single-thread NWChem original (the reference's speedup baseline)
    use_ref = O.ref_available()
    dims = (30, 30, 30, 30)
    n = 30 ** 4
    pairs = max(1, -(-POOL_BYTES // (16 * n)))
    xs = [O.uniform_c(n, 500 + i) for i in range(pairs)]
    ys = [O.uniform_c(n, 700 + i) for i in range(pairs)]
    t_used1, bytes1, calls1 = 0.0, 0, 0
    t_end = time.time() + seconds * 0.15
    while time.time() < t_end:
        for perm in list(O.KNIGHT_4D)[::3]:
            p = [int(c) for c in perm]
            x, y = xs[calls1 % pairs], ys[calls1 % pairs]
            t0 = time.perf_counter()
            if use_ref:
                O.ref_old_sort(x, y, dims, p, 0.5, False)
            else:
                O.old_sort(x, y, dims, p, 0.5, False)
            t_used1 += time.perf_counter() - t0
            bytes1 += 16 * n
            calls1 += 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except Exception:
        affinity = os.cpu_count()
    return {
        "value": round(value, 2), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": f"4D sweep sample: sizes 20^4, 30^4, 100^4, 10x30x20x40 x 24 perms x {{sort, "
                  f"accumulate}}, factor 0.5, {calls} calls, {tot_t:.1f} s, arrays rotated over >= 512 MB, "
                  f"loop orders of trunk/makelib.py:39-62; OpenMP {threads} threads "
                  f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, OMP_PLACES={os.environ.get('OMP_PLACES')}), "
                  f"{affinity} host CPUs available, host {_host_model()}; gcc -O3 -fopenmp -ffp-contract=off",
        "slices_4d": slices,
        "six_d": six,
        "single_thread_old_sort": {
            "value": round(bytes1 / t_used1 / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "reference" if use_ref else "port",
            "sample": f"old_sort_4 30^4, {calls1} calls, {t_used1:.1f} s"},
        "seconds": round(time.time() - t_start, 1),
    }


class CallList:
    """One spg_sort per tile through the C ABI, argument arrays built once:
    the per-call leg of the batched comparisons (the call a Fortran or C
    caller makes per tile)."""

    def __init__(self, torch, tiles, ptrs):
        import ctypes
        from spaghetty_amd import _lib
        self.L = _lib.lib()
        self.st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.args = [(x, y, len(t.dims), _lib.ints(t.dims), _lib.ints(t.perm), float(t.factor),
                      float(t.beta)) for t, (x, y) in zip(tiles, ptrs)]

    def run(self):
        L, st = self.L, self.st
        for x, y, r, d, p, f, b in self.args:
            rc = L.spg_sort(x, y, r, d, p, f, b, st)
            if rc:
                raise RuntimeError(f"spg_sort failed with status {rc}")


def config1_bench(torch, dev, cpu: bool) -> dict:
    """The reference's own CPU-runnable case: tce_sort_4 on a 30x30x30x30
    FP64 array, permutation 4321, factor 1.0, overwrite.  GPU: one spg_sort
    per call through the C ABI, back to back on rotating arrays (>= 2 GB),
    and the same calls replayed from a graph; CPU (rank 0 host): the
    reference's old_sort_4 on one thread (oracle/_ref, as compiled from
    trunk/old_sort.f) and the restated library kernel with its packaged loop
    order on every core."""
    import ctypes
    from spaghetty_amd import _lib
    dims, perm, n = (30, 30, 30, 30), (4, 3, 2, 1), 30 ** 4
    nbytes = 16 * n
    pool = Pool(torch, dims, dev)
    L = _lib.lib()
    st = torch.cuda.Stream(device=dev)
    sp = ctypes.c_void_p(st.cuda_stream)
    d, p = _lib.ints(dims), _lib.ints(perm)
    reps = 4 * pool.count

    def calls():
        for _ in range(reps):
            x, y = pool.next()
            rc = L.spg_sort(x, y, 4, d, p, 1.0, 0.0, sp)
            if rc:
                raise RuntimeError(f"spg_sort failed with status {rc}")
    with torch.cuda.stream(st):
        calls()
        t_call = _timed(torch, calls, 3, st) / reps
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            calls()
        g.replay()
        t_graph = _timed(torch, g.replay, 3, st) / reps
    out = {"case": "tce_sort_4 30x30x30x30, perm 4321, factor 1.0, overwrite", "bytes_per_call": nbytes,
           "gpu_us_per_call": round(t_call * 1e6, 2), "gpu_gbs": round(nbytes / t_call / 1e9, 1),
           "gpu_graph_us_per_call": round(t_graph * 1e6, 2),
           "gpu_graph_gbs": round(nbytes / t_graph / 1e9, 1)}
    del pool, g
    torch.cuda.empty_cache()
    if cpu:
        from oracle import oracle as O
        threads = O.max_threads()
        pairs = max(1, -(-POOL_BYTES // nbytes))
        xs = [O.uniform_c(n, 900 + i) for i in range(pairs)]
        ys = [O.uniform_c(n, 950 + i) for i in range(pairs)]
        loop = [int(c) for c in O.KNIGHT_4D["4321"]]
        use_ref = O.ref_available()
        for name, fn in (("cpu_reference_1_thread", lambda x, y: (O.ref_old_sort if use_ref else O.old_sort)(
                             x, y, dims, list(perm), 1.0, False)),
                         ("cpu_port_all_cores", lambda x, y: O.trans(x, y, dims, list(perm), loop, 1.0, 0.0, 0,
                                                                     threads))):
            k, t = 0, 0.0
            t_end = time.time() + 2.0
            while time.time() < t_end:
                x, y = xs[k % pairs], ys[k % pairs]
                t0 = time.perf_counter()
                fn(x, y)
                t += time.perf_counter() - t0
                k += 1
            out[name] = {"us_per_call": round(t / k * 1e6, 1), "gbs": round(nbytes * k / t / 1e9, 2),
                         "calls": k, "threads": 1 if "1_thread" in name else threads,
                         "kind": ("reference" if use_ref else "port") if "1_thread" in name else "port"}
        out["gpu_over_cpu_reference"] = round(out["cpu_reference_1_thread"]["us_per_call"] / out["gpu_us_per_call"], 1)
    return out


def _timed(torch, fn, reps, stream=None):
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def batched_bench(S, torch, dev, reps=5, budget=64 << 30) -> dict:
    """The 4,096-tile coupled-cluster mix: 70% 4D / 30% 6D, occupied dims
    U[8,20], virtual U[16,40], random perm, factor from {+-1, +-0.5},
    accumulate on 50%.
      small_4d  its 2,837 4D tiles (the small uneven ones) in one persistent
                launch (plain and as a replayed hipGraph);
      full_mix  all 4,096 tiles in waves under a 64 GB budget (the 6D tiles
                average 482 MB, ~1.2 TB in all): each wave reuses one 32 GB
                input and one 32 GB output region, and runs as one persistent
                launch over its tiles below 2^22 elements plus one launch per
                larger tile on its per-call plan; the whole job is one
                captured hipGraph.
    Both are compared with one spg_sort per tile through the C ABI on the
    same arrays and stream."""
    import ctypes
    from spaghetty_amd.workloads import cc_tile_mix, packed_offsets, waves_by_budget
    out = {}
    mix = cc_tile_mix(4096, seed=7)
    # -- the 4D tiles in one launch; every tile's arrays start 256-byte
    # aligned, as separately allocated device arrays (and the full mix's
    # waves) do
    tiles = [t for t in mix if len(t.dims) == 4]
    offs, total = packed_offsets(tiles)
    inp = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    outp = torch.rand(total, dtype=torch.float64, device=dev) * 2 - 1
    ptrs = [(inp.data_ptr() + 8 * o, outp.data_ptr() + 8 * o) for o in offs]
    b = S.Batch()
    for t, o in zip(tiles, offs):
        b.add(inp[o:o + t.nelem], outp[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta)
    b.commit()
    nbytes = b.bytes
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        b.launch(stream)
        t_batch = _timed(torch, lambda: b.launch(stream), reps, stream)
        b.graph_launch(stream)
        t_graph = _timed(torch, lambda: b.graph_launch(stream), reps, stream)
        calls = CallList(torch, tiles, ptrs)
        calls.run()
        t_call = _timed(torch, calls.run, reps, stream)
    info = b.info()
    b.close()
    gbs = lambda t: nbytes / t / 1e9  # noqa: E731
    out["small_4d"] = {
        "tiles": len(tiles), "of": len(mix), "bytes_per_launch": nbytes, "grid": info["grid"],
        "batched_gbs": round(gbs(t_batch), 1), "graph_gbs": round(gbs(t_graph), 1),
        "per_call_gbs": round(gbs(t_call), 1),
        "batched_over_per_call": round(t_call / min(t_batch, t_graph), 2),
        "batched_ms": round(t_batch * 1e3, 3), "per_call_ms": round(t_call * 1e3, 3)}
    del inp, outp
    torch.cuda.empty_cache()
    # -- the full mix in waves under the budget
    place = waves_by_budget(mix, budget)
    half = budget // 2 // 8
    inp = torch.empty(half, dtype=torch.float64, device=dev)
    outp = torch.empty(half, dtype=torch.float64, device=dev)
    for k in range(0, half, 1 << 28):  # fill in slices: no full-size temporaries
        inp[k:k + (1 << 28)].uniform_(-1, 1)
        outp[k:k + (1 << 28)].uniform_(-1, 1)
    b = S.Batch()
    wave = 0
    ptrs = []
    for t, (w, o) in zip(mix, place):
        if w != wave:
            b.next_wave()
            wave = w
        b.add(inp[o:o + t.nelem], outp[o:o + t.nelem], t.dims, t.perm, t.factor, t.beta)
        ptrs.append((inp.data_ptr() + 8 * o, outp.data_ptr() + 8 * o))
    b.commit()
    info = b.info()
    nbytes = b.bytes
    r2 = max(1, reps // 2)
    with torch.cuda.stream(stream):
        b.graph_launch(stream)
        t_graph = _timed(torch, lambda: b.graph_launch(stream), r2, stream)
        calls = CallList(torch, mix, ptrs)
        calls.run()
        t_call = _timed(torch, calls.run, r2, stream)
    b.close()
    out["full_mix"] = {
        "tiles": len(mix), "tiles_6d": sum(len(t.dims) == 6 for t in mix), "waves": info["waves"],
        "budget_bytes": budget, "big_tasks": info["big_tasks"], "bytes_per_job": nbytes,
        "graph_gbs": round(nbytes / t_graph / 1e9, 1), "per_call_gbs": round(nbytes / t_call / 1e9, 1),
        "batched_over_per_call": round(t_call / t_graph, 3),
        "graph_ms": round(t_graph * 1e3, 2), "per_call_ms": round(t_call * 1e3, 2)}
    del inp, outp
    torch.cuda.empty_cache()
    return out


def serving_bench(torch, dev, n_req=2000, budget=64 << 30, slots=4, pool_bytes=40 << 30) -> dict:
    """Poisson streams of requests with the coupled-cluster tile mix (70% 4D /
    30% 6D, the batched distribution), outputs allocated by the server inside
    a 64 GB budget.  Every request reads its own input: consecutive slices of
    a 40 GB pool (more than the inputs in flight under the budget), so
    requests on the device at the same time share no data through the L2 or
    the 256 MB Infinity Cache.  The saturated throughput (every request
    present at t = 0) sets the capacity; the three offered loads are 0.3 /
    0.6 / 0.9 of it."""
    from spaghetty_amd.serving import Server, poisson_arrivals
    from spaghetty_amd.workloads import cc_tile_mix
    tiles = cc_tile_mix(n_req, seed=11)
    npool = pool_bytes // 8
    pool = torch.empty(npool, dtype=torch.float64, device=dev)
    for k in range(0, npool, 1 << 28):
        pool[k:k + (1 << 28)].uniform_(-1, 1)
    offs, o = [], 0
    for t in tiles:
        if o + t.nelem > npool:
            o = 0
        offs.append(o)
        o += t.nelem + (-t.nelem) % 32
    mean_bytes = sum(t.bytes for t in tiles) / len(tiles)

    def run(arrivals):
        srv = Server(budget, slots=slots, max_batch=256)
        for t, a, o in zip(tiles, arrivals, offs):
            srv.submit(pool[o:o + t.nelem], None, t.dims, t.perm, t.factor, t.beta, arrival_us=float(a))
        srv.start()
        srv.wait(600)
        st = srv.stats()
        srv.close()
        return st

    sat = run(np.zeros(len(tiles)))
    cap = sat["gbs"]
    out = {"requests": n_req, "budget_bytes": budget, "slots": slots, "input_pool_bytes": pool_bytes,
           "mean_request_bytes": round(mean_bytes), "saturated": sat}
    for load in (0.3, 0.6, 0.9):
        rate = load * cap * 1e3 / mean_bytes  # requests per microsecond
        st = run(poisson_arrivals(n_req, rate, seed=int(load * 10)))
        out[f"load_{load}"] = {"offered_gbs": round(load * cap, 1), **st}
    del pool
    torch.cuda.empty_cache()
    return out


def host_dropin(reps=5) -> dict:
    """tce_sort_4 through the NWChem drop-in with HOST arrays (the reference's
    call: host -> device copy, kernel, device -> host copy, synchronous):
    the PCIe-inclusive rate, reported beside the HBM-resident figures."""
    import spaghetty_amd as S
    from oracle import oracle as O
    out = {}
    for dims in ((30, 30, 30, 30), (100, 100, 100, 100)):
        n = int(np.prod(dims))
        x = O.uniform(n, 3)
        y = np.zeros(n)
        S.tce_sort_4(x, y, *dims, 4, 3, 2, 1, 1.0)
        t0 = time.perf_counter()
        for _ in range(reps):
            S.tce_sort_4(x, y, *dims, 4, 3, 2, 1, 1.0)
        dt = (time.perf_counter() - t0) / reps
        out["x".join(map(str, dims))] = {"ms_per_call": round(dt * 1e3, 3),
                                          "gbs_algorithmic": round(16 * n / dt / 1e9, 2)}
    return out


def sweep_6d(torch, dev, reps=3) -> dict:
    """tce_sort_6 per call, HBM-resident, sort and accumulate, factor 1, on
    rotating buffers: the 54 (T) permutations at 16^6, and the 64 seeded
    random permutations of the 6D configuration at 16^6 and at 20^6.  The
    calls of one pass over the permutations run back to back between two HIP
    events, so the figure is the per-call rate of a caller issuing them in a
    row (no synchronisation between calls).  `tuned` counts the problems
    the dispatch table holds.

    The arrays rotate over >= 2 GB, so every output is written with none of
    it still in the 256 MB last-level cache (MALL).  The ":warm" lines
    rotate 16^6 arrays over 512 MB instead: each output (134 MB) is then
    written again two calls later while much of it is still in that cache,
    and writes that hit there are cheaper (a loop that reuses its output
    arrays runs so).  The ":streamed" lines rotate over >= 2 GB like the
    plain ones, with the library told that outputs are written cold
    (set_output_reuse(False): the plans tuned for that regime)."""
    import ctypes
    from spaghetty_amd import _lib
    from spaghetty_amd.api import dispatch_key
    from spaghetty_amd.autotune import DEFAULT_TABLE, TRIPLES_6D, random_perms_6d
    L = _lib.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with open(DEFAULT_TABLE) as f:
        table = {json.loads(line)["key"] for line in f if line.strip()}
    out = {}
    import spaghetty_amd as S
    for name, dims, perms, var in (("triples", (16,) * 6, TRIPLES_6D, ""),
                                   ("random", (16,) * 6, random_perms_6d(), ""),
                                   ("random", (20,) * 6, random_perms_6d(), ""),
                                   ("triples", (16,) * 6, TRIPLES_6D, "warm"),
                                   ("random", (16,) * 6, random_perms_6d(), "warm"),
                                   ("triples", (16,) * 6, TRIPLES_6D, "streamed"),
                                   ("random", (16,) * 6, random_perms_6d(), "streamed")):
        pool = Pool(torch, dims, dev, WARM_POOL_BYTES if var == "warm" else None)
        S.set_output_reuse(var != "streamed")
        n = pool.n
        for beta in (0.0, 1.0):
            d = _lib.ints(dims)
            args = [_lib.ints([int(c) for c in p]) for p in perms]
            for p in args:  # warm the plan cache
                x, y = pool.next()
                L.spg_sort(x, y, 6, d, p, 1.0, beta, st)
            torch.cuda.synchronize()
            tot, cnt = 0.0, 0
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for p in args:
                    x, y = pool.next()
                    L.spg_sort(x, y, 6, d, p, 1.0, beta, st)
                e1.record()
                e1.synchronize()
                tot += e0.elapsed_time(e1)
                cnt += len(args)
            gbs = 8 * n * (3 if beta else 2) / (tot / cnt / 1e3) / 1e9
            tuned = sum(dispatch_key(dims, [int(c) for c in p], beta) in table for p in perms)
            out[f"{'x'.join(map(str, dims))}:{name}:{'acc' if beta else 'sort'}{':' + var if var else ''}"] = {
                "perms": len(perms), "tuned": tuned, "gbs": round(gbs, 1),
                "pct_of_peak": round(100 * gbs / PEAK_GBS, 1), "us_per_call": round(tot / cnt * 1e3, 2)}
        S.set_output_reuse(True)
        del pool
        torch.cuda.empty_cache()
    return out


def ttdt_bench(torch, dev, reps=5) -> dict:
    """The contraction around the sort (ccsd_t2_8, ccsd_t2_8_skeleton.F:162-183)
    on the device: two sorts, one DGEMM (rocBLAS), one sort-accumulate, at
    tilesizes 20, 30 and 60 (all six dims equal).  Reports the time per
    contraction, its FP64 rate (2*M*N*K flops) and the share of the three
    sorts, timed alone on the same shapes."""
    import ctypes
    import spaghetty_amd as S
    from spaghetty_amd import _lib
    L = _lib.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for ts in (20, 30, 60):
        d = (ts,) * 6
        n4 = ts ** 4
        t = torch.rand(n4, dtype=torch.float64, device=dev)
        v = torch.rand(n4, dtype=torch.float64, device=dev)
        r = torch.zeros(n4, dtype=torch.float64, device=dev)
        w = torch.empty(3 * n4, dtype=torch.float64, device=dev)
        S.ccsd_t2_8(t, v, r, *d)
        torch.cuda.synchronize()

        def timed(fn):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps / 1e3

        t_all = timed(lambda: S.ccsd_t2_8(t, v, r, *d))
        d4 = _lib.ints((ts,) * 4)
        p4321, p2143 = _lib.ints((4, 3, 2, 1)), _lib.ints((2, 1, 4, 3))

        def sorts():
            L.spg_sort(t.data_ptr(), w.data_ptr(), 4, d4, p4321, 1.0, 0.0, st)
            L.spg_sort(v.data_ptr(), w.data_ptr() + 8 * n4, 4, d4, p2143, 1.0, 0.0, st)
            L.spg_sort(w.data_ptr() + 16 * n4, r.data_ptr(), 4, d4, p2143, 1.0, 0.0, st)
        t_sorts = timed(sorts)
        flops = 2.0 * ts ** 6
        out[f"{ts}^6"] = {"us_per_contraction": round(t_all * 1e6, 1),
                          "tflops": round(flops / t_all / 1e12, 2),
                          "sorts_us": round(t_sorts * 1e6, 1),
                          "sorts_share": round(t_sorts / t_all, 3),
                          "m_n_k": ts * ts}
        del t, v, r, w
        torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part "
                                   "of it (tiles shard across GPUs with no exchange)"}))
        sys.exit(2)
    import torch
    import spaghetty_amd as S
    from spaghetty_amd import _lib
    _lib.lib()  # fail loudly if the HIP library is missing
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("allocating pools")
    pools = {dims: Pool(torch, dims, dev) for dims in SIZES}
    calls = sweep_calls()
    if args.calls_log:
        os.makedirs(os.path.dirname(os.path.abspath(args.calls_log)), exist_ok=True)
        with open(args.calls_log, "w") as f:
            json.dump({"warmup": args.warmup, "steps": args.steps,
                       "calls": [[list(d), list(p), f_, b] for d, p, f_, b in calls]}, f)
    step_bytes = sum(8 * int(np.prod(d)) * (3 if b else 2) for d, _, _, b in calls)
    caller = Caller(torch, calls, pools)
    dom_idx = [i for i, (d, _, _, b) in enumerate(calls) if d == (100,) * 4 and b == 0.0]

    def new_events(idx):
        return {i: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                for i in idx}

    log("warmup")
    for _ in range(args.warmup):
        caller.run()
    torch.cuda.synchronize()
    # the timed steps; HIP events around each launch of the dominant kernel
    # (the 100^4 sorts), recorded on the stream the kernels run on
    events_all = [new_events(dom_idx) for _ in range(args.steps)]
    log("timed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        caller.run(events_all[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps
    value = step_bytes / dt / 1e9
    # per-size / mode breakdown: one more (untimed) step with events on every call
    ev_b = new_events(range(len(calls)))
    caller.run(ev_b)
    torch.cuda.synchronize()
    per = {}
    for i, (dims, perm, f, beta) in enumerate(calls):
        ms = ev_b[i][0].elapsed_time(ev_b[i][1])
        key = ("x".join(map(str, dims)), "acc" if beta else "sort")
        s = per.setdefault(key, [0.0, 0, 0])
        s[0] += ms
        s[1] += 8 * int(np.prod(dims)) * (3 if beta else 2)
        s[2] += 1
    dom = [0.0, 0, 0]
    for ev in events_all:
        for i in dom_idx:
            dom[0] += ev[i][0].elapsed_time(ev[i][1])
            dom[1] += 16 * 100 ** 4
            dom[2] += 1
    per_size = {f"{k[0]}:{k[1]}": {"gbs": round(v[1] / (v[0] / 1e3) / 1e9, 1),
                                   "us_per_call": round(v[0] / v[2] * 1e3, 2)}
                for k, v in per.items()}
    dom_us = dom[0] / dom[2] * 1e3
    dom_bytes = 16 * 100 ** 4
    achieved = dom_bytes / (dom_us / 1e6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": "k_vtile/k_tiled/k_rows (tce_sort_4 100^4 sorts, 24 perms x 3 factors)",
                "launches": dom[2], "avg_us": round(dom_us, 2), "bytes_per_launch": dom_bytes}
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "tce_sort_4/tce_sortacc_4 sweep: 24 perms x {20^4, 30^4, 100^4, "
                               "10x30x20x40} x factor {1,-1,0.5} x {sort, accumulate}, FP64, "
                               "uniform(-1,1) inputs resident in HBM",
                   "calls_per_step": len(calls), "bytes_per_step": step_bytes,
                   "parallelism": "single GPU"},
        "pct_of_peak": round(100.0 * value / PEAK_GBS, 2),
        "per_size": per_size,
        "roofline": roofline,
    }
    del pools
    torch.cuda.empty_cache()
    log("config 1")
    result["config1"] = config1_bench(torch, dev, not args.no_cpu)
    log("6d")
    result["sweep_6d"] = sweep_6d(torch, dev)
    if not args.no_batched:
        log("batched")
        result["batched"] = batched_bench(S, torch, dev)
    if not args.no_serving:
        log("serving")
        result["serving"] = serving_bench(torch, dev)
    log("ttdt")
    result["ttdt"] = ttdt_bench(torch, dev)
    log("host drop-in")
    result["host_dropin_pcie_inclusive"] = host_dropin()
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
