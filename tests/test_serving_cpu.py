"""The serving engine's scheduling on the CPU, with the fake executor (a
bandwidth model in place of the GPU; no device calls): admission under the
HBM budget, FIFO order, serialisation of requests on one output, slot reuse
and overlap, budget rejection, latency statistics."""
from __future__ import annotations

import numpy as np
import pytest

from spaghetty_amd.serving import Server, poisson_arrivals
from spaghetty_amd import _lib

MB = 1 << 20


def fake(budget, slots=2, max_batch=64, gbs=100.0, lat=50.0):
    return Server(budget, slots=slots, max_batch=max_batch, fake=True, fake_gbs=gbs,
                  fake_latency_us=lat)


def test_budget_bounds_in_flight_bytes():
    # 40 requests of 2 MiB footprint (1 MiB in + out) under a 10 MiB budget
    s = fake(10 * MB, slots=4, max_batch=8)
    dims = (16, 16, 16, 32)  # 131072 doubles = 1 MiB
    base = 1 << 30
    for i in range(40):
        s.submit(base, base + (1 + i) * 2 * MB, dims, (4, 3, 2, 1), 1.0, 0.0, arrival_us=0.0)
    s.start()
    s.wait(60)
    st = s.stats()
    assert st["completed"] == 40
    assert st["peak_in_flight"] <= 10 * MB
    assert st["batches"] >= 8  # at most 5 requests fit at once
    r = s.results()
    assert np.all(r["done_us"] >= r["admit_us"]) and np.all(r["admit_us"] >= r["arrival_us"])
    s.close()


def test_fifo_admission_order():
    s = fake(6 * MB, slots=1, max_batch=2)
    base = 1 << 30
    for i in range(12):
        # increasing arrival times, all of them in the past by the time the
        # engine's first pass runs
        s.submit(base, base + (1 + i) * 2 * MB, (16, 16, 16, 32), (4, 3, 2, 1), 1.0, 0.0,
                 arrival_us=float(i) * 1e-6)
    s.start()
    s.wait(60)
    r = s.results()
    # admitted (and finished) in arrival order, two per group
    assert np.all(np.diff(r["admit_us"]) >= 0)
    assert list(r["batch"]) == [i // 2 for i in range(12)]
    s.close()


def test_same_output_serialised():
    # accumulations into one tensor never share a group and run in order
    s = fake(1 << 34, slots=4, max_batch=64)
    base = 1 << 30
    out = base + 64 * MB
    ids = []
    for i in range(6):
        ids.append(s.submit(base + i * MB, out, (8, 8, 8, 8), (2, 1, 4, 3), 1.0, 1.0, arrival_us=0.0))
    other = s.submit(base, base + 128 * MB, (8, 8, 8, 8), (2, 1, 4, 3), 1.0, 0.0, arrival_us=0.0)
    s.start()
    s.wait(60)
    r = s.results()
    b = r["batch"][ids]
    assert len(set(b)) == 6 and np.all(np.diff(b) > 0)
    assert np.all(np.diff(r["admit_us"][ids]) > 0)
    for i in range(1, 6):  # each starts after the previous one finished
        assert r["admit_us"][ids[i]] >= r["done_us"][ids[i - 1]]
    assert r["batch"][other] == r["batch"][ids[0]]  # unrelated work joins the first group
    s.close()


def test_groups_overlap_across_slots():
    s = fake(1 << 34, slots=4, max_batch=1, gbs=10.0, lat=100.0)
    base = 1 << 30
    for i in range(8):
        s.submit(base, base + (1 + i) * 4 * MB, (16, 16, 16, 32), (4, 3, 2, 1), 1.0, 0.0,
                 arrival_us=0.0)
    s.start()
    s.wait(60)
    r = s.results()
    # with four slots, the second request starts before the first finishes
    assert r["admit_us"][1] < r["done_us"][0]
    s.close()


def test_continuous_admission_of_late_arrivals():
    s = fake(1 << 34, slots=2, max_batch=256, gbs=50.0, lat=20.0)
    base = 1 << 30
    # arrivals spread over ~40 ms, so the engine admits groups while later
    # requests are still arriving even when its thread starts late
    t = poisson_arrivals(200, rate_per_us=0.005, seed=3)
    for i, ti in enumerate(t):
        s.submit(base, base + (1 + i) * MB, (8, 8, 8, 16), (3, 1, 4, 2), 0.5, 1.0, arrival_us=ti)
    s.start()
    s.wait(60)
    st = s.stats()
    r = s.results()
    assert st["completed"] == 200 and st["batches"] > 1
    assert np.all(r["admit_us"] >= r["arrival_us"] - 1e-6)
    assert st["p99_us"] >= st["p50_us"] > 0


def test_rejections():
    s = fake(1 * MB)
    with pytest.raises(_lib.SpgError) as e:
        s.submit(1 << 30, (1 << 30) + 64 * MB, (64, 64, 64, 1), (1, 2, 3, 4))
    assert e.value.code == _lib.SPG_EBUDGET
    with pytest.raises(ValueError):
        s.submit(1 << 30, (1 << 30) + 64 * MB, (4, 4), (1, 1))
    with pytest.raises(ValueError):
        s.submit(1 << 30, 1 << 30, (4, 4), (2, 1))  # in == out
    s.close()


def test_submissions_from_other_threads_while_running():
    # requests submitted from several host threads while the engine admits
    # and retires groups (the request list grows under the engine's feet)
    import threading
    s = fake(1 << 34, slots=3, max_batch=8, gbs=200.0, lat=5.0)
    base = 1 << 30
    s.start()
    per, nthreads = 150, 4
    errors = []

    def producer(k):
        try:
            for i in range(per):
                j = k * per + i
                s.submit(base, base + (1 + j) * MB, (8, 8, 8, 16), (3, 1, 4, 2), 1.0, 0.0)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    th = [threading.Thread(target=producer, args=(k,)) for k in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors
    s.wait(60)
    st = s.stats()
    r = s.results()
    assert st["completed"] == per * nthreads and st["error"] == 0
    assert np.all(r["done_us"] >= r["admit_us"]) and np.all(r["admit_us"] >= r["arrival_us"] - 1e-6)
    s.close()


def test_idle_engine_takes_no_core():
    # a started server with nothing to do blocks instead of polling
    import time
    s = fake(1 << 30, slots=2)
    s.start()
    time.sleep(0.05)
    c0, w0 = time.process_time(), time.perf_counter()
    time.sleep(0.4)
    used = time.process_time() - c0
    wall = time.perf_counter() - w0
    assert used < 0.25 * wall, (used, wall)
    # and it still wakes up for a submission
    s.submit(1 << 30, (1 << 30) + 64 * MB, (16, 16, 16, 32), (4, 3, 2, 1), 1.0, 0.0)
    s.wait(10)
    assert s.stats()["completed"] == 1
    s.close()
