"""Serving form: a continuous-batching request loop over the HIP library.

The engine (C++, spaghetty_amd/csrc/serve.hip) runs on its own thread:
requests arrive over time (each carries an arrival time), are admitted in
FIFO order while the input+output bytes of unfinished requests fit the HBM
budget, and every admitted group runs as one batched persistent kernel on a
free stream, overlapping with groups already on the device.  Requests whose
data overlap an earlier unfinished request's output wait for it, so two
accumulations into one tensor keep their arrival order.

    srv = Server(budget_bytes=64 << 30, slots=4)
    for t_us, (u, s, dims, perm, f, beta) in trace:
        srv.submit(u, s, dims, perm, f, beta, arrival_us=t_us)
    srv.start(); srv.wait()
    srv.stats()   # p50 / p99 latency (us), aggregate GB/s, batches ...
"""
from __future__ import annotations

import ctypes
import json

import numpy as np

from ._lib import check, ints, lib

SPG_SERVER_FAKE = 1


class Server:
    def __init__(self, budget_bytes: int, slots: int = 4, max_batch: int = 1024,
                 fake: bool = False, fake_gbs: float = 4000.0, fake_latency_us: float = 5.0):
        self._h = ctypes.c_void_p()
        self.fake = fake
        if not fake:
            import torch
            torch.cuda.init()
        check(lib().spg_server_create(int(budget_bytes), int(slots), int(max_batch),
                                      SPG_SERVER_FAKE if fake else 0, ctypes.byref(self._h)),
              "server_create")
        if fake:
            check(lib().spg_server_set_fake_rate(self._h, float(fake_gbs), float(fake_latency_us)),
                  "server_set_fake_rate")
        self._keep = []
        self.n = 0

    @staticmethod
    def _ptr(x):
        if x is None:
            return None
        if isinstance(x, int):
            return x
        return x.data_ptr()

    def submit(self, unsorted, sorted, dims, perm, factor=1.0, beta=0.0,
               arrival_us: float = -1.0) -> int:
        """Queues one request; `sorted` None lets the server allocate (and
        discard) the output.  Device tensors, or raw device addresses."""
        rid = ctypes.c_int64()
        check(lib().spg_server_submit(self._h, self._ptr(unsorted), self._ptr(sorted), len(dims),
                                      ints(dims), ints(perm), float(factor), float(beta),
                                      float(arrival_us), ctypes.byref(rid)), "server_submit")
        self._keep.append((unsorted, sorted))
        self.n += 1
        return int(rid.value)

    def start(self) -> None:
        check(lib().spg_server_start(self._h), "server_start")

    def wait(self, timeout_s: float = 600.0) -> None:
        check(lib().spg_server_wait(self._h, float(timeout_s)), "server_wait")

    def results(self) -> dict:
        n = self.n
        arr = np.zeros(n)
        adm = np.zeros(n)
        done = np.zeros(n)
        batch = np.zeros(n, dtype=np.int32)
        check(lib().spg_server_results(self._h, n, arr.ctypes.data, adm.ctypes.data,
                                       done.ctypes.data, batch.ctypes.data), "server_results")
        return {"arrival_us": arr, "admit_us": adm, "done_us": done, "batch": batch}

    def stats(self) -> dict:
        buf = ctypes.create_string_buffer(1024)
        check(lib().spg_server_stats(self._h, buf, 1024), "server_stats")
        return json.loads(buf.value.decode())

    def close(self) -> None:
        if self._h:
            lib().spg_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(n: int, rate_per_us: float, seed: int = 0) -> np.ndarray:
    """Arrival times (us) of a Poisson stream with the given mean rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_us, size=n))
