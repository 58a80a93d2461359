"""Batched form: many independent sorts in one persistent (graph-capturable)
launch -- the device counterpart of calling tce_sort_4/6 once per
coupled-cluster tile (trunk/makelib.py:64-95) in a loop.

    b = Batch()
    for (u, s, dims, perm, f, acc) in tiles:       # device tensors
        b.add(u, s, dims, perm, f, 1.0 if acc else 0.0)
    b.commit()
    b.launch()            # or b.graph_launch()  (captured once, replayed)
"""
from __future__ import annotations

import ctypes
import json

from ._lib import check, ints, lib


class Batch:
    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().spg_batch_create(ctypes.byref(self._h)), "batch_create")
        self._keep = []  # tensors referenced by the work list stay alive
        self._device = None

    def add(self, unsorted, sorted, dims, perm, factor=1.0, beta=0.0) -> None:
        import torch
        for t, name in ((unsorted, "unsorted"), (sorted, "sorted")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous()):
                raise TypeError(f"{name} must be a contiguous float64 device tensor")
        n = 1
        for d in dims:
            n *= int(d)
        if unsorted.numel() < n or sorted.numel() < n:
            raise ValueError("array smaller than the dims")
        if self._device is None:
            self._device = unsorted.device
        if unsorted.device != self._device or sorted.device != self._device:
            raise ValueError("all tensors of a batch must be on one device")
        check(lib().spg_batch_add(self._h, unsorted.data_ptr(), sorted.data_ptr(), len(dims),
                                  ints(dims), ints(perm), float(factor), float(beta)),
              "batch_add")
        self._keep.append((unsorted, sorted))

    def next_wave(self) -> None:
        """Tasks added from now on form the next wave: waves run one after
        the other on the stream, so they may reuse each other's buffers (a
        job larger than device memory runs wave after wave through the same
        arrays).  Overlap checks apply within a wave."""
        check(lib().spg_batch_next_wave(self._h), "batch_next_wave")

    def set_flags(self, flags: int) -> None:
        """Kernel flags before commit: SPG_TILE_NT = 1 (nontemporal access),
        SPG_TILE_PIPE = 2 (next tile's loads overlap this tile's stores),
        SPG_TILE_LDS1 = 32 (one LDS buffer per workgroup: more workgroups
        per CU, a second barrier per tile), SPG_BATCH_WIDE = 64 (tasks that
        move element pairs take tiles of up to 4096 elements instead of
        2048), SPG_BATCH_INLINE = 256 (tasks of 2^22 elements and more run in
        the persistent launch too, instead of on their own per-call plans).
        The default is 0."""
        check(lib().spg_batch_set_flags(self._h, int(flags)), "batch_set_flags")

    def _stream(self, stream):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self._device)
        return ctypes.c_void_p(stream.cuda_stream)

    def commit(self, stream=None) -> None:
        import torch
        with torch.cuda.device(self._device):
            check(lib().spg_batch_commit(self._h, self._stream(stream)), "batch_commit")

    def launch(self, stream=None) -> None:
        import torch
        with torch.cuda.device(self._device):
            check(lib().spg_batch_launch(self._h, self._stream(stream)), "batch_launch")

    def graph_launch(self, stream=None) -> None:
        """Replays the batch as a captured hipGraph (captured on first use).
        Needs a non-default stream."""
        import torch
        with torch.cuda.device(self._device):
            check(lib().spg_batch_graph_launch(self._h, self._stream(stream)), "batch_graph_launch")

    @property
    def bytes(self) -> int:
        """Algorithmic HBM bytes of one launch: 8*N*(2 + acc) summed over tasks."""
        return int(lib().spg_batch_bytes(self._h))

    def info(self) -> dict:
        buf = ctypes.create_string_buffer(512)
        check(lib().spg_batch_info(self._h, buf, 512), "batch_info")
        return json.loads(buf.value.decode())

    def close(self) -> None:
        if self._h:
            lib().spg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_noalloc(tasks, flags: int = 0, stream=None) -> None:
    """The batched form in one call that allocates no device memory
    (spg_batch_run_noalloc): `tasks` is a sequence of (unsorted, sorted, dims,
    perm, factor, beta) with contiguous float64 device tensors on the current
    device.  The work list is staged in the library's fixed per-GPU scratch
    array (SPG_NOALLOC_SCRATCH_BYTES); calls that overlap on different
    streams of one GPU race on it."""
    import torch
    from ._lib import SPG_MAX_RANK
    n = len(tasks)
    ins = (ctypes.c_void_p * max(n, 1))()
    outs = (ctypes.c_void_p * max(n, 1))()
    ranks = (ctypes.c_int * max(n, 1))()
    dims = (ctypes.c_int * max(n * SPG_MAX_RANK, 1))()
    perms = (ctypes.c_int * max(n * SPG_MAX_RANK, 1))()
    fs = (ctypes.c_double * max(n, 1))()
    gs = (ctypes.c_double * max(n, 1))()
    device = None
    for i, (u, s, d, p, f, g) in enumerate(tasks):
        for t, name in ((u, "unsorted"), (s, "sorted")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous()):
                raise TypeError(f"{name} must be a contiguous float64 device tensor")
        if len(d) != len(p) or len(d) > SPG_MAX_RANK:
            raise ValueError("dims and perm must have the same length, at most 8")
        m = 1
        for v in d:
            m *= int(v)
        if u.numel() < m or s.numel() < m:
            raise ValueError("array smaller than the dims")
        device = device or u.device
        if u.device != device or s.device != device:
            raise ValueError("all tensors must be on one device")
        ins[i], outs[i], ranks[i] = u.data_ptr(), s.data_ptr(), len(d)
        for q in range(len(d)):
            dims[i * SPG_MAX_RANK + q] = int(d[q])
            perms[i * SPG_MAX_RANK + q] = int(p[q])
        fs[i], gs[i] = float(f), float(g)
    if device is None:
        return
    with torch.cuda.device(device):
        st = stream if stream is not None else torch.cuda.current_stream(device)
        check(lib().spg_batch_run_noalloc(n, ins, outs, ranks, dims, perms, fs, gs, int(flags),
                                          ctypes.c_void_p(st.cuda_stream)), "batch_run_noalloc")
