"""Python surface of the library, mirroring spaghetty's generated tce_sort_4.

The reference's packaged library is the Fortran routine written by
trunk/makelib.py:64-95:

    subroutine tce_sort_4(u, s, a, b, c, d, i, j, k, l, factor)

(``u`` unsorted, ``s`` sorted, ``a..d`` dims slowest -> fastest, ``i..l`` the
1-based permutation, ``factor`` the scale).  The same names, argument order
and meaning are kept here, plus NWChem's siblings ``tce_sortacc_4``,
``tce_sort_6``, ``tce_sortacc_6`` (semantics: trunk/old_sort.f:33-95,
97-234, 512-574, 576-713).

Arrays may be

* NumPy ``float64`` arrays (host): the call stages through the device and
  returns when ``sorted`` holds the result (like the Fortran routine);
* ``torch.float64`` CUDA/HIP tensors (device): the sort is enqueued on the
  current torch stream and the call returns immediately.

Error behaviour follows the reference: the ``tce_*`` routines print
``something is wrong...`` and leave ``sorted`` untouched when the
permutation is invalid (makelib.py:90-91).  :func:`sort` is the
status-checked generic form and raises ``ValueError`` instead.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import TileParams, check, ints, lib

__all__ = [
    "tce_sort_4", "tce_sortacc_4", "tce_sort_6", "tce_sortacc_6", "sort",
    "trans", "version_of", "plan", "dispatch_key", "plan_check", "TileParams",
    "set_output_reuse",
]


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _nelem(dims) -> int:
    n = 1
    for d in dims:
        n *= int(d)
    return n


def _check_array(x, n: int, name: str):
    if _is_torch(x):
        import torch
        if x.dtype != torch.float64:
            raise TypeError(f"{name} must be float64")
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if x.numel() < n:
            raise ValueError(f"{name} holds {x.numel()} elements, needs {n}")
        return "device" if x.is_cuda else "torch-cpu"
    if not isinstance(x, np.ndarray):
        raise TypeError(f"{name} must be a NumPy array or a torch tensor")
    if x.dtype != np.float64:
        raise TypeError(f"{name} must be float64")
    if not x.flags.c_contiguous:
        raise ValueError(f"{name} must be contiguous")
    if x.size < n:
        raise ValueError(f"{name} holds {x.size} elements, needs {n}")
    return "host"


def _addr(x) -> int:
    return x.data_ptr() if _is_torch(x) else x.ctypes.data


def _stream_of(x):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)


def _run(unsorted, sorted, dims, perm, factor, beta, params=None, where="sort") -> int:
    n = _nelem(dims)
    ku = _check_array(unsorted, n, "unsorted")
    ks = _check_array(sorted, n, "sorted")
    if ku != ks:
        raise TypeError("unsorted and sorted must both be host arrays or both device tensors")
    if ku == "torch-cpu":
        raise TypeError("CPU torch tensors are not accepted; pass NumPy arrays or device tensors")
    L = lib()
    d, p = ints(dims), ints(perm)
    if ku == "host":
        if params is not None:
            raise ValueError("explicit tile parameters need device tensors")
        return L.spg_sort_host(_addr(unsorted), _addr(sorted), len(dims), d, p,
                               float(factor), float(beta))
    if unsorted.device != sorted.device:
        raise ValueError("unsorted and sorted are on different devices")
    import torch
    with torch.cuda.device(unsorted.device):
        stream = _stream_of(unsorted)
        if params is None:
            return L.spg_sort(_addr(unsorted), _addr(sorted), len(dims), d, p,
                              float(factor), float(beta), stream)
        return L.spg_sort_with(_addr(unsorted), _addr(sorted), len(dims), d, p,
                               float(factor), float(beta), ctypes.byref(params), stream)


def _dropin(unsorted, sorted, dims, perm, factor, beta, name):
    rc = _run(unsorted, sorted, dims, perm, factor, beta, where=name)
    if rc == _lib.SPG_EINVAL_PERM:
        print(" something is wrong...")  # trunk/makelib.py:90-91
        return
    check(rc, name)


def tce_sort_4(unsorted, sorted, a, b, c, d, i, j, k, l, factor):
    """sorted(perm) = factor * unsorted (trunk/makelib.py:71; old_sort.f:33-95)."""
    _dropin(unsorted, sorted, (a, b, c, d), (i, j, k, l), factor, 0.0, "tce_sort_4")


def tce_sortacc_4(unsorted, sorted, a, b, c, d, i, j, k, l, factor):
    """sorted(perm) += factor * unsorted (old_sort.f:512-574)."""
    _dropin(unsorted, sorted, (a, b, c, d), (i, j, k, l), factor, 1.0, "tce_sortacc_4")


def tce_sort_6(unsorted, sorted, a, b, c, d, e, f, i, j, k, l, m, n, factor):
    """6D sort (old_sort.f:97-234)."""
    _dropin(unsorted, sorted, (a, b, c, d, e, f), (i, j, k, l, m, n), factor, 0.0, "tce_sort_6")


def tce_sortacc_6(unsorted, sorted, a, b, c, d, e, f, i, j, k, l, m, n, factor):
    """6D sort-accumulate (old_sort.f:576-713)."""
    _dropin(unsorted, sorted, (a, b, c, d, e, f), (i, j, k, l, m, n), factor, 1.0,
            "tce_sortacc_6")


def tce_sort_2(unsorted, sorted, a, b, i, j, factor):
    """2D sort (old_sort.f:10-31)."""
    _dropin(unsorted, sorted, (a, b), (i, j), factor, 0.0, "tce_sort_2")


def tce_sortacc_2(unsorted, sorted, a, b, i, j, factor):
    """2D sort-accumulate (old_sort.f:489-510)."""
    _dropin(unsorted, sorted, (a, b), (i, j), factor, 1.0, "tce_sortacc_2")


def tce_sort_8(unsorted, sorted, a, b, c, d, e, f, g, h, i, j, k, l, m, n, o, p, factor):
    """8D sort (old_sort.f:236-478)."""
    _dropin(unsorted, sorted, (a, b, c, d, e, f, g, h), (i, j, k, l, m, n, o, p), factor, 0.0,
            "tce_sort_8")


def tce_sortacc_8(unsorted, sorted, a, b, c, d, e, f, g, h, i, j, k, l, m, n, o, p, factor):
    """8D sort-accumulate (old_sort.f:715-957)."""
    _dropin(unsorted, sorted, (a, b, c, d, e, f, g, h), (i, j, k, l, m, n, o, p), factor, 1.0,
            "tce_sortacc_8")


def ttdt(a, adims, aperm, b, bdims, bperm, c, cdims, cperm, m: int, n: int, k: int,
         alpha: float = 1.0, beta: float = 0.0) -> None:
    """The contraction around the sort (spg_ttdt), on device tensors:
    A sorted by aperm into K x M (column-major), B by bperm into K x N,
    C' = alpha * A'^T B' (DGEMM 'T','N'), C = beta*C + C' sorted by cperm
    (branches/old/python/macaroni/ccsd_t2_8_skeleton.F:162-183)."""
    for x, name, dims in ((a, "a", adims), (b, "b", bdims), (c, "c", cdims)):
        if _check_array(x, _nelem(dims), name) != "device":
            raise TypeError(f"{name} must be a float64 device tensor")
    if not (a.device == b.device == c.device):
        raise ValueError("a, b and c are on different devices")
    import torch
    with torch.cuda.device(c.device):
        rc = lib().spg_ttdt(_addr(a), len(adims), ints(adims), ints(aperm), _addr(b), len(bdims),
                            ints(bdims), ints(bperm), _addr(c), len(cdims), ints(cdims), ints(cperm),
                            int(m), int(n), int(k), float(alpha), float(beta), None, _stream_of(c))
    check(rc, "ttdt")


def ccsd_t2_8(t, v, r, dim1, dim2, dim3, dim4, dim5, dim6, accumulate: bool = False) -> None:
    """i0(p3,p4,h1,h2) = Sum(p5,p6) v(p3,p4,p5,p6) * t(p5,p6,h1,h2) as the
    reference's ccsd_t2_8 computes it (ccsd_t2_8_skeleton.F:162-183):
    t (dims dim5,dim6,dim1,dim2) sorted 4321, v (dim3,dim4,dim5,dim6) sorted
    2143, dgemm('T','N', dim1*dim2, dim3*dim4, dim5*dim6), result (dims
    dim4,dim3,dim2,dim1) sorted 2143 into r (dims dim3,dim4,dim1,dim2),
    overwriting r as the skeleton's final tce_sort_4 does (:178); with
    accumulate=True the last sort adds into r, as tce_sortacc_4 would."""
    ttdt(t, (dim5, dim6, dim1, dim2), (4, 3, 2, 1), v, (dim3, dim4, dim5, dim6), (2, 1, 4, 3),
         r, (dim4, dim3, dim2, dim1), (2, 1, 4, 3), dim1 * dim2, dim3 * dim4, dim5 * dim6,
         1.0, 1.0 if accumulate else 0.0)


def sort(unsorted, sorted, dims: Sequence[int], perm: Sequence[int], factor: float = 1.0,
         beta: float = 0.0, params: TileParams | None = None) -> None:
    """Generic rank-0..8 form: sorted = beta*sorted + factor*perm(unsorted).

    beta = 0 sorts, beta = 1 accumulates (the generated five-case routines of
    trunk/generator2.py:133-443 / generator6.py:137-220 take (f, g) = (factor,
    beta)).  Raises ValueError for an invalid permutation or dims."""
    check(_run(unsorted, sorted, dims, perm, factor, beta, params), "sort")


def trans(unsorted, sorted, dims, perm, factor=1.0, acc_factor=0.0):
    """trans_<P>_loop_<L>_<omp>(u, s, d1..dN, f, g) of the generated libraries
    (generator6.py:137-220): loop order is a CPU notion and has no GPU
    counterpart, so it is not an argument."""
    sort(unsorted, sorted, dims, perm, factor, acc_factor)


def version_of(perm: Sequence[int]) -> int:
    """The reference's dispatch key: version = 1000*i + 100*j + 10*k + l
    (trunk/makelib.py:79; tce_sort4_switch.F:9), generalised to any rank."""
    v = 0
    for p in perm:
        v = 10 * v + int(p)
    return v


def plan(dims, perm, beta: float = 0.0, params: TileParams | None = None) -> dict:
    """The execution plan the library would use (host only, no GPU needed)."""
    import json
    buf = ctypes.create_string_buffer(8192)
    rc = lib().spg_plan_describe(len(dims), ints(dims), ints(perm), float(beta),
                                 ctypes.byref(params) if params is not None else None, buf, 8192)
    check(rc, "plan")
    return json.loads(buf.value.decode())


def set_output_reuse(reused: bool) -> bool:
    """Whether later calls (this whole process) plan for outputs that are
    written again soon (True, the default: the dispatch table's plans) or for
    outputs written cold, as when streaming through many distinct arrays
    (False: the plans of dispatch_gfx950_streamed.jsonl first, where it has
    the problem).  Returns the previous setting."""
    return bool(lib().spg_set_output_reuse(1 if reused else 0))


def dispatch_key(dims, perm, beta: float = 0.0) -> str:
    buf = ctypes.create_string_buffer(512)
    check(lib().spg_dispatch_key(len(dims), ints(dims), ints(perm), float(beta), buf, 512),
          "dispatch_key")
    return buf.value.decode()


def plan_check(unsorted: np.ndarray, sorted: np.ndarray, dims, perm, factor=1.0, beta=0.0,
               params: TileParams | None = None) -> None:
    """Walks the plan's tile/slot index math on the host (no GPU) and checks
    that it produces every output element exactly once.  A test hook for the
    plan logic, not a compute path: the values it writes come from the same
    walk the kernels perform."""
    n = _nelem(dims)
    if _check_array(unsorted, n, "unsorted") != "host" or _check_array(sorted, n, "sorted") != "host":
        raise TypeError("plan_check takes host arrays")
    rc = lib().spg_plan_check(_addr(unsorted), _addr(sorted), len(dims), ints(dims), ints(perm),
                              float(factor), float(beta),
                              ctypes.byref(params) if params is not None else None)
    check(rc, "plan_check")
