"""CPU oracle for spaghetty's tensor-sort path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker (or the timed CPU
baseline).  The ``spaghetty_amd`` package never imports it; its product path
runs the HIP library or raises.

Three independent statements of the same contract live here:

* ``liboracle.so`` (``oracle.c``): C restatement of NWChem's ``old_sort_4/6``
  and ``old_sortacc_4/6`` (trunk/old_sort.f:33-95, 97-234, 512-574, 576-713)
  and of the generated library kernel ``trans_<P>_loop_<L>_omp_any_f``
  (trunk/generator.py:107-159) used as the timed CPU baseline;
* ``_ref/libold_sort.so``: the reference's own ``trunk/old_sort.f``, compiled
  from the reference checkout where it exists (oracle/Makefile);
* :func:`closed_form`: the NumPy statement
  ``S = f * U.reshape(dims).transpose(perm - 1)`` (C order, dims slowest ->
  fastest), derived from the index rule at trunk/old_sort.f:51-52.

Inputs: :func:`uniform` (seeded, index-addressable splitmix64 -> uniform(-1,1),
bit-identical to ``oracle_fill_uniform`` in C) and :func:`kat` (the reference
testers' integer ramp, trunk/tester_futil.F:139-185).
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libold_sort.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
_ref = None


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def lib() -> ctypes.CDLL:
    """The C restatement (oracle.c); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        i, d = ctypes.c_int, ctypes.c_double
        L.oracle_old_sort_4.argtypes = [_dp, _dp] + [i] * 8 + [d]
        L.oracle_old_sortacc_4.argtypes = [_dp, _dp] + [i] * 8 + [d]
        L.oracle_old_sort_6.argtypes = [_dp, _dp] + [i] * 12 + [d]
        L.oracle_old_sortacc_6.argtypes = [_dp, _dp] + [i] * 12 + [d]
        L.oracle_sort_n.argtypes = [_dp, _dp, i, _ip, _ip, d, d, i]
        L.oracle_sort_n.restype = i
        L.oracle_trans.argtypes = [_dp, _dp, i, _ip, _ip, _ip, d, d, i, i]
        L.oracle_trans.restype = i
        L.oracle_init_kat.argtypes = [_dp, i, _ip]
        L.oracle_fill_uniform.argtypes = [_dp, ctypes.c_int64, ctypes.c_uint64]
        L.oracle_max_threads.restype = i
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    """The reference's own old_sort.f (gfortran-style ABI: trailing underscore,
    every argument by reference)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_PATH} missing (reference checkout absent)")
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def _ints(v: Sequence[int]):
    arr = (ctypes.c_int * len(v))(*[int(t) for t in v])
    return arr


# --------------------------------------------------------------------------
# the four NWChem routines, three ways
# --------------------------------------------------------------------------

def old_sort(x: np.ndarray, y: np.ndarray, dims, perm, factor: float, acc: bool):
    """C restatement of old_sort_{4,6} / old_sortacc_{4,6}; y updated in place."""
    L = lib()
    R = len(dims)
    if R == 4:
        fn = L.oracle_old_sortacc_4 if acc else L.oracle_old_sort_4
    elif R == 6:
        fn = L.oracle_old_sortacc_6 if acc else L.oracle_old_sort_6
    else:
        rc = L.oracle_sort_n(_ptr(x), _ptr(y), R, _ints(dims), _ints(perm),
                             float(factor), 1.0 if acc else 0.0, 1 if acc else 0)
        if rc != 0:
            raise ValueError("bad permutation or dims")
        return y
    fn(_ptr(x), _ptr(y), *[int(t) for t in dims], *[int(t) for t in perm], float(factor))
    return y


def ref_old_sort(x: np.ndarray, y: np.ndarray, dims, perm, factor: float, acc: bool):
    """The reference's compiled old_sort_{2,4,6,8} / old_sortacc_{...}."""
    R = len(dims)
    name = ("old_sortacc_%d_" if acc else "old_sort_%d_") % R
    fn = getattr(ref(), name)
    ints = [ctypes.byref(ctypes.c_int(int(t))) for t in list(dims) + list(perm)]
    fn(_ptr(x), _ptr(y), *ints, ctypes.byref(ctypes.c_double(float(factor))))
    return y


def sort_n(x, y, dims, perm, f: float, g: float, mode: int):
    """Any rank; mode 0 sort, 1 sort-accumulate, 2 y = g*y + f*x."""
    rc = lib().oracle_sort_n(_ptr(x), _ptr(y), len(dims), _ints(dims), _ints(perm),
                             float(f), float(g), int(mode))
    if rc != 0:
        raise ValueError("bad permutation or dims")
    return y


def closed_form(x: np.ndarray, dims, perm, factor: float, y0=None):
    """NumPy statement: f * X.reshape(dims).transpose(perm-1), C order.
    With ``y0`` given, returns y0 + (f*x) permuted (two roundings)."""
    t = np.asarray(x).reshape(tuple(dims)).transpose([p - 1 for p in perm])
    s = np.ascontiguousarray(t).reshape(-1) * np.float64(factor)
    if y0 is not None:
        s = np.asarray(y0).reshape(-1) + s
    return s


def source_positions(dims, perm, positions) -> np.ndarray:
    """Input position ia of each output position ib (old_sort.f:51-52): the
    element that lands at ib.  Lets a test gather only the inputs it checks."""
    dims = [int(d) for d in dims]
    out_dims = [dims[p - 1] for p in perm]
    pos = np.asarray(positions, dtype=np.int64)
    o = np.unravel_index(pos, out_dims)            # output multi-index, slowest first
    coords = [None] * len(dims)
    for q, p in enumerate(perm):
        coords[p - 1] = o[q]                       # output axis q is input axis p-1
    return np.ravel_multi_index(coords, dims)


def sampled_values(x_at_src, factor: float, y0_at_pos=None, g: float = 1.0):
    """The closed form's arithmetic on gathered values: f*x (sort),
    y0 + f*x (accumulate, two roundings) or g*y0 + f*x (five-case form)."""
    s = np.asarray(x_at_src, dtype=np.float64) * np.float64(factor)
    if y0_at_pos is None:
        return s
    y = np.asarray(y0_at_pos, dtype=np.float64)
    return (y if g == 1.0 else y * np.float64(g)) + s


def closed_form_at(x: np.ndarray, dims, perm, factor: float, positions, y0=None):
    """closed_form at the given output positions only (for arrays too large
    to permute whole in a test): output element ib holds f * x[ia] (+ y0[ib])
    with ia the input position of the same multi-index (old_sort.f:51-52)."""
    pos = np.asarray(positions, dtype=np.int64)
    ia = source_positions(dims, perm, pos)
    xs = np.asarray(x).reshape(-1)[ia]
    return sampled_values(xs, factor, None if y0 is None else np.asarray(y0).reshape(-1)[pos])


def ttdt(a, adims, aperm, b, bdims, bperm, cdims, cperm, m, n, k, alpha=1.0, beta=0.0, c0=None):
    """The contraction around the sort, restated from
    branches/old/python/macaroni/ccsd_t2_8_skeleton.F:162-183 with this
    module's closed form for the sorts: A sorted into a K x M column-major
    matrix (C-order shape (M, K)), B into K x N ((N, K)), C' = alpha A'^T B'
    column-major M x N (C-order shape (N, M)), then C = beta*C0 + C' sorted.
    The sum runs in NumPy's order: bit-exact only where every partial sum is
    exact (integer-valued data)."""
    As = closed_form(a, adims, aperm, 1.0).reshape(m, k)
    Bs = closed_form(b, bdims, bperm, 1.0).reshape(n, k)
    Cp = (Bs @ As.T) * np.float64(alpha) if alpha != 1.0 else Bs @ As.T   # (N, M): M fastest
    s = closed_form(Cp.reshape(-1), cdims, cperm, 1.0)
    if beta != 0.0:
        y = np.asarray(c0).reshape(-1)
        s = (y if beta == 1.0 else y * np.float64(beta)) + s
    return s


# --------------------------------------------------------------------------
# generated library kernel (CPU baseline)
# --------------------------------------------------------------------------

# Best loop order per 4D permutation for the packaged library, as tuned by the
# reference on its "knight" machine (trunk/makelib.py:39-62).  Data kept alike
# on purpose: it is the reference's dispatch table for its CPU library.
KNIGHT_4D = {
    "4321": "3241", "4312": "3142", "4231": "2341", "4132": "1342",
    "4213": "2134", "4123": "1234", "3421": "3241", "3412": "3142",
    "2431": "2341", "1432": "1342", "2413": "2143", "1423": "1243",
    "3241": "3214", "3142": "3124", "2341": "2314", "1342": "1324",
    "2143": "2143", "1243": "1234", "3214": "2134", "3124": "1243",
    "2314": "2341", "1324": "1234", "2134": "2143", "1234": "1243",
}


def trans(u, s, dims, perm, loop, f: float, g: float = 0.0, mode: int = 0,
          nthreads: int = 0):
    """trans_<P>_loop_<L>_omp_any_f restated (OpenMP, collapse(2)/(3))."""
    rc = lib().oracle_trans(_ptr(u), _ptr(s), len(dims), _ints(dims), _ints(perm),
                            _ints(loop), float(f), float(g), int(mode), int(nthreads))
    if rc != 0:
        raise ValueError("bad arguments")
    return s


def max_threads() -> int:
    return int(lib().oracle_max_threads())


# --------------------------------------------------------------------------
# inputs
# --------------------------------------------------------------------------

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(n: int, seed: int, start: int = 0) -> np.ndarray:
    """uniform(-1,1) doubles, element i = splitmix64(seed + (i+1)*golden);
    identical to oracle_fill_uniform (and to the device fill used by bench)."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return 2.0 * u - 1.0


def uniform_c(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().oracle_fill_uniform(_ptr(out), int(n), ctypes.c_uint64(seed))
    return out


def kat(fortran_dims) -> np.ndarray:
    """init_4d_array / init_6d_array values (tester_futil.F:152,176)."""
    n = int(np.prod(fortran_dims))
    out = np.empty(n, dtype=np.float64)
    lib().oracle_init_kat(_ptr(out), len(fortran_dims), _ints(fortran_dims))
    return out
