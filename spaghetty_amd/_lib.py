"""ctypes binding of libspaghetty_amd.so (the C ABI in include/spaghetty_amd.h).

The library is the only compute path: if it is missing, every entry point
raises ``RuntimeError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspaghetty_amd.so")

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
_ip = ctypes.POINTER(c_int)
_dp = ctypes.POINTER(c_double)

SPG_OK = 0
SPG_EINVAL_PERM = 1
SPG_EINVAL_DIM = 2
SPG_EALIAS = 3
SPG_ENOMEM = 4
SPG_EHIP = 5
SPG_EINVAL = 6
SPG_EBUDGET = 7
SPG_ESTATE = 8
SPG_MAX_RANK = 8
SPG_BATCH_WIDE = 64
SPG_BATCH_INLINE = 256
SPG_NOALLOC_SCRATCH_BYTES = 16 << 20


class TileParams(ctypes.Structure):
    """spg_tile_params: the autotuner's search space (see the C header)."""

    _fields_ = [
        ("kind", c_int32),
        ("nx", c_int32),
        ("tx", c_int32),
        ("ny", c_int32),
        ("ty", c_int32),
        ("block", c_int32),
        ("grid_per_cu", c_int32),
        ("xcd_remap", c_int32),
        ("max_tile", c_int32),
        ("flags", c_int32),
        ("walk_x", c_int32),
        ("walk_mode", c_int32),
        ("walk_chunk", c_int32),
        ("tw", c_int32),
    ]

    @classmethod
    def default(cls, **kw) -> "TileParams":
        p = cls(-1, 0, 0, 0, 0, 0, 0, -1, 0, -1, 0, 0, 0, 0)
        for k, v in kw.items():
            setattr(p, k, int(v))
        return p

    @classmethod
    def from_record(cls, rec: dict) -> "TileParams":
        """The parameters of a dispatch-table record, read as the library
        reads them (a record without "tw" keeps one block per tile)."""
        names = [name for name, _ in cls._fields_]
        return cls.default(**{"tw": 1, **{k: rec[k] for k in names if k in rec}})

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_ARGS2 = [c_void_p, c_void_p] + [_ip] * 4 + [_dp]
_ARGS4 = [c_void_p, c_void_p] + [_ip] * 8 + [_dp]
_ARGS6 = [c_void_p, c_void_p] + [_ip] * 12 + [_dp]
_ARGS8 = [c_void_p, c_void_p] + [_ip] * 16 + [_dp]

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "tce_sort_4_": (None, _ARGS4),
    "tce_sort_4": (None, _ARGS4),
    "tce_sort_4__": (None, _ARGS4),
    "tce_sortacc_4_": (None, _ARGS4),
    "tce_sortacc_4": (None, _ARGS4),
    "tce_sortacc_4__": (None, _ARGS4),
    "tce_sort_6_": (None, _ARGS6),
    "tce_sort_6": (None, _ARGS6),
    "tce_sort_6__": (None, _ARGS6),
    "tce_sortacc_6_": (None, _ARGS6),
    "tce_sortacc_6": (None, _ARGS6),
    "tce_sortacc_6__": (None, _ARGS6),
    "tce_sort_2_": (None, _ARGS2),
    "tce_sort_2": (None, _ARGS2),
    "tce_sort_2__": (None, _ARGS2),
    "tce_sortacc_2_": (None, _ARGS2),
    "tce_sortacc_2": (None, _ARGS2),
    "tce_sortacc_2__": (None, _ARGS2),
    "tce_sort_8_": (None, _ARGS8),
    "tce_sort_8": (None, _ARGS8),
    "tce_sort_8__": (None, _ARGS8),
    "tce_sortacc_8_": (None, _ARGS8),
    "tce_sortacc_8": (None, _ARGS8),
    "tce_sortacc_8__": (None, _ARGS8),
    "spg_sort_host": (c_int, [c_void_p, c_void_p, c_int, _ip, _ip, c_double, c_double]),
    "spg_sort": (c_int, [c_void_p, c_void_p, c_int, _ip, _ip, c_double, c_double, c_void_p]),
    "spg_tile_params_default": (None, [ctypes.POINTER(TileParams)]),
    "spg_sort_with": (c_int, [c_void_p, c_void_p, c_int, _ip, _ip, c_double, c_double,
                              ctypes.POINTER(TileParams), c_void_p]),
    "spg_dispatch_load": (c_int, [c_char_p]),
    "spg_dispatch_set": (c_int, [c_char_p, ctypes.POINTER(TileParams)]),
    "spg_dispatch_clear": (c_int, []),
    "spg_dispatch_load_streamed": (c_int, [c_char_p]),
    "spg_set_output_reuse": (c_int, [c_int]),
    "spg_dispatch_key": (c_int, [c_int, _ip, _ip, c_double, c_char_p, c_int]),
    "spg_plan_describe": (c_int, [c_int, _ip, _ip, c_double, ctypes.POINTER(TileParams),
                                  c_char_p, c_int]),
    "spg_plan_check": (c_int, [c_void_p, c_void_p, c_int, _ip, _ip, c_double, c_double,
                               ctypes.POINTER(TileParams)]),
    "spg_selftest_fastdiv": (c_int, [c_int64]),
    "spg_batch_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "spg_batch_add": (c_int, [c_void_p, c_void_p, c_void_p, c_int, _ip, _ip, c_double, c_double]),
    "spg_batch_set_flags": (c_int, [c_void_p, c_int]),
    "spg_batch_next_wave": (c_int, [c_void_p]),
    "spg_batch_commit": (c_int, [c_void_p, c_void_p]),
    "spg_batch_launch": (c_int, [c_void_p, c_void_p]),
    "spg_batch_graph_launch": (c_int, [c_void_p, c_void_p]),
    "spg_batch_destroy": (c_int, [c_void_p]),
    "spg_batch_bytes": (c_int64, [c_void_p]),
    "spg_batch_info": (c_int, [c_void_p, c_char_p, c_int]),
    "spg_batch_run_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_int, c_void_p]),
    "spg_server_create": (c_int, [c_int64, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "spg_server_set_fake_rate": (c_int, [c_void_p, c_double, c_double]),
    "spg_server_submit": (c_int, [c_void_p, c_void_p, c_void_p, c_int, _ip, _ip, c_double,
                                  c_double, c_double, ctypes.POINTER(c_int64)]),
    "spg_server_start": (c_int, [c_void_p]),
    "spg_server_wait": (c_int, [c_void_p, c_double]),
    "spg_server_results": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "spg_server_stats": (c_int, [c_void_p, c_char_p, c_int]),
    "spg_server_destroy": (c_int, [c_void_p]),
    "spg_ttdt": (c_int, [c_void_p, c_int, _ip, _ip, c_void_p, c_int, _ip, _ip, c_void_p, c_int, _ip,
                         _ip, c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p]),
    "spg_last_error": (c_int, []),
    "spg_last_error_message": (c_char_p, []),
    "spg_strerror": (c_char_p, [c_int]),
    "spg_version": (c_int, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m spaghetty_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ints(v) -> ctypes.Array:
    v = [int(t) for t in v]
    return (c_int * max(len(v), 1))(*v)


def error_message(rc: int) -> str:
    L = lib()
    msg = L.spg_last_error_message()
    base = L.spg_strerror(rc).decode()
    msg = msg.decode() if msg else ""
    return f"{base}: {msg}" if msg and msg != base else base


class SpgError(RuntimeError):
    def __init__(self, rc: int, where: str):
        self.code = rc
        super().__init__(f"{where}: {error_message(rc)} (status {rc})")


def check(rc: int, where: str) -> None:
    if rc != SPG_OK:
        if rc in (SPG_EINVAL_PERM, SPG_EINVAL_DIM, SPG_EINVAL, SPG_EALIAS):
            raise SpgValueError(rc, where)
        raise SpgError(rc, where)


class SpgValueError(SpgError, ValueError):
    pass
