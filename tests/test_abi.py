"""The C ABI (CPU only): the library loads, exports every symbol that
include/spaghetty_amd.h declares, and handles invalid arguments the way the
reference does -- before touching a GPU."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import sys

import numpy as np

from spaghetty_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spaghetty_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and (n.startswith("spg_") or n.startswith("tce_"))})


def test_header_symbols_exported():
    L = _lib.lib()
    names = declared_functions()
    assert "tce_sort_4_" in names and "spg_sort" in names and "spg_batch_launch" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert not set(names) - set(_lib.SIGNATURES)


def test_exports_with_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in declared_functions():
        assert n in exported, n


def test_version_and_strerror():
    L = _lib.lib()
    assert L.spg_version() >= 10000
    assert L.spg_strerror(_lib.SPG_EINVAL_PERM) == b"invalid permutation"


def _by_ref(vals):
    return [ctypes.byref(ctypes.c_int(v)) for v in vals]


def test_dropin_bad_permutation_is_a_noop():
    # trunk/makelib.py:90-91: print 'something is wrong...' and return
    code = r"""
import ctypes, numpy as np, sys
sys.path.insert(0, %r)
from spaghetty_amd import _lib
L = _lib.lib()
u = np.arange(24, dtype=np.float64); s = np.full(24, 7.0)
ints = [ctypes.byref(ctypes.c_int(v)) for v in (2, 3, 2, 2, 1, 1, 3, 4)]
L.tce_sort_4_(u.ctypes.data, s.ctypes.data, *ints, ctypes.byref(ctypes.c_double(1.0)))
assert (s == 7.0).all()
assert L.spg_last_error() == _lib.SPG_EINVAL_PERM
print("done")
""" % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "something is wrong..." in r.stdout
    assert "done" in r.stdout


def test_python_dropin_bad_permutation(capsys):
    import spaghetty_amd as S
    u = np.arange(24, dtype=np.float64)
    s = np.full(24, 7.0)
    S.tce_sort_4(u, s, 2, 3, 2, 2, 4, 4, 1, 2, 1.0)
    assert "something is wrong..." in capsys.readouterr().out
    assert (s == 7.0).all()


def test_status_api_errors():
    import pytest
    import spaghetty_amd as S
    u = np.zeros(24)
    s = np.zeros(24)
    with pytest.raises(ValueError):
        S.sort(u, s, (2, 3, 4), (1, 1, 2))
    with pytest.raises(ValueError):
        S.sort(u, s, (2, 3, 5), (1, 2, 3))  # array too small
    with pytest.raises(TypeError):
        S.sort(u.astype(np.float32), s, (2, 3, 4), (1, 2, 3))
    # aliasing is refused before any device work
    with pytest.raises(ValueError):
        S.sort(u, u, (2, 3, 4), (3, 2, 1))
    # empty problems are a no-op
    S.sort(u, s, (0, 3, 4), (3, 2, 1))
