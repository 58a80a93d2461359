"""spaghetty_amd -- MI355X (gfx950) tensor-sort library with the call surface of
spaghetty's generated ``tce_sort_4`` library (trunk/makelib.py:64-95).

    from spaghetty_amd import tce_sort_4
    tce_sort_4(unsorted, sorted, a, b, c, d, i, j, k, l, factor)

Compute runs only in the HIP library ``libspaghetty_amd.so`` (built with
``python -m spaghetty_amd.build``); importing works without it, calling does
not.
"""
from .api import (  # noqa: F401
    TileParams,
    ccsd_t2_8,
    dispatch_key,
    plan,
    plan_check,
    set_output_reuse,
    sort,
    tce_sort_2,
    tce_sort_4,
    tce_sort_6,
    tce_sort_8,
    tce_sortacc_2,
    tce_sortacc_4,
    tce_sortacc_6,
    tce_sortacc_8,
    trans,
    ttdt,
    version_of,
)
from .batched import Batch  # noqa: F401

__version__ = "0.1.0"
