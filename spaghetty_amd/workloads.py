"""Synthetic coupled-cluster tile workloads (batched and serving forms).

NWChem TCE calls tce_sort_4 / tce_sort_6 once per tile of a blocked tensor
(papers/old/spaghetty_chapter_v1.tex:60; the (T) permutations in
background/nwchem-perturbative-triples-usage:5-58).  A tile's dims are tile
sizes of occupied (o) and virtual (v) index ranges: 4D tiles are o,o,v,v
(T2-like) and 6D tiles o,o,o,v,v,v (T3-like), in a random order.

The mix here: 70% 4D / 30% 6D; occupied dims uniform in [8,20], virtual dims
uniform in [16,40]; a random permutation per tile; factor from
{+1, -1, +0.5, -0.5}; accumulate on for half the tiles.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Tile:
    dims: tuple
    perm: tuple
    factor: float
    beta: float

    @property
    def nelem(self) -> int:
        n = 1
        for d in self.dims:
            n *= d
        return n

    @property
    def bytes(self) -> int:
        """Algorithmic HBM bytes: 8*N*(2 + acc)."""
        return 8 * self.nelem * (3 if self.beta else 2)


def cc_tile_mix(count: int, seed: int = 0, frac_4d: float = 0.7) -> list[Tile]:
    rng = np.random.default_rng(seed)
    tiles = []
    for _ in range(count):
        rank = 4 if rng.random() < frac_4d else 6
        k = rank // 2
        occ = [int(v) for v in rng.integers(8, 21, size=k)]
        vir = [int(v) for v in rng.integers(16, 41, size=k)]
        dims = occ + vir
        order = rng.permutation(rank)
        dims = tuple(dims[i] for i in order)
        perm = tuple(int(v) + 1 for v in rng.permutation(rank))
        factor = float(rng.choice([1.0, -1.0, 0.5, -0.5]))
        beta = 1.0 if rng.random() < 0.5 else 0.0
        tiles.append(Tile(dims, perm, factor, beta))
    return tiles


def waves_by_budget(tiles, budget_bytes: int):
    """Cuts a tile list, in order, into waves whose input + output footprint
    (16 bytes per element) fits `budget_bytes`: wave after wave, a job larger
    than device memory runs through one input and one output region of
    budget/2 bytes each.  Returns, per tile, (wave, element offset of its
    input and output inside the wave's regions)."""
    half = budget_bytes // 2
    out, wave, off = [], 0, 0
    for t in tiles:
        n = t.nelem
        if 8 * n > half:
            raise ValueError(f"tile of {8 * n} bytes per array exceeds half the budget")
        if 8 * (off + n) > half:
            wave, off = wave + 1, 0
        out.append((wave, off))
        off += n
        off += (-off) % 32  # 256-byte aligned arrays
    return out


def packed_offsets(tiles, align: int = 32):
    """Element offsets of the tiles' arrays packed one after another in one
    buffer, each starting on a multiple of `align` elements (32: 256 bytes,
    as separately allocated device arrays start); returns (offsets, total)."""
    offs, total = [], 0
    for t in tiles:
        offs.append(total)
        total += t.nelem
        total += (-total) % align
    return offs, total
