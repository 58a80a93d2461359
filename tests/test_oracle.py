"""Pins the CPU oracle (oracle/oracle.c) -- CPU only.

The oracle is checked against
  * the reference's own trunk/old_sort.f, compiled from the reference
    checkout (oracle/_ref), wherever that exists: all 24 4D and 720 6D
    permutations, sort and sort-accumulate, KAT and random inputs;
  * the committed golden digests made from that library
    (tests/golden/make_golden.py), so the pin holds without the reference;
  * mathematics: the NumPy closed form, inverse-permutation round trips,
    rank-2 = matrix transpose, identity = scale, linearity, brute force;
  * planted mistakes: a wrong index rule or a fused multiply-add must fail
    one of the checks above (checked on the CPU by perturbing the oracle's
    inputs, never the product).
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden_digests.json")
SMALL = os.path.join(HERE, "golden", "golden_small.npz")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


# --------------------------------------------------------------------------
# inputs
# --------------------------------------------------------------------------

def test_uniform_numpy_matches_c():
    for seed in (0, 1, 12345, 2**63 + 7):
        a = O.uniform(10007, seed)
        b = O.uniform_c(10007, seed)
        assert same(a, b)
        assert a.min() >= -1.0 and a.max() < 1.0
    # index-addressable: a slice equals the stream from an offset
    assert same(O.uniform(100, 5)[37:], O.uniform(63, 5, start=37))


def test_kat_matches_tester_formula():
    # x(i,j,k,l) = i + j*a + k*a*b + l*a*b*c, first index fastest
    a, b, c, d = 3, 4, 5, 6
    x = O.kat([a, b, c, d]).reshape(d, c, b, a)
    for (i, j, k, l) in [(1, 1, 1, 1), (3, 4, 5, 6), (2, 1, 3, 5)]:
        assert x[l - 1, k - 1, j - 1, i - 1] == i + j * a + k * a * b + l * a * b * c


# --------------------------------------------------------------------------
# oracle vs the reference's compiled old_sort.f
# --------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checkout absent")


@needs_ref
@pytest.mark.parametrize("dims", [(3, 4, 5, 2), (1, 7, 2, 9), (6, 5, 4, 3)])
def test_old_sort_4_vs_reference(dims):
    n = int(np.prod(dims))
    for x in (O.kat(list(dims)), O.uniform(n, 3)):
        for perm in itertools.permutations([1, 2, 3, 4]):
            for f in (1.0, -1.0, 0.5, 37.0):
                for acc in (False, True):
                    y0 = O.uniform(n, 4)
                    a = O.old_sort(x, y0.copy(), dims, perm, f, acc)
                    b = O.ref_old_sort(x, y0.copy(), dims, perm, f, acc)
                    assert same(a, b), (dims, perm, f, acc)


@needs_ref
def test_old_sort_6_vs_reference_all_720():
    dims = (2, 3, 2, 1, 3, 2)
    n = int(np.prod(dims))
    x = O.uniform(n, 5)
    for perm in itertools.permutations(range(1, 7)):
        for acc in (False, True):
            y0 = O.uniform(n, 6)
            a = O.old_sort(x, y0.copy(), dims, perm, 37.0, acc)
            b = O.ref_old_sort(x, y0.copy(), dims, perm, 37.0, acc)
            assert same(a, b), (perm, acc)


@needs_ref
def test_ranks_2_and_8_vs_reference():
    for dims, R in (((5, 7), 2), ((2, 1, 2, 3, 1, 2, 2, 3), 8)):
        n = int(np.prod(dims))
        x = O.uniform(n, 8)
        for perm in list(itertools.permutations(range(1, R + 1)))[:60]:
            for acc in (False, True):
                y0 = O.uniform(n, 9)
                a = O.sort_n(x, y0.copy(), dims, perm, -0.5, 1.0 if acc else 0.0, int(acc))
                b = O.ref_old_sort(x, y0.copy(), dims, perm, -0.5, acc)
                assert same(a, b), (dims, perm, acc)


# --------------------------------------------------------------------------
# oracle vs committed golden vectors (made from the reference)
# --------------------------------------------------------------------------

def _golden_cases(max_elems):
    with open(GOLD) as f:
        data = json.load(f)
    return [c for c in data["cases"] if int(np.prod(c["dims"])) <= max_elems]


def _gen(spec, n, dims):
    if spec["kind"] == "uniform":
        return O.uniform(n, spec["seed"])
    if spec["kind"] == "kat":
        return O.kat(list(dims))
    return np.zeros(n)


def golden_inputs(c):
    n = int(np.prod(c["dims"]))
    x = _gen(c["input"], n, c["dims"])
    y0 = _gen(c["y0"], n, c["dims"]) if c["beta"] != 0.0 else np.zeros(n)
    return x, y0


def test_oracle_reproduces_golden_digests():
    cases = _golden_cases(2_000_000)
    assert len(cases) > 1000
    for c in cases:
        x, y0 = golden_inputs(c)
        perm = [int(t) for t in c["perm"]]
        y = O.old_sort(x, y0, c["dims"], perm, c["factor"], c["beta"] != 0.0)
        assert digest(y) == c["sha256"], c


def test_small_golden_arrays_match_digests():
    data = np.load(SMALL)
    with open(GOLD) as f:
        cases = json.load(f)["cases"]
    for key in data.files:
        i = int(key[4:])
        assert digest(data[key]) == cases[i]["sha256"]


# --------------------------------------------------------------------------
# mathematics
# --------------------------------------------------------------------------

@pytest.mark.parametrize("R", [2, 3, 4, 5, 6])
def test_closed_form(R):
    rng = np.random.default_rng(R)
    for _ in range(40):
        dims = [int(v) for v in rng.integers(1, 5, size=R)]
        perm = [int(v) + 1 for v in rng.permutation(R)]
        n = int(np.prod(dims))
        x, y0 = O.uniform(n, 1), O.uniform(n, 2)
        f = float(rng.uniform(-3, 3))
        assert same(O.sort_n(x, np.zeros(n), dims, perm, f, 0.0, 0), O.closed_form(x, dims, perm, f))
        assert same(O.sort_n(x, y0.copy(), dims, perm, f, 1.0, 1), O.closed_form(x, dims, perm, f, y0))


def _axpby_closed_form(x, y0, dims, perm, f, g):
    # the five-case form's general case, sorted = g*sorted + f*unsorted
    # (trunk/generator2.py:396, 431-432; generator6.py:205-208): two products,
    # each rounded, then one rounded sum -- three separate NumPy operations
    p = O.closed_form(x, dims, perm, f)
    q = np.asarray(y0) * np.float64(g)
    return q + p


@pytest.mark.parametrize("R", [2, 3, 4, 5, 6, 7, 8])
def test_general_g_mode_closed_form(R):
    # oracle mode 2 (s = g*s + f*u) against the closed form, for g values
    # whose products are exact (-2), inexact (0.7) and odd (3)
    rng = np.random.default_rng(100 + R)
    for trial in range(12):
        dims = [int(v) for v in rng.integers(1, 5 if R <= 6 else 3, size=R)]
        perm = [int(v) + 1 for v in rng.permutation(R)]
        n = int(np.prod(dims))
        x, y0 = O.uniform(n, 7 + trial), O.uniform(n, 8 + trial)
        for f in (1.0, -0.5, 0.3, 37.0):
            for g in (-2.0, 0.7, 3.0):
                got = O.sort_n(x, y0.copy(), dims, perm, f, g, 2)
                assert same(got, _axpby_closed_form(x, y0, dims, perm, f, g)), (dims, perm, f, g)


def test_general_g_mode_planted_mistakes():
    # each plausible mistake in s = g*s + f*u changes the result on these
    # inputs, so the closed-form check above has teeth
    dims, perm = (3, 4, 5, 2), (2, 4, 1, 3)
    n = int(np.prod(dims))
    x, y0 = O.uniform(n, 51), O.uniform(n, 52)
    f, g = 0.3, 0.7
    good = O.sort_n(x, y0.copy(), dims, perm, f, g, 2)
    assert same(good, _axpby_closed_form(x, y0, dims, perm, f, g))
    p = O.closed_form(x, dims, perm, f)
    # f and g swapped
    assert not same(good, _axpby_closed_form(x, y0, dims, perm, g, f))
    # g applied to the input instead of the output
    assert not same(good, y0 + O.closed_form(x, dims, perm, f * g))
    # accumulate instead of scale-and-add (g ignored)
    assert not same(good, y0 + p)
    # the two products fused into one rounding: g*y0 + p evaluated in
    # extended precision and rounded once
    fused = (np.asarray(y0, dtype=np.longdouble) * np.longdouble(g) + p.astype(np.longdouble))
    assert not same(good, fused.astype(np.float64))
    # the permutation applied to the old output as well
    assert not same(good, O.closed_form(y0, dims, perm, g) + p)
    # and mode 2 with g = 1 is the accumulate of mode 1
    assert same(O.sort_n(x, y0.copy(), dims, perm, f, 1.0, 2), O.sort_n(x, y0.copy(), dims, perm, f, 1.0, 1))


def test_closed_form_at_matches_closed_form():
    # the sampled closed form used by the full-size GPU tests
    rng = np.random.default_rng(3)
    for dims in ((5, 6, 7, 8), (3, 4, 2, 5, 3, 2)):
        n = int(np.prod(dims))
        x, y0 = O.uniform(n, 1), O.uniform(n, 2)
        for perm in list(itertools.permutations(range(1, len(dims) + 1)))[::7]:
            pos = rng.integers(0, n, size=200)
            full = O.closed_form(x, dims, perm, -0.5, y0)
            assert same(O.closed_form_at(x, dims, perm, -0.5, pos, y0), full[pos])
            assert same(O.closed_form_at(x, dims, perm, 37.0, pos), O.closed_form(x, dims, perm, 37.0)[pos])
            # the five-case form on gathered values
            ia = O.source_positions(dims, perm, pos)
            assert same(O.sampled_values(x[ia], 0.3, y0[pos], 0.7),
                        _axpby_closed_form(x, y0, dims, perm, 0.3, 0.7)[pos])
    # a swapped permutation is caught
    dims, perm = (5, 6, 7, 8), (4, 3, 2, 1)
    pos = np.arange(5 * 6 * 7 * 8)
    x = O.uniform(len(pos), 4)
    assert not same(O.closed_form_at(x, dims, (3, 4, 2, 1), 1.0, pos), O.closed_form(x, dims, perm, 1.0))


def test_brute_force_index_rule():
    # literal transcription of ib for one tiny case per perm
    dims = (2, 3, 4, 2)
    n = int(np.prod(dims))
    x = O.uniform(n, 17)
    for perm in itertools.permutations([1, 2, 3, 4]):
        y = np.full(n, np.nan)
        for j1, j2, j3, j4 in itertools.product(*[range(1, d + 1) for d in dims]):
            jd = dims
            idv = (j1, j2, j3, j4)
            ia = j4 + jd[3] * (j3 - 1 + jd[2] * (j2 - 1 + jd[1] * (j1 - 1)))
            i, j, k, l = perm
            ib = idv[l - 1] + jd[l - 1] * (idv[k - 1] - 1 + jd[k - 1] * (
                idv[j - 1] - 1 + jd[j - 1] * (idv[i - 1] - 1)))
            y[ib - 1] = x[ia - 1] * 2.5
        assert same(O.old_sort(x, np.zeros(n), dims, perm, 2.5, False), y)


def test_inverse_round_trip_and_transpose():
    dims = (4, 5, 3, 6)
    n = int(np.prod(dims))
    x = O.uniform(n, 21)
    for perm in itertools.permutations([1, 2, 3, 4]):
        y = O.old_sort(x, np.zeros(n), dims, perm, 1.0, False)
        odims = [dims[p - 1] for p in perm]
        inv = [0] * 4
        for q, p in enumerate(perm):
            inv[p - 1] = q + 1
        z = O.old_sort(y, np.zeros(n), odims, inv, 1.0, False)
        assert same(z, x), perm
    m = O.uniform(35, 3)
    assert same(O.sort_n(m, np.zeros(35), (5, 7), (2, 1), 1.0, 0.0, 0), m.reshape(5, 7).T.reshape(-1))
    assert same(O.old_sort(x, np.zeros(n), dims, (1, 2, 3, 4), -0.5, False), -0.5 * x)


def test_two_roundings_not_fma():
    # y + x*f must round the product first; a fused multiply-add differs here
    x = np.array([1.0 + 2.0 ** -30] * 4)
    f = 1.0 + 2.0 ** -30
    y0 = np.array([-1.0] * 4)
    y = O.old_sort(x, y0.copy(), (1, 1, 2, 2), (1, 2, 3, 4), f, True)
    two = np.float64(x[0] * f) + y0[0]
    assert same(y, np.full(4, two))
    fma = np.longdouble(x[0]) * np.longdouble(f) + np.longdouble(y0[0])
    assert np.float64(fma) != two  # the planted mistake is detectable


def test_planted_mistakes_are_caught():
    # a swapped pair of permutation indices or a transposed operand must
    # change the output -- i.e. the checks above have teeth
    dims = (3, 4, 5, 2)
    n = int(np.prod(dims))
    x = O.uniform(n, 31)
    good = O.old_sort(x, np.zeros(n), dims, (4, 3, 2, 1), 1.0, False)
    assert not same(good, O.old_sort(x, np.zeros(n), dims, (4, 3, 1, 2), 1.0, False))
    assert not same(good, O.closed_form(x, dims[::-1], (4, 3, 2, 1), 1.0))


def test_generated_kernel_port_equals_old_sort():
    # trans_<P>_loop_<L>_omp_any_f (generator.py:107-159) with the packaged
    # loop table (makelib.py:39-62) computes old_sort_4's values
    dims = (7, 6, 5, 4)
    n = int(np.prod(dims))
    x = O.uniform(n, 41)
    for perm, loop in O.KNIGHT_4D.items():
        p = [int(c) for c in perm]
        lp = [int(c) for c in loop]
        for mode, acc in ((0, False), (1, True)):
            y0 = O.uniform(n, 42)
            a = O.trans(x, y0.copy(), dims, p, lp, 0.5, 0.0, mode)
            b = O.old_sort(x, y0.copy(), dims, p, 0.5, acc)
            assert same(a, b)
        # the five-case general form (generator2.py:383-439) in the same loop nest
        y0 = O.uniform(n, 44)
        a = O.trans(x, y0.copy(), dims, p, lp, 0.3, 0.7, 2)
        assert same(a, _axpby_closed_form(x, y0, dims, p, 0.3, 0.7))
    # 6D, every loop order of one permutation, collapse(3)
    d6 = (3, 2, 4, 2, 3, 2)
    n6 = int(np.prod(d6))
    x6 = O.uniform(n6, 43)
    ref = O.old_sort(x6, np.zeros(n6), d6, (6, 5, 4, 3, 2, 1), -1.0, False)
    for loop in itertools.permutations(range(1, 7)):
        y = O.trans(x6, np.zeros(n6), d6, (6, 5, 4, 3, 2, 1), loop, -1.0, 0.0, 0)
        assert same(y, ref)


# --------------------------------------------------------------------------
# the contraction around the sort (TTDT)
# --------------------------------------------------------------------------

def test_ttdt_oracle_matches_the_contraction_it_computes():
    # ccsd_t2_8 (branches/old/python/macaroni/ccsd_t2_8_skeleton.F:162-183):
    # i0(p3,p4,h1,h2) = Sum(p5,p6) v(p3,p4,p5,p6) * t(p5,p6,h1,h2); NWChem
    # dims slowest -> fastest, so t is T[p5][p6][h1][h2] in C order.  With
    # integer-valued data every partial sum is exact, whatever the order.
    rng = np.random.default_rng(5)
    d1, d2, d3, d4, d5, d6 = 3, 4, 5, 2, 6, 3
    T = rng.integers(-8, 9, size=(d5, d6, d1, d2)).astype(np.float64)
    V = rng.integers(-8, 9, size=(d3, d4, d5, d6)).astype(np.float64)
    R0 = rng.integers(-8, 9, size=(d3, d4, d1, d2)).astype(np.float64)
    want = np.einsum("abef,efij->abij", V, T)
    got = O.ttdt(T.reshape(-1), (d5, d6, d1, d2), (4, 3, 2, 1), V.reshape(-1), (d3, d4, d5, d6),
                 (2, 1, 4, 3), (d4, d3, d2, d1), (2, 1, 4, 3), d1 * d2, d3 * d4, d5 * d6)
    assert same(got, want.reshape(-1))
    got = O.ttdt(T.reshape(-1), (d5, d6, d1, d2), (4, 3, 2, 1), V.reshape(-1), (d3, d4, d5, d6),
                 (2, 1, 4, 3), (d4, d3, d2, d1), (2, 1, 4, 3), d1 * d2, d3 * d4, d5 * d6,
                 alpha=-2.0, beta=1.0, c0=R0.reshape(-1))
    assert same(got, (R0 - 2.0 * want).reshape(-1))
    # a wrong sort of t (2143 for 4321) is caught
    bad = O.ttdt(T.reshape(-1), (d5, d6, d1, d2), (3, 4, 1, 2), V.reshape(-1), (d3, d4, d5, d6),
                 (2, 1, 4, 3), (d4, d3, d2, d1), (2, 1, 4, 3), d1 * d2, d3 * d4, d5 * d6)
    assert not same(bad, want.reshape(-1))
