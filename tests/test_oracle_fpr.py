"""Pins the oracle's false-positive-rate model (oracle.c orc_fpr_*, a restatement
of src/dtl/filter/blocked_bloomfilter/math.hpp:17-145 and fpr.hpp:11-60) against
mathematics that does not reuse its formulas:

* the Poisson pmf against the exact lambda^i e^-lambda / i! (50-digit decimal
  arithmetic), and its mass against 1;
* the blocked and sectorized Poisson mixtures, for integer bits per sector and
  no self-collision adjustment, against their closed forms: with
  a = (1 - 1/S)^(k/s), (1 - a^i)^k expands binomially and
  E[x^i] = exp(lambda (x - 1)) for i ~ Poisson(lambda) (the generating
  function), so the mixture is sum_j C(k, j) (-1)^j exp(lambda (a^j - 1));
* the self-collision adjustment against the expected number of distinct bits
  among k uniform draws, counted by enumerating every draw sequence;
* a sectorized filter with one sector against the blocked formula;
* the zoned inner term p_cache against its closed form where its truncation
  (j <= i) removes no mass, and the outer zoned mixture against a sum of
  p_cache weighted by exact Poisson probabilities;
* the cuckoo formula against its exponential form in the load.

A plausible slip in the oracle -- lgamma(i) for lgamma(i + 1), lambda = m / B,
a dropped self-collision term, the wrong power of a sector's rate, a zone
mean of i * S / B -- fails one of these.  No GPU is needed.
"""
import itertools
import math
from decimal import Decimal, getcontext

import pytest

import oracle as O

getcontext().prec = 50


def exact_pmf(lam, i):
    lam_d = Decimal(repr(lam))
    return float((lam_d ** i) * (-lam_d).exp() / Decimal(math.factorial(i)))


@pytest.mark.parametrize("lam", [0.25, 1.0, 3.0, 12.7, 42.0, 85.3, 512.0 / 12.0])
def test_poisson_pmf_is_exact(lam):
    L = O.lib()
    top = int(lam * 3 + 30)
    total = 0.0
    for i in range(top):
        got = L.orc_poisson_pmf(lam, i)
        exp = exact_pmf(lam, i)
        total += got
        if exp > 1e-300:
            assert abs(got - exp) <= 1e-11 * exp, (lam, i, got, exp)
    assert abs(total - 1.0) < 1e-9, total


def mixture_closed_form(lam, a, k):
    """E[(1 - a^i)^k] for i ~ Poisson(lam), integer k."""
    return sum(math.comb(k, j) * (-1) ** j * math.exp(lam * (a ** j - 1.0)) for j in range(k + 1))


@pytest.mark.parametrize("B", [32, 64, 256, 512, 1024])
@pytest.mark.parametrize("bits_per_key", [4, 8, 12, 20])
@pytest.mark.parametrize("k", [1, 3, 6, 8])
def test_blocked_mixture_closed_form(B, bits_per_key, k):
    n = 100000
    m = n * bits_per_key
    lam = B * n / m
    got = O.lib().orc_fpr_blocked(m, n, float(k), B, 0, 1e-13)
    exp = mixture_closed_form(lam, (1.0 - 1.0 / B) ** k, k)
    assert abs(got - exp) <= 1e-8 * exp + 1e-12, (got, exp)


@pytest.mark.parametrize("B,S", [(512, 64), (512, 32), (256, 64), (1024, 128), (64, 8)])
@pytest.mark.parametrize("bits_per_key", [6, 10, 16])
@pytest.mark.parametrize("k_per_sector", [1, 2])
def test_sectorized_mixture_closed_form(B, S, bits_per_key, k_per_sector):
    s = B // S
    k = k_per_sector * s
    n = 100000
    m = n * bits_per_key
    lam = B * n / m
    got = O.lib().orc_fpr_sectorized(m, n, float(k), B, S, 0, 1e-13)
    exp = mixture_closed_form(lam, (1.0 - 1.0 / S) ** k_per_sector, k)
    assert abs(got - exp) <= 1e-8 * exp + 1e-12, (got, exp)


def expected_distinct(B, k):
    total = 0
    for seq in itertools.product(range(B), repeat=k):
        total += len(set(seq))
    return total / B ** k


@pytest.mark.parametrize("B,k", [(4, 2), (4, 3), (8, 3), (8, 4), (16, 3), (6, 4)])
def test_self_collision_adjustment_is_expected_distinct_bits(B, k):
    """With self collisions the blocked formula uses the expected number of
    distinct bits of k draws in place of k (math.hpp:42-44)."""
    kd = expected_distinct(B, k)
    n = 1000
    for bpk in (2, 4, 8):
        m = n * bpk
        a = O.lib().orc_fpr_blocked(m, n, float(k), B, 1, 1e-13)
        b = O.lib().orc_fpr_blocked(m, n, kd, B, 0, 1e-13)
        assert abs(a - b) <= 1e-12 * b, (a, b)
        # sectorized: the same adjustment per sector, with k / s draws per sector
        s = 2
        a = O.lib().orc_fpr_sectorized(m, n, float(k * s), B * s, B, 1, 1e-13)
        b = O.lib().orc_fpr_sectorized(m, n, kd * s, B * s, B, 0, 1e-13)
        assert abs(a - b) <= 1e-12 * b, (a, b)


@pytest.mark.parametrize("B", [32, 64, 512])
@pytest.mark.parametrize("k", [1, 4, 7, 16])
def test_one_sector_is_the_blocked_filter(B, k):
    for bpk in (5, 10, 16):
        n = 50000
        m = n * bpk
        for sc in (0, 1):
            a = O.lib().orc_fpr_sectorized(m, n, float(k), B, B, sc, 1e-5)
            b = O.lib().orc_fpr_blocked(m, n, float(k), B, sc, 1e-5)
            assert a == b


def test_config_dispatch():
    """fpr.hpp:11-60: B = word bits x w, S = B / s; blocked when S == B and
    z == 1, sectorized when S < B and z == 1, zoned when z > 1 (the
    self-collision form and epsilon = 1e-5 throughout)."""
    L = O.lib()
    m, n = 1 << 20, 1 << 16
    assert L.orc_fpr_config(m, n, 4, 4, 1, 1, 1) == L.orc_fpr_blocked(m, n, 4.0, 32, 1, 1e-5)
    assert L.orc_fpr_config(m, n, 8, 8, 8, 8, 1) == L.orc_fpr_sectorized(m, n, 8.0, 512, 64, 1, 1e-5)
    assert L.orc_fpr_config(m, n, 8, 4, 2, 1, 1) == L.orc_fpr_blocked(m, n, 8.0, 64, 1, 1e-5)
    assert L.orc_fpr_config(m, n, 6, 4, 16, 16, 2) == L.orc_fpr_zoned(m, n, 6.0, 512, 32, 2.0, 1, 1e-5)


@pytest.mark.parametrize("w,word_bits,z,k", [(16, 32, 2, 6), (8, 32, 2, 4), (16, 64, 4, 8), (16, 64, 2, 16)])
def test_zoned_inner_term_closed_form(w, word_bits, z, k):
    """p_cache(z, S, B, k, i) (math.hpp:99-113): the probability that all z
    zones of a block holding i keys report a hit, each zone's word holding
    j ~ Poisson(i z S / B) keys, j >= 1, with k / z bits per zone.  Where
    the truncation j <= i drops no mass, the inner sum is the closed form
    E[(1 - a^j)^(k/z)] (the j = 0 term is 0), raised to the power z."""
    S = word_bits
    B = w * word_bits
    kz = k // z
    a = (1.0 - 1.0 / S) ** kz
    for i in (80, 120, 200):
        ev = i * z * S / B
        ev_d = Decimal(repr(ev))
        term = (-ev_d).exp()
        mass = term
        for j in range(1, i + 1):
            term = term * ev_d / j
            mass += term
        assert 1 - mass < Decimal("1e-20")
        got = O.lib().orc_fpr_p_cache(z, S, B, float(k), i)
        exp = mixture_closed_form(ev, a, kz) ** z
        assert abs(got - exp) <= 1e-9 * exp, (i, got, exp)
    # a block with no keys never reports a hit; with keys, more keys -> more hits
    assert O.lib().orc_fpr_p_cache(z, S, B, float(k), 0) == 0.0
    vals = [O.lib().orc_fpr_p_cache(z, S, B, float(k), i) for i in range(1, 60)]
    assert all(b > a_ for a_, b in zip(vals, vals[1:]))


@pytest.mark.parametrize("bpk", [6, 10, 14])
def test_zoned_outer_mixture(bpk):
    """fpr_zoned (math.hpp:119-145): keys per block ~ Poisson(B n / m), each
    weighted by p_cache; the self-collision k_per_s it computes is unused."""
    L = O.lib()
    w, S, z, k = 16, 32, 2, 6
    B = w * S
    n = 20000
    m = n * bpk
    lam = B * n / m
    exp = 0.0
    mass = 0.0
    for i in range(int(lam * 6 + 60)):
        p = exact_pmf(lam, i)
        mass += p
        exp += p * L.orc_fpr_p_cache(z, S, B, float(k), i)
    assert mass > 1 - 1e-12
    got = L.orc_fpr_zoned(m, n, float(k), B, S, float(z), 1, 1e-13)
    assert abs(got - exp) <= 1e-9 * exp
    assert L.orc_fpr_zoned(m, n, float(k), B, S, float(z), 0, 1e-13) == got


def test_cuckoo_fpr_exponential_in_load():
    """math.hpp:198-205: 1 - (1 - 1/(2^t - 1))^(2 b load): the miss
    probabilities of two loads multiply, one 1-bit tag always matches, and
    the rate is -2 b load ln(1 - 1/(2^t - 1)) to first order in the load."""
    L = O.lib()
    for t, b in ((16, 4), (8, 2), (12, 4), (4, 1)):
        for l1, l2 in ((0.2, 0.3), (0.5, 0.45), (0.01, 0.9)):
            p1, p2, p12 = (L.orc_fpr_cuckoo(b, t, x) for x in (l1, l2, l1 + l2))
            assert abs((1 - p1) * (1 - p2) - (1 - p12)) < 1e-12
        tiny = 1e-6
        first = -2 * b * tiny * math.log1p(-1.0 / (2 ** t - 1))
        assert abs(L.orc_fpr_cuckoo(b, t, tiny) / first - 1) < 1e-5
    assert L.orc_fpr_cuckoo(4, 1, 0.5) == 1.0
