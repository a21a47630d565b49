"""The FP64 step divides by constants (Delta_T, C_m, tau_w, ratio_ER_cyt,
tau_IP3) through their reciprocals: q0 = RN(a y), then twice
q <- RN(q + RN(a - b q) y) with y = RN(1/b), each inner step one fused
multiply-add. The claim is that this is RN(a / b), the correctly rounded
quotient an IEEE division gives, for every finite a and every positive normal
b with a normal reciprocal (a zero or non-finite a takes q0 = a y, which is
a / b there). Here the method runs in exact rational arithmetic (each
operation rounded once, as the instructions round) against the exact
quotient rounded once, on the constants of the shipped configurations and on
constructed worst cases: quotients just below and above binade boundaries,
divisors at both ends of their binade, where the first estimate q0 can be
more than one ulp off."""

import json
from fractions import Fraction

import numpy as np
import pytest

from conftest import config_path


def _rn(x: Fraction) -> float:
    """x rounded to the nearest double (ties to even): Python's int/int
    true division rounds correctly."""
    return x.numerator / x.denominator


def _fma(a: float, b: float, c: float) -> float:
    return _rn(Fraction(a) * Fraction(b) + Fraction(c))


def div_rn(a: float, b: float) -> float:
    y = _rn(Fraction(1) / Fraction(b))
    q0 = float(np.float64(a) * np.float64(y))  # one IEEE product (keeps -0, inf, NaN)
    if a == 0.0 or not np.isfinite(a):
        return q0
    q1 = _fma(_fma(-b, q0, a), y, q0)
    return _fma(_fma(-b, q1, a), y, q1)


def _check(pairs):
    bad = []
    for a, b in pairs:
        want = _rn(Fraction(a) / Fraction(b))
        got = div_rn(a, b)
        if np.float64(got).view(np.uint64) != np.float64(want).view(np.uint64):
            bad.append((a, b, got, want))
    assert not bad, bad[:5]


def _config_divisors():
    out = set()
    for name in ("quick", "use_case", "bursting", "random_pool", "paper_sparse"):
        with open(config_path(name)) as f:
            cfg = json.load(f)
        for pop in ("neuron_exc", "neuron_inh"):
            p = cfg.get(pop, {})
            for k in ("Delta_T", "C_m", "tau_w"):
                if k in p:
                    out.add(float(p[k]))
        a = cfg.get("astrocyte", {})
        for k in ("ratio_ER_cyt", "tau_IP3"):
            if k in a:
                out.add(float(a[k]))
    return sorted(out)


def test_config_constants_with_random_numerators():
    rng = np.random.default_rng(1)
    divisors = _config_divisors()
    assert len(divisors) >= 5
    pairs = []
    for b in divisors:
        # tau_IP3 is drawn per astrocyte around its configured value
        for bb in (b, b * rng.uniform(0.9, 1.05)):
            mant = rng.uniform(1.0, 2.0, 1500)
            expo = rng.integers(-60, 61, 1500)
            sign = rng.choice([-1.0, 1.0], 1500)
            pairs += [(float(s * m * 2.0 ** int(e)), float(bb)) for s, m, e in zip(sign, mant, expo)]
    _check(pairs)


def test_quotients_at_binade_edges_and_divisors_at_binade_ends():
    rng = np.random.default_rng(2)
    pairs = []
    ulp = 2.0 ** -52
    divisors = [2.0 - k * ulp for k in range(1, 40)] + [1.0 + k * ulp for k in range(1, 40)]
    divisors += list(rng.uniform(1.0, 2.0, 40)) + [1.5, 1.75, 1.9999999, 1.0000001]
    for b in divisors:
        for s in range(-3, 4):
            scale = 2.0 ** s
            # q = a / b just below 2 * scale, at it, just above it, and around 1
            for target in (2.0 - 3 * ulp, 2.0 - ulp, 2.0, 1.0, 1.0 + ulp, 1.5):
                a0 = _rn(Fraction(target) * Fraction(scale) * Fraction(b))
                for k in range(-3, 4):
                    a = float(np.nextafter(a0, np.inf if k > 0 else -np.inf)) if k else a0
                    for _ in range(abs(k) - 1):
                        a = float(np.nextafter(a, np.inf if k > 0 else -np.inf))
                    pairs.append((a, float(b)))
                    pairs.append((-a, float(b)))
    _check(pairs)


def test_random_divisors_over_wide_ranges():
    rng = np.random.default_rng(3)
    n = 12000
    b = rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-200, 201, n).astype(np.float64)
    a = rng.uniform(-2.0, 2.0, n) * 2.0 ** rng.integers(-300, 301, n).astype(np.float64)
    _check(list(zip(a.tolist(), b.tolist())))


@pytest.mark.parametrize("a", [0.0, -0.0, np.inf, -np.inf, np.nan])
def test_zero_and_non_finite_numerators_take_the_product(a):
    for b in (2.0, 130.0, 0.185):
        got = div_rn(a, b)
        want = np.float64(a) / np.float64(b)
        if np.isnan(want):
            assert np.isnan(got)
        else:
            assert np.float64(got).view(np.uint64) == want.view(np.uint64)
