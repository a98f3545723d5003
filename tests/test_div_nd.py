"""The calcnu kernels divide the dipolar potential by Nx*Ny*Nz (the reference's
psidd2 / (Nx * Ny * Nz), imag3d-mpicuda.cu:1776, real3d-mpicuda.cu calcnu)
without a divide instruction sequence: nd = b 2^e with b odd, q = RN(phi RN(1/b)),
r = fma(-q, b, phi), RN(q + r RN(1/b)), times 2^-e (nu_device.hpp div_nd).
These tests check that this sequence equals the IEEE quotient phi / nd, with
every FMA evaluated exactly (fractions) and rounded once, for the odd parts of
every supported grid product and for values next to exact multiples."""
import math
import random
from fractions import Fraction

import pytest

LENGTHS = (16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512)


def fma(a, b, c):
    # one rounding of the exact a*b + c (int / int true division rounds to nearest even)
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div_nd(phi, nd):
    b, p = nd, 1.0
    while b % 2 == 0:
        b //= 2
        p *= 0.5
    y = 1.0 / b
    q = phi * y
    r = fma(-q, float(b), phi)
    return fma(r, y, q) * p


def odd_part(n):
    while n % 2 == 0:
        n //= 2
    return n


def grid_products():
    return sorted({a * b * c for a in LENGTHS for b in LENGTHS for c in LENGTHS})


def test_every_odd_part_is_covered():
    odds = {odd_part(n) for n in grid_products()}
    assert odds == {1, 3, 5, 9, 15, 25, 27, 45, 75, 125}


@pytest.mark.parametrize("nd", [256 ** 3, 256 * 256 * 384, 192 ** 3, 48 * 80 * 96, 160 ** 3, 80 * 160 * 320,
                                96 * 192 * 384, 48 * 48 * 48, 80 * 80 * 80, 160 * 320 * 320])
def test_random_values_match_the_quotient(nd):
    rng = random.Random(nd)
    for _ in range(4000):
        phi = math.ldexp(rng.getrandbits(53) | (1 << 52), rng.randint(-120, 60)) * rng.choice((-1.0, 1.0))
        assert div_nd(phi, nd) == phi / nd, (phi, nd)


@pytest.mark.parametrize("b", [3, 5, 9, 15, 25, 27, 45, 75, 125])
def test_values_next_to_exact_multiples(b):
    # phi = b m 2^k (exact quotient) and its neighbours, where a quotient
    # computed from a rounded reciprocal is most likely to be off by one ulp
    rng = random.Random(b)
    nd = b * 2 ** 21
    for _ in range(1500):
        m = rng.getrandbits(rng.randint(1, 50)) | 1
        phi = float(b * m) * 2.0 ** rng.randint(-30, 30)
        for v in (phi, math.nextafter(phi, math.inf), math.nextafter(phi, -math.inf)):
            assert div_nd(v, nd) == v / nd, (v, nd)


def test_power_of_two_grids_are_a_scaling():
    for phi in (1.0, -3.25, 1e-300, 7.5e250):
        assert div_nd(phi, 256 ** 3) == phi * 2.0 ** -24
