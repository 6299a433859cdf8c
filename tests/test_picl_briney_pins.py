"""Pins of the oracle's Briney added-mass model (AMFORCE 2).

The helper functions (libpicl/user_files/Added_Mass/
ppiclf_user_AddedMass_functions.f, _subroutines.f) are compiled from the
reference tree here (oracle/Makefile `ref`, amdflang) and their outputs kept
as golden vectors (tests/golden/briney_am.json, tests/golden/
make_briney_am.py).  The restatement must reproduce them to a few ulps: the
reference's build uses gfortran, whose expansion of integer powers can
differ from flang's in the last bit.

Beyond the vectors, the mathematics the functions claim:
* the distant-neighbour corrections are the integrals
  alpha * int_m^inf x^2 (B_11 + 2 B_22) dx of the curves (Briney et al. 2024,
  appendix D) -- II exactly so; IA with its volume-fraction term lacking the
  factor 70.7731 that B11_11's carries (the closed form's coefficients
  0.333333, 0.348933, 0.146105 are (10 m^2 + 5 a m + a^2) / 30 without it);
* the numerical branch is Simpson's rule on the same integrand weighted by
  the radial distribution function, over the dimensional distance (so a
  factor of the radius apart from the analytical one);
* the radial distribution function is 0 inside contact, continuous at its
  branch point r*, and tends to 1;
* resistance_pair rotates the axial curves: along the pair axis n the
  blocks act as B_11 n, across it as B_22 t, and each block is symmetric.

Not pinned beyond restating the code: the W store's reads of the particle
loop's left-over rphif and relative velocity (ppiclf_user.f:934-990).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "briney_am.json")))
REF = os.path.join(ROOT, "oracle", "_ref", "libppiclf_am_ref.so")
ULPS = 4e-15


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def built():
    orc.build()


def test_curves_equal_the_compiled_reference():
    worst = max(rel(orc.am_curve(name, d, al), v) for name, d, al, v in GOLD["curves"])
    assert worst <= ULPS, worst


def test_distant_neighbour_corrections_equal_the_compiled_reference():
    worst = max(rel(orc.am_curve(name, rmax, rad, al), v) for name, rmax, rad, al, v in GOLD["corrections"])
    assert worst <= ULPS, worst


def test_radial_distribution_equals_the_compiled_reference():
    for r, al, v in GOLD["rdf"]:
        g = orc.am_curve("rdf", r, al)
        assert abs(g - v) <= ULPS * max(abs(v), 1.0), (r, al, g, v)


def test_resistance_pair_equals_the_compiled_reference():
    for x, y, z, al, rad, Rm in GOLD["resistance_pair"]:
        Rf = np.array(Rm)  # R(6,6), row i = R(i+1, :)
        R = orc.am_resistance_pair(x, y, z, al, rad)
        scale = np.abs(Rf[:3, :]).max()
        assert np.abs(R - Rf[:3, :]).max() <= ULPS * scale
        # the lower blocks repeat the upper ones (R(4:6,1:3) = B12, R(4:6,4:6) = B11)
        assert np.array_equal(Rf[3:, :3], Rf[:3, 3:]) and np.array_equal(Rf[3:, 3:], Rf[:3, :3])


@pytest.mark.skipif(not os.path.exists(REF), reason="compiled reference not built (make -C oracle ref)")
def test_random_inputs_equal_the_compiled_reference():
    L = C.CDLL(REF)
    D = C.c_double

    def call(name, *a):
        f = getattr(L, name + "_")
        f.restype = D
        return f(*[C.byref(D(float(v))) for v in a])

    rng = np.random.default_rng(3)
    for _ in range(300):
        d, al = rng.uniform(2.0, 30.0), rng.uniform(0.0, 0.4)
        for name in ("b11_11", "b11_22", "b12_11", "b12_22"):
            assert rel(orc.am_curve(name, d, al), call(name, d, al)) <= ULPS
        m, rad = rng.uniform(3.49, 40.0), rng.uniform(1e-5, 2e-4)
        for name in ("ia_analytical", "ii_analytical"):
            assert rel(orc.am_curve(name, m * rad, rad, al), call(name, m * rad, rad, al)) <= ULPS


@pytest.mark.parametrize("al", [0.05, 0.2, 0.4])
@pytest.mark.parametrize("m", [3.49, 4.0, 6.0, 8.0])
def test_corrections_are_integrals_of_the_curves(m, al):
    a32 = float(np.float32(2.0936))
    c11 = lambda x: (al * al - float(np.float32(0.0902)) * al) * float(np.float32(70.7731)) / (x + a32) ** 6
    # IA: B11_11 with its volume-fraction term divided by 70.7731, as the closed form has it
    fa = lambda x: x * x * ((orc.am_curve("b11_11", x, al) - c11(x) + c11(x) / float(np.float32(70.7731))) +
                            2.0 * orc.am_curve("b11_22", x, al))
    fi = lambda x: x * x * (orc.am_curve("b12_11", x, al) + 2.0 * orc.am_curve("b12_22", x, al))
    qa = al * quad(fa, m, np.inf, limit=200, epsabs=0.0, epsrel=1e-11)[0]
    qi = al * quad(fi, m, np.inf, limit=200, epsabs=0.0, epsrel=1e-11)[0]
    assert rel(orc.am_curve("ia_analytical", m, 1.0, al), qa) <= 1e-3
    assert rel(orc.am_curve("ii_analytical", m, 1.0, al), qi) <= 1e-4


@pytest.mark.parametrize("al", [0.1, 0.3])
@pytest.mark.parametrize("m", [2.0, 3.0])
def test_numerical_corrections_are_simpson_sums_over_the_distance(m, al):
    rad = 5.75e-5
    for name, (b1, b2) in (("ia_numerical", ("b11_11", "b11_22")), ("ii_numerical", ("b12_11", "b12_22"))):
        f = lambda d: al * (d / rad) ** 2 * orc.am_curve("rdf", d / (2.0 * rad), al) * (
            orc.am_curve(b1, d / rad, al) + 2.0 * orc.am_curve(b2, d / rad, al))
        q = quad(f, m * rad, 20.0 * rad, limit=400)[0]
        assert rel(orc.am_curve(name, m * rad, rad, al), q) <= 1e-4, name


@pytest.mark.parametrize("al", [0.05, 0.2, 0.4])
def test_radial_distribution_contact_continuity_and_limit(al):
    assert orc.am_curve("rdf", 0.999, al) == 0.0
    assert orc.am_curve("rdf", 1.0, al) > 1.0  # contact value of a dense hard-sphere fluid
    eta = al
    rstar = (np.float32(2.0116) - np.float32(1.0647) * eta) + (np.float32(0.0538) * eta) * eta
    lo, hi = orc.am_curve("rdf", rstar * (1 - 1e-12), al), orc.am_curve("rdf", rstar * (1 + 1e-12), al)
    assert abs(lo - hi) <= 1e-9 * abs(hi)
    assert abs(orc.am_curve("rdf", 40.0, al) - 1.0) < 1e-6


def test_resistance_pair_rotates_the_axial_curves():
    rng = np.random.default_rng(5)
    for _ in range(50):
        v = rng.normal(size=3)
        rad, al = 6e-5, rng.uniform(0.0, 0.4)
        x = v / np.linalg.norm(v) * rad * rng.uniform(2.0, 8.0)
        d = np.linalg.norm(x) / rad
        R = orc.am_resistance_pair(*x, al, rad)
        n = x / np.linalg.norm(x)
        t = np.cross(n, [1.0, 0.0, 0.0])
        t /= np.linalg.norm(t)
        for blk, (c1, c2) in ((R[:, :3], ("b11_11", "b11_22")), (R[:, 3:], ("b12_11", "b12_22"))):
            s = max(abs(orc.am_curve(c1, d, al)), abs(orc.am_curve(c2, d, al)))
            assert np.abs(blk @ n - orc.am_curve(c1, d, al) * n).max() <= 1e-14 * s
            assert np.abs(blk @ t - orc.am_curve(c2, d, al) * t).max() <= 1e-14 * s
            assert np.abs(blk - blk.T).max() <= 1e-14 * s
