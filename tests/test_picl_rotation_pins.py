"""Pins of the oracle's COLLISIONS 3-4 laws (ppiclf_user_Torque.f:72-197,
ppiclf_user_Lift.f:61-181) against the classical results they reduce to.

* Torque_Hydro in the creeping regime (Re_r < 1, C_t = 16 pi / Re_r) is
  Stokes' rotational torque 8 pi mu a^3 (Omega_f - omega), a = d_p / 2,
  Omega_f = vorticity / 2, whatever the particle density; it stops where
  ppiclF stops (Re_r >= 100).
* Saffman's lift with Mei's correction tends to Saffman's result
  1.615 mu d^2 |v| sqrt(|Omega| / nu) as epsilon grows (J -> 1.0002), along
  v x Omega.
* Magnus' lift tends to rho d^3 / 8 (v x omega) as Re -> 0 (C_L -> 1), and
  vanishes without spin.

The material derivative of the vorticity (PICL_TEMP_Runge.F90:747-809) is
restated from the reference's expression; beyond its vorticity, which comes
from the LSQ gradients pinned exact on linear fields, it is parity unpinned.
"""
import numpy as np
import pytest

from oracle import oracle as orc


@pytest.fixture(scope="module", autouse=True)
def built():
    orc.build()


@pytest.mark.parametrize("rhop", [1000.0, 2500.0, 8000.0])
def test_creeping_torque_is_stokes(rhop):
    rng = np.random.default_rng(int(rhop))
    mu, rhof, dp = 1.8e-5, 1.2, 50e-6
    for _ in range(20):
        vort, spin = rng.normal(size=3) * 200.0, rng.normal(size=3) * 100.0
        rel = 0.5 * vort - spin
        assert rhof * dp * dp * np.linalg.norm(rel) / (4.0 * mu) < 1.0
        rmass = rhop * np.pi / 6.0 * dp ** 3
        t = orc.torque_hydro(vort, spin, rhof, dp, mu, rmass, rhop)
        stokes = 8.0 * np.pi * mu * (dp / 2.0) ** 3 * rel
        assert np.allclose(t, stokes, rtol=1e-13, atol=0.0)


def test_torque_stops_at_rotational_reynolds_100():
    mu, rhof, dp = 1.8e-5, 1.2, 1e-3
    w = 100.0 * 4.0 * mu / (rhof * dp * dp)  # Re_r = 100 for a spin w
    assert orc.torque_hydro([0, 0, 0], [0, 0, 0.99 * w], rhof, dp, mu, 1.0, 2500.0) is not None
    assert orc.torque_hydro([0, 0, 0], [0, 0, 1.01 * w], rhof, dp, mu, 1.0, 2500.0) is None


def test_saffman_lift_tends_to_saffman():
    mu, rhof, dp = 1.8e-5, 1.2, 100e-6
    v = np.array([0.01, 0.0, 0.0])  # slow slip, strong shear: large epsilon
    vort = np.array([0.0, 0.0, 5e4])
    rep = max(0.1, np.linalg.norm(v) * dp * rhof / mu)
    f = orc.lift(vort, [0, 0, 0], v, dp, rep, mu, rhof)
    eps = np.sqrt(np.linalg.norm(vort) * dp / np.linalg.norm(v) / rep)
    assert eps > 10.0
    saffman = 1.615 * mu * dp * dp * np.linalg.norm(v) * np.sqrt(np.linalg.norm(vort) / (mu / rhof))
    direction = np.cross(v, vort) / np.linalg.norm(np.cross(v, vort))
    assert np.allclose(f, 1.0002 * saffman * direction, rtol=1e-3)


def test_magnus_lift_at_vanishing_reynolds():
    mu, rhof, dp = 1.8e-5, 1.2, 100e-6
    v = np.array([1e-6, 2e-6, 0.5e-6])
    spin = np.array([3.0, -1.0, 2.0])
    rep = 1e-12  # C_L = 1 - d1 tanh(0.18 sqrt(Re)) -> 1
    f = orc.lift([0, 0, 0], spin, v, dp, rep, mu, rhof)
    assert np.allclose(f, rhof * dp ** 3 / 8.0 * np.cross(v, spin), rtol=1e-6)
    assert np.array_equal(orc.lift([0, 0, 0], [0, 0, 0], v, dp, rep, mu, rhof), np.zeros(3))
