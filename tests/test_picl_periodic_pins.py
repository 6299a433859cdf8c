"""Pins of the oracle's particle periodicity (PERIODICX/Y/Z): the linear wrap
of ppiclf_solve_InvokeLinearPeriodic (ppiclf_solve.f:1319-1353) after every
stage and the periodic ghost images of the Gaussian projection
(ppiclf_comm.f:1115-1530, CheckPeriodicBC :2136-2206).  CPU only.

What the geometry fixes, not a retyped formula: a wrapped coordinate differs
from the unwrapped run's by exactly one period and lies inside the range,
nothing else in the stage changes, the wrapped particles stay located while
without periodicity they are lost; a particle on the low face and one on the
high face are the same point of a periodic domain and project the same node
field; and the Gaussian filter conserves the deposit of a particle next to a
periodic face only when its images count.
"""
import numpy as np
import pytest

from rocflupicl_amd import configs

N, H = 12, 0.0005
LY = LZ = 4 * H


def stage_pair(oracle_mod, periodic, nstages=1):
    """A 12x4x4 curtain whose particles sit within 2e-6 of the y and z faces,
    moving outward at 300 m/s, so that the first stage carries them across."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=N, ny=4, nz=4, npart=64, jitter=0.05, seed=7)
    rng = np.random.default_rng(3)
    y = y.copy()
    side_y = rng.integers(0, 2, len(y))
    side_z = rng.integers(0, 2, len(y))
    eps_y = 2e-6 * rng.random(len(y))
    eps_z = 2e-6 * rng.random(len(y))
    y[:, 1] = np.where(side_y == 1, LY - eps_y, eps_y)
    y[:, 2] = np.where(side_z == 1, LZ - eps_z, eps_z)
    y[:, 4] = np.where(side_y == 1, 300.0, -300.0)
    y[:, 5] = np.where(side_z == 1, 300.0, -300.0)
    y[::4, 2] = 0.5 * LZ  # some stay clear of the z faces
    y[::4, 5] = 0.0
    picl.update(periodicY=periodic, periodicZ=periodic)
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y, pc, **picl)
    for s in range(1, nstages + 1):
        o.rk_stage(s, 1e-8)
    return o


def test_wrap_moves_by_one_period_and_nothing_else(oracle_mod):
    on = stage_pair(oracle_mod, 1).picl_particles()
    off = stage_pair(oracle_mod, 0).picl_particles()
    yo, yf = on["y"], off["y"]
    crossed = 0
    for d, length in ((1, LY), (2, LZ)):
        lo_out, hi_out = yf[:, d] < 0.0, yf[:, d] > length
        crossed += lo_out.sum() + hi_out.sum()
        # re-entry one period away, inside the range
        assert np.allclose(yo[lo_out, d], yf[lo_out, d] + length, rtol=0, atol=4 * np.spacing(length))
        assert np.allclose(yo[hi_out, d], yf[hi_out, d] - length, rtol=0, atol=4 * np.spacing(length))
        assert np.all((yo[:, d] >= 0.0) & (yo[:, d] <= length))
        inside = ~(lo_out | hi_out)
        assert np.array_equal(yo[inside, d], yf[inside, d])
    assert crossed >= 40
    # x, the velocities and the stage's other rows are untouched
    assert np.array_equal(yo[:, [0, 3, 4, 5]], yf[:, [0, 3, 4, 5]])
    for k in ("y1", "ydot", "ydotc"):
        assert np.array_equal(on[k], off[k]), k


def test_wrapped_particles_stay_located(oracle_mod):
    on = stage_pair(oracle_mod, 1, nstages=2)
    off = stage_pair(oracle_mod, 0, nstages=2)
    assert oracle_mod.lib().orc_picl_nlost(on.h) == 0
    assert np.all(on.picl_particles()["elem"] >= 0)
    assert oracle_mod.lib().orc_picl_nlost(off.h) >= 40


def gauss_nodes(oracle_mod, yp, periodic):
    h = 0.5
    mesh, params, ic, _, picl = configs.curtain(nx=16, ny=16, nz=16, h=h, npart=4)
    y = np.zeros((len(yp), 12))
    y[:, :3] = yp
    y[:, 6] = 300.0
    pc = np.stack([np.full(len(yp), 2500.0), np.full(len(yp), 1e-3), np.full(len(yp), 1e-3),
                   np.ones(len(yp))], axis=1)
    picl.update(projMode=1, filterWidth=2 * h, filterAlpha=1e-12, periodicY=periodic)
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y, pc, **picl)
    return o.picl_nodes(), o.cells()[0]


def test_low_and_high_face_are_the_same_point(oracle_mod):
    lo, _ = gauss_nodes(oracle_mod, [[4.1, 0.0, 3.9]], 1)
    hi, _ = gauss_nodes(oracle_mod, [[4.1, 8.0, 3.9]], 1)
    assert np.array_equal(lo, hi)
    lo0, _ = gauss_nodes(oracle_mod, [[4.1, 0.0, 3.9]], 0)
    hi0, _ = gauss_nodes(oracle_mod, [[4.1, 8.0, 3.9]], 0)
    assert not np.array_equal(lo0, hi0)


@pytest.mark.parametrize("yp", [0.1, 7.8])
def test_gaussian_deposit_is_conserved_at_a_periodic_face(oracle_mod, yp):
    for periodic in (1, 0):
        nodes, vol = gauss_nodes(oracle_mod, [[4.1, yp, 3.9]], periodic)
        total = (vol * (0.125 * nodes[:, 0, :]).sum(axis=1)).sum()
        if periodic:
            assert abs(total / 1e-3 - 1.0) < 1e-4
        else:
            assert total / 1e-3 < 0.9  # half the filter falls outside the box
