"""GPU parity of the quasi-steady force fluctuations (QSFLUCT 1 Lattanzi,
2 Osnes; libpicl/user_files/ppiclf_user_Fluctuations.f:32-418) against the
CPU oracle.

The box means (ppiclf_user_EvalNearestNeighbor.f:318-360) run over the same
cube-grid neighbour search as Briney's pair terms, summed with compensated
(TwoSum) sums so that their order does not show; the six uniforms of a
particle's call are a counter-based hash of (seed, call, particle, draw) on
both sides.  The SDE step calls log/cos/sin/exp/pow, so states are held to
1e-12 per element (positions, velocities and momenta as vectors), as the
other particle models are.
The particles carry a seeded velocity scatter of a few m/s, so the granular
temperature, and with it the fluctuations, is far from zero.
"""
import numpy as np
import pytest

from rocflupicl_amd import configs, rflu

pytestmark = pytest.mark.gpu

TOL = 1e-12
FLOOR = 1e-9
D0 = 115e-6


def close(a, b, tol=TOL, vectors=(), floor=FLOOR):
    """Largest element-wise relative error, |a - b| / max(|b|, FLOOR * the
    column's largest magnitude); each group of columns in `vectors` (a
    position, a velocity, a momentum) is measured as one vector, the error
    vector's length over the vector's length, with the same floor."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.ndim == 1:
        a, b = a[:, None], b[:, None]
    grouped = {c for g in vectors for c in g}
    groups = [list(g) for g in vectors] + [[c] for c in range(b.shape[1]) if c not in grouped]
    err = 0.0
    for g in groups:
        num = np.linalg.norm(a[:, g] - b[:, g], axis=1)
        mag = np.linalg.norm(b[:, g], axis=1)
        scale = np.maximum(np.maximum(mag, floor * mag.max()), 1e-300)
        err = max(err, float((num / scale).max()))
    return err <= tol, err


def cloud(npart=800, vel_sd=3.0, mono=False, **pp):
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=npart, jitter=0.05, seed=4, phi=0.3)
    y = y.copy()
    if vel_sd > 0.0:
        y[:, 3:6] = np.random.default_rng(9).normal([20.0, 1.0, -1.0], vel_sd, size=(npart, 3))
    if mono:
        pc = pc.copy()
        pc[:, 1] = D0
        pc[:, 2] = np.pi / 6.0 * D0 ** 3
    picl.update(qsFluct=1, fluctSeed=3)
    picl.update(pp)
    return mesh, params, ic, y, pc, picl


def gpu_run(mesh, params, ic, y, pc, picl, steps, graph=True):
    g = rflu.Region(mesh, **params)
    g.set_state(ic(g.mesh_arrays()["cofg"]))
    g.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(steps, use_graph=graph)
    return g


def oracle_run(oracle_mod, mesh, params, ic, y, pc, picl, steps):
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y, pc, **picl)
    o.step(steps)
    return o


def check(g, o, floor=FLOOR):
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = close(gp[k], op[k], vectors=([0, 1, 2], [3, 4, 5]), floor=floor)
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv(), vectors=([1, 2, 3],), floor=floor)
    assert ok, err
    return op


def fluct_share(oracle_mod, case, op, steps):
    """How far the accelerations move from the run without fluctuations."""
    mesh, params, ic, y, pc, picl = case
    o0 = oracle_run(oracle_mod, mesh, params, ic, y, pc, dict(picl, qsFluct=0), steps)
    a0, a1 = o0.picl_particles()["ydot"][:, 3:6], op["ydot"][:, 3:6]
    return np.abs(a1 - a0).max() / np.abs(a0).max()


def test_lattanzi_fluctuations_match_oracle(oracle_mod):
    case = cloud()
    g = gpu_run(*case, steps=4)
    o = oracle_run(oracle_mod, *case, steps=4)
    op = check(g, o)
    assert np.abs(o.picl_rprop()[:, 24:27]).max() > 0.0
    assert fluct_share(oracle_mod, case, op, 4) > 1e-3


def test_osnes_fluctuations_with_adaptive_box_and_periodic_images(oracle_mod):
    case = cloud(qsFluct=2, qsFluctAdapt=1, periodicY=1, periodicZ=1)
    g = gpu_run(*case, steps=4)
    o = oracle_run(oracle_mod, *case, steps=4)
    op = check(g, o)
    assert fluct_share(oracle_mod, case, op, 4) > 1e-3


def test_osnes_fluctuations_with_explicit_box_and_neighbour_width(oracle_mod):
    case = cloud(qsFluct=2, fluctHalfWidth=1.5e-4, neighborWidth=3.0e-4)
    g = gpu_run(*case, steps=3)
    o = oracle_run(oracle_mod, *case, steps=3)
    check(g, o)


def test_particles_at_rest_are_skipped_until_the_shock_arrives(oracle_mod):
    """Particles at rest in still gas are skipped by SetYdot (relative speed
    at most 1e-3 m/s, ppiclf_user.f:335) before any force, the fluctuation's
    step included: they keep their fluctuating force and take no box means,
    while the particles the shock has reached do."""
    case = cloud(vel_sd=0.0, qsFluct=2)
    g = gpu_run(*case, steps=2)
    o = oracle_run(oracle_mod, *case, steps=2)
    op = check(g, o)
    assert np.all(np.isfinite(op["y"]))


def test_fluctuations_with_briney_added_mass_and_collisions(oracle_mod):
    """The AirShock deck's particle physics as far as it is built: Osnes'
    fluctuations, Briney's added mass (monodisperse: its binary curves hold
    for equal spheres only) and collisions with tangential forces, periodic
    in y and z."""
    case = cloud(mono=True, qsFluct=2, amFlag=2, collisionalFlag=2, ksp=10.0, erest=0.7, periodicY=1, periodicZ=1)
    g = gpu_run(*case, steps=3)
    o = oracle_run(oracle_mod, *case, steps=3)
    # with collisions the pair forces come in opposite pairs and partly
    # cancel, and the GPU adds them in cube order, the oracle in particle-
    # index order: a few sums end far below their terms, so this case is
    # held to 1e-12 of each column's largest magnitude (per element it
    # measured up to 3e-12 on ydotc and 5e-11 on the cell momenta)
    check(g, o, floor=1.0)


def test_fluctuation_graph_replays_equal_eager_steps():
    case = cloud(qsFluct=2, qsFluctAdapt=1)
    a = gpu_run(*case, steps=3, graph=True)
    b = gpu_run(*case, steps=3, graph=False)
    pa, pb = a.picl_particles(), b.picl_particles()
    for k in ("y", "y1", "ydot", "ydotc"):
        assert np.array_equal(pa[k].view(np.uint64), pb[k].view(np.uint64)), k


@pytest.mark.parametrize("bad", [dict(qsFluct=3), dict(qsFluct=-1), dict(qsFluctAdapt=2), dict(fluctHalfWidth=-1.0)])
def test_invalid_fluctuation_settings_are_rejected(bad):
    mesh, params, ic, y, pc, picl = cloud(npart=50, **bad)
    g = rflu.Region(mesh, **params)
    g.set_state(ic(g.mesh_arrays()["cofg"]))
    with pytest.raises(Exception):
        g.picl_init(y, pc, **picl)
