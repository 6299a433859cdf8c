"""GPU parity of the ppiclF coupling against the CPU oracle.

Integer work (element ids, particle counts) is compared exactly, and so is
everything computed without pow/exp/log: the particle set-up (projection,
piclVF, scaled cv), the gas side of the coupled stage (grad rho_g, the 20
interpolation inputs) and the interpolated values.  The particle forces call
pow, exp and log, whose device versions may differ from glibc by an ulp, so
particle and gas states after coupled stages are compared element by
element with a relative tolerance of 1e-12: |a - b| / max(|b|, FLOOR times
the column's largest magnitude), the floor only keeping values that are zero
but for rounding (a transverse velocity, a force far from the cloud) from
dividing by their own noise.
"""
import numpy as np
import pytest

from rocflupicl_amd import configs, rflu

pytestmark = pytest.mark.gpu

TOL = 1e-12
FLOOR = 1e-9


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def close(a, b, tol=TOL, vectors=()):
    """Largest element-wise relative error, |a - b| / max(|b|, FLOOR * the
    column's largest magnitude); each group of columns in `vectors` (a
    position or a velocity) is measured as one vector, the error vector's
    length over the vector's length, with the same floor."""
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
        scale = np.maximum(np.maximum(mag, FLOOR * mag.max()), 1e-300)
        err = max(err, float((num / scale).max()))
    return err <= tol, err


def pclose(a, b):
    """close() for particle states: position and velocity as vectors."""
    return close(a, b, vectors=([0, 1, 2], [3, 4, 5]))


def pair(oracle_mod, jitter=0.0, nx=12, ny=4, nz=4, npart=600, seed=4, **pp):
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=nx, ny=ny, nz=nz, npart=npart, jitter=jitter,
                                                      seed=seed)
    picl.update(pp)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    return g, o, y, pc


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_particle_setup_matches_oracle(jitter, oracle_mod):
    g, o, y, pc = pair(oracle_mod, jitter=jitter)
    gc, oc = g.picl_cells(), o.picl_cells()
    assert same_bits(gc["phi"], oc["phi"])
    assert same_bits(gc["proFld"], oc["proFld"])
    assert np.array_equal(g.picl_particles()["elem"], o.picl_particles()["elem"])
    gcv, gdv = g.get_state()
    assert same_bits(gcv, o.cv()) and same_bits(gdv, o.dv())


def test_point_location_matches_oracle(oracle_mod):
    g, o, y, pc = pair(oracle_mod, jitter=0.15, nx=7, ny=5, nz=4, npart=10)
    rng = np.random.default_rng(9)
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([7, 5, 4]) * 0.0005
    pts = lo + (hi - lo) * rng.random((3000, 3)) * 1.02 - 0.01 * (hi - lo)
    assert np.array_equal(g.picl_locate(pts), o.picl_locate(pts))


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_first_stage_gas_side_is_bit_exact(jitter, oracle_mod):
    """From identical states: grad(rho_g), the 20 interpolation inputs and the
    interpolated particle values of the first coupled stage, bit for bit."""
    g, o, y, pc = pair(oracle_mod, jitter=jitter)
    g.picl_set_debug(True)
    g.RFLU_TimeStepInviscid()
    dt = g.RFLU_MinimumTimeStep()
    g.rk_stage(1)
    o.rk_stage(1, dt)
    gc, oc = g.picl_cells(), o.picl_cells()
    assert same_bits(gc["gradRhog"], oc["gradRhog"])
    assert same_bits(gc["fields"], oc["fields"])
    rp = o.picl_rprop()[:, np.array(rflu.PICL_FIELD_RPROP) - 1]
    assert same_bits(g.picl_interp(), rp)
    assert np.array_equal(g.picl_particles()["elem"], o.picl_particles()["elem"])


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_coupled_stage_after_shock_arrival_matches_oracle(jitter, oracle_mod):
    g, o, y, pc = pair(oracle_mod, jitter=jitter)
    g.RFLU_TimeStepping(3, use_graph=False)
    o.step(3)
    g.picl_set_debug(True)
    g.RFLU_TimeStepInviscid()
    dt = g.RFLU_MinimumTimeStep()
    g.rk_stage(1)
    o.rk_stage(1, dt)
    gc, oc = g.picl_cells(), o.picl_cells()
    for k in ("fields", "gradRhog", "proFld", "phi"):
        ok, err = close(gc[k], oc[k])
        assert ok, (k, err)
    rp = o.picl_rprop()[:, np.array(rflu.PICL_FIELD_RPROP) - 1]
    ok, err = close(g.picl_interp(), rp)
    assert ok, err
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["ydot"][:, 3]).max() > 0.0  # forces are acting
    for k in ("y", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


def test_coupled_steps_match_oracle(oracle_mod):
    g, o, y, pc = pair(oracle_mod, jitter=0.05, npart=800)
    n = 10
    g.RFLU_TimeStepping(n, use_graph=True)
    o.step(n)
    t, dt = g.time()
    assert abs(t - o.time) <= 1e-12 * o.time
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert g.picl_sizes() == (o.nPart, o.picl_nlost())
    for k in ("y", "ydot", "ydotc", "y1"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    assert np.abs(gp["y"][:, 3]).max() > 0.0  # the shock has reached the curtain
    gcv, gdv = g.get_state()
    ok, err = close(gcv, o.cv())
    assert ok, err
    ok, err = close(g.picl_cells()["phi"], o.picl_cells()["phi"])
    assert ok, err


def test_coupled_stage_by_routine_equals_fused_and_graph(oracle_mod):
    a, _, y, pc = pair(oracle_mod)
    b, _, _, _ = pair(oracle_mod)
    c, _, _, _ = pair(oracle_mod)
    for r in (a, b):
        r.RFLU_TimeStepInviscid()
        r.RFLU_MinimumTimeStep()
    for s in (1, 2, 3):
        a.RKInitMP(s)
        a.CellGradientsMP()
        a.ConvectiveFluxesMP()
        a.SourceTermsMP()
        a.PICL_TEMP_Runge(s)
        a.IntegrateSourceTermsMP()
        a.RKUpdateMP(s)
        b.rk_stage(s)
    c.RFLU_TimeStepping(1, use_graph=False)
    ca, cb, cc = a.get_state()[0], b.get_state()[0], c.get_state()[0]
    assert same_bits(ca, cb) and same_bits(ca, cc)
    for k in ("y", "ydotc", "elem"):
        assert np.array_equal(a.picl_particles()[k], b.picl_particles()[k])
    # graph replays against eager steps, with eager work between replays
    d, _, _, _ = pair(oracle_mod)
    e, _, _, _ = pair(oracle_mod)
    for _ in range(3):
        d.RFLU_TimeStepping(2, use_graph=True)
        e.RFLU_TimeStepping(2, use_graph=False)
        assert same_bits(d.get_state()[0], e.get_state()[0])
        assert same_bits(d.picl_particles()["y"], e.picl_particles()["y"])


# ---- optional models: drag laws, trilinear interpolation, Gaussian filter ----
@pytest.mark.parametrize("qs", [2, 3, 4, 5])
def test_coupled_steps_with_each_drag_law(oracle_mod, qs):
    g, o, y, pc = pair(oracle_mod, jitter=0.05, npart=500, qsFlag=qs)
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["y"][:, 3]).max() > 0.0
    for k in ("y", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (qs, k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, (qs, err)


def test_trilinear_interpolation_matches_oracle(oracle_mod):
    g, o, y, pc = pair(oracle_mod, jitter=0.1, interpMode=1)
    g.picl_set_debug(True)
    g.RFLU_TimeStepInviscid()
    dt = g.RFLU_MinimumTimeStep()
    g.rk_stage(1)
    o.rk_stage(1, dt)
    gc, oc = g.picl_cells(), o.picl_cells()
    assert same_bits(gc["fields"], oc["fields"])
    rp = o.picl_rprop()[:, np.array(rflu.PICL_FIELD_RPROP) - 1]
    assert same_bits(g.picl_interp(), rp)  # same Newton (r,s,t), same contraction
    g, o, y, pc = pair(oracle_mod, jitter=0.1, interpMode=1)
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    for k in ("y", "ydot"):
        ok, err = pclose(g.picl_particles()[k], o.picl_particles()[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


def test_gaussian_projection_matches_oracle(oracle_mod):
    """exp() and the vertex gather's summation order differ from the oracle's
    (ascending particle index over all particles), so the node values are
    held to the FP64 tolerance."""
    h = 0.0005
    g, o, y, pc = pair(oracle_mod, jitter=0.05, npart=400, projMode=1, filterWidth=2 * h, filterAlpha=1e-2)
    ok, err = close(g.picl_nodes().reshape(-1, 40), o.picl_nodes().reshape(-1, 40))
    assert ok, err
    ok, err = close(g.picl_cells()["phi"], o.picl_cells()["phi"])
    assert ok, err
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    ok, err = close(g.picl_nodes().reshape(-1, 40), o.picl_nodes().reshape(-1, 40))
    assert ok, err
    for k in ("y", "ydot", "ydotc"):
        ok, err = pclose(g.picl_particles()[k], o.picl_particles()[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


def test_gaussian_graph_replays_equal_eager_steps(oracle_mod):
    h = 0.0005
    kw = dict(jitter=0.05, npart=400, projMode=1, filterWidth=2 * h, filterAlpha=1e-2)
    d, _, _, _ = pair(oracle_mod, **kw)
    e, _, _, _ = pair(oracle_mod, **kw)
    for _ in range(3):
        d.RFLU_TimeStepping(2, use_graph=True)
        e.RFLU_TimeStepping(2, use_graph=False)
        assert same_bits(d.get_state()[0], e.get_state()[0])
        assert same_bits(d.picl_nodes(), e.picl_nodes())


@pytest.mark.parametrize("discr", [4, 5])
def test_coupled_steps_with_weno_reconstruction(oracle_mod, discr):
    """The gas side with WENO-XYZ reconstruction (the deck default) under the
    particle coupling, with AUSM+ and with AUSM+up (whose particle branch
    divides the face density by the gas fraction and scales the convected
    flux by the face's gas fraction)."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=500, jitter=0.05, seed=4,
                                                      reconst=2, discr=discr)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["y"][:, 3]).max() > 0.0
    for k in ("y", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


@pytest.mark.parametrize("ht", [1, 2, 3, 4])
def test_coupled_steps_with_heat_transfer(oracle_mod, ht):
    """HEATTRANSFER 1-4: the particle temperature evolves and the heat flux
    joins the energy feedback (a dilute curtain in the post-shock stream)."""
    n, h = 12, 0.0005
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=n, ny=4, nz=4, npart=500, jitter=0.05, seed=4, phi=0.01,
                                                      x_shock=0.9 * n * h, x_curtain=0.3 * n * h)
    y = y.copy()
    y[:, 6] = 250.0
    picl.update(heatTransfer=ht)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["y"][:, 6] - 250.0).max() > 0.0  # the temperature moved
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


@pytest.mark.parametrize("proj", [0, 1])
def test_periodic_particles_match_oracle(oracle_mod, proj):
    """PERIODICY and PERIODICZ: a quarter of the particles start within 1e-5
    of a y or z face moving outward at 300 m/s, leave during the run and
    re-enter at the opposite face; with the Gaussian filter their images
    across the faces deposit too."""
    n, h = 12, 0.0005
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=n, ny=4, nz=4, npart=400, jitter=0.05, seed=4)
    rng = np.random.default_rng(11)
    y = y.copy()
    y[:, 4] = rng.uniform(-300.0, 300.0, len(y))
    y[:, 5] = rng.uniform(-300.0, 300.0, len(y))
    edge = np.arange(len(y)) % 4 == 0
    for c in (4, 5):  # fast enough to cross within the run
        y[edge, c] = np.sign(y[edge, c]) * rng.uniform(200.0, 300.0, edge.sum())
    hi_y = edge & (y[:, 4] > 0)
    y[edge, 1] = np.where(y[edge, 4] > 0, 4 * h - 1e-5 * rng.random(edge.sum()), 1e-5 * rng.random(edge.sum()))
    y[edge, 2] = np.where(y[edge, 5] > 0, 4 * h - 1e-5 * rng.random(edge.sum()), 1e-5 * rng.random(edge.sum()))
    picl.update(periodicY=1, periodicZ=1, projMode=proj, filterWidth=2 * h, filterAlpha=1e-2)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(6, use_graph=True)
    o.step(6)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.all(gp["elem"] >= 0)
    assert np.all(gp["y"][hi_y, 1] < 2 * h)  # left through y = 4h, back in near y = 0
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    if proj == 1:
        ok, err = close(g.picl_nodes().reshape(-1, 40), o.picl_nodes().reshape(-1, 40))
        assert ok, err
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


@pytest.mark.parametrize("dilute", [True, False])
def test_coupled_steps_with_viscous_unsteady_force(oracle_mod, dilute):
    """ViscousUnsteady_flag 1: the history kernel over the last five steps'
    d(rho u)/dt of gas and particle, kept per particle and moved with it
    through the per-stage sort; eight steps fill the window and shift it.
    Graph replays on the GPU."""
    n, h = 12, 0.0005
    kw = dict(phi=0.01, x_shock=0.9 * n * h, x_curtain=0.3 * n * h) if dilute else {}
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=n, ny=4, nz=4, npart=500, jitter=0.05, seed=4, **kw)
    picl.update(viscousUnsteady=1)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(8, use_graph=True)
    o.step(8)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["ydot"][:, 3]).max() > 0.0
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


@pytest.mark.parametrize("periodic", [0, 1])
def test_coupled_steps_with_collisions(oracle_mod, periodic):
    """COLLISIONS 1: a curtain of randomly placed particles at phi 0.3, so
    that many overlap (about 2 contacts each) and feel the soft-sphere
    force; with periodic y and z also across those faces.  The GPU's hash
    search must find exactly the oracle's contacts and sum them in the same
    order."""
    n = 12
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=n, ny=4, nz=4, npart=2000, jitter=0.05, seed=4, phi=0.3)
    picl.update(collisionalFlag=1, ksp=10.0, erest=0.7, periodicY=periodic, periodicZ=periodic)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(4, use_graph=True)
    o.step(4)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["ydot"][:, 3]).max() > 1e6  # collision-sized accelerations
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


def test_coupled_steps_with_tangential_collisions(oracle_mod):
    """COLLISIONS 2: random spins and the dense curtain; the tangential
    force, the contact and rolling torques and the angular velocity they
    drive, with periodic y and z."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=2000, jitter=0.05, seed=4, phi=0.3)
    y = y.copy()
    y[:, 7:10] = np.random.default_rng(9).uniform(-3e4, 3e4, (len(y), 3))
    picl.update(collisionalFlag=2, ksp=10.0, erest=0.7, periodicY=1, periodicZ=1)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(4, use_graph=True)
    o.step(4)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(op["ydot"][:, 7:10]).max() > 0.0  # the spins changed
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


@pytest.mark.parametrize("level", [3, 4])
def test_coupled_steps_with_hydrodynamic_torque_and_lift(oracle_mod, level):
    """COLLISIONS 3 adds the hydrodynamic and undisturbed torques, from the
    gas vorticity and its material derivative (with the feedback force's
    LSQ/WENO/limited gradient); 4 adds the Saffman and Magnus lift.  Random
    spins, the dense curtain, periodic y and z, and a jittered mesh so the
    vorticity is not zero."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=2000, jitter=0.05, seed=4, phi=0.3)
    y = y.copy()
    y[:, 7:10] = np.random.default_rng(9).uniform(-3e4, 3e4, (len(y), 3))
    picl.update(collisionalFlag=level, ksp=10.0, erest=0.7, periodicY=1, periodicZ=1)
    g = rflu.Region(mesh, **params)
    o = oracle_mod.Region(mesh, **params)
    cv = ic(g.mesh_arrays()["cofg"])
    g.set_state(cv)
    o.set_cv(cv)
    g.picl_init(y, pc, **picl)
    o.picl_init(y, pc, **picl)
    g.RFLU_TimeStepping(4, use_graph=True)
    o.step(4)
    gp, op = g.picl_particles(), o.picl_particles()
    assert np.array_equal(gp["elem"], op["elem"])
    assert np.abs(o.picl_rprop()[:, 30:33]).max() > 0.0  # vorticity reaches the particles
    for k in ("y", "y1", "ydot", "ydotc"):
        ok, err = pclose(gp[k], op[k])
        assert ok, (k, err)
    ok, err = close(g.get_state()[0], o.cv())
    assert ok, err


def test_collision_list_overflow_is_reported():
    """Forty particles packed within a tenth of a diameter of one point each
    overlap 39 neighbours, more than the 32 a particle's contact list holds:
    the kernel finishes normally and the stage reports the overflow."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=40, seed=4)
    y = y.copy()
    rng = np.random.default_rng(2)
    y[:, :3] = y[0, :3] + 0.1 * pc[:, 1:2].min() * rng.uniform(-1.0, 1.0, (40, 3))
    picl.update(collisionalFlag=1, ksp=10.0, erest=0.7)
    g = rflu.Region(mesh, **params)
    g.set_state(ic(g.mesh_arrays()["cofg"]))
    g.picl_init(y, pc, **picl)
    with pytest.raises(Exception, match="32 neighbours"):
        g.RFLU_TimeStepping(1, use_graph=False)


@pytest.mark.parametrize("bad", [dict(collisionalFlag=5), dict(amFlag=3), dict(amFlag=2, neighborWidth=-1.0),
                                 dict(viscousUnsteady=2),
                                 dict(collisionalFlag=1, ksp=0.0), dict(collisionalFlag=1, erest=0.0),
                                 dict(collisionalFlag=1, erest=1.5), dict(periodicY=2), dict(heatTransfer=5)])
def test_unbuilt_or_invalid_particle_physics_is_rejected(bad):
    """Options outside their range fail at set-up with an error."""
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=8, ny=4, nz=4, npart=50, seed=4)
    picl.update(bad)
    g = rflu.Region(mesh, **params)
    g.set_state(ic(g.mesh_arrays()["cofg"]))
    with pytest.raises(Exception):
        g.picl_init(y, pc, **picl)
