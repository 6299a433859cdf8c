"""The workloads of the Rocflu hot path, built from seeds.

* ``sod``: 3-D Sod shock tube, 200x4x4 hexes (h = 0.005) stored as an
  unstructured grid, nodes jittered +-10% of h (interior freely, boundary
  nodes within their plane), perfect gas gamma = 1.4 with Cp = 3.5 (R = 1),
  rho/p = 1/1 left and 0.125/0.1 right of x = 0.5, first-order AUSM+,
  Wray RK3, CFL 0.5; x-ends supersonic outflow, sides slip walls.
* ``shock_tets``: [0,1]^3 box of n^3 hexes split into 5 tets each
  (64^3 -> 1,310,720 tets), nodes jittered +-5%, quiescent air rho = 1.2,
  p = 1e5, gamma = 1.4, Cp = 1004.64, with a Mach 1.5 shock at x = 0.25
  moving +x (Rankine-Hugoniot post-shock state behind it); second order with
  LSQ gradients + Venkatakrishnan limiter; Roe (ENTROPY 0.05) or HLLC.
* ``curtain``: 400x80x80 hexes over 0.2 x 0.04 x 0.04 m, Mach 2 shock at
  x = 0.02; the particle curtain lives in picl_configs.

Each builder takes a scale argument so tests can run the same workload at
sizes the CPU oracle finishes in seconds, and ``bc`` (one
(bcType, bcOpt, pout) per box side x-, x+, y-, y+, z-, z+) to replace the
default outflow ends and slip-wall sides, e.g. by a closed box.
"""
import numpy as np

from . import meshgen as mg

AIR = dict(refCp=1004.64, refGamma=1.4)


def sod(nx=200, ny=4, nz=4, jitter=0.1, seed=1, discr=4, order=1, **kw):
    h = 0.005
    mesh = mg.hex_box(nx, ny, nz, nx * h, ny * h, nz * h, jitter=jitter, seed=seed)
    params = dict(spaceDiscr=discr, spaceOrder=order, refCp=3.5, refGamma=1.4, cfl=0.5,
                  dtImposed=3.0e-4, dtMinLimit=0.0)
    if order == 2:
        params.update(reconst=11)
    params.update(kw)
    x0 = 0.5 * nx * h

    def ic(cofg):
        x = cofg[:, 0]
        rho = np.where(x < x0, 1.0, 0.125)
        p = np.where(x < x0, 1.0, 0.1)
        z = np.zeros_like(rho)
        return mg.prim_to_cons(rho, z, z, z, p, 1.4).T.copy()

    return mesh, params, ic


def shock_tets(n=64, jitter=0.05, seed=2, discr=1, mach=1.5, order=2, gamma=1.4, bc=None, **kw):
    mesh = mg.tet_box(n, n, n, 1.0, 1.0, 1.0, jitter=jitter, seed=seed, bc=bc)
    params = dict(spaceDiscr=discr, spaceOrder=order, reconst=11 if order == 2 else 0,
                  cfl=0.5, dtImposed=1.0e-7, dtMinLimit=0.0, epsentr=0.05, **AIR)
    params.update(refGamma=gamma)
    params.update(kw)
    rho1, p1, g = 1.2, 1.0e5, gamma
    rho2, u2, p2 = mg.normal_shock_state(mach, rho1, p1, g)

    def ic(cofg):
        x = cofg[:, 0]
        post = x < 0.25
        rho = np.where(post, rho2, rho1)
        u = np.where(post, u2, 0.0)
        p = np.where(post, p2, p1)
        z = np.zeros_like(rho)
        return mg.prim_to_cons(rho, u, z, z, p, g).T.copy()

    return mesh, params, ic


def shock_hex(nx=400, ny=80, nz=80, lx=0.2, ly=0.04, lz=0.04, mach=2.0, x_shock=0.02,
              discr=4, order=2, jitter=0.0, seed=3, gamma=1.4, **kw):
    mesh = mg.hex_box(nx, ny, nz, lx, ly, lz, jitter=jitter, seed=seed)
    params = dict(spaceDiscr=discr, spaceOrder=order, reconst=11 if order == 2 else 0,
                  cfl=0.5, dtImposed=1.0e-8, dtMinLimit=0.0, **AIR)
    params.update(refGamma=gamma)
    params.update(kw)
    rho1, p1, g = 1.2, 1.0e5, gamma
    rho2, u2, p2 = mg.normal_shock_state(mach, rho1, p1, g)

    def ic(cofg):
        x = cofg[:, 0]
        post = x < x_shock
        rho = np.where(post, rho2, rho1)
        u = np.where(post, u2, 0.0)
        p = np.where(post, p2, p1)
        z = np.zeros_like(rho)
        return mg.prim_to_cons(rho, u, z, z, p, g).T.copy()

    return mesh, params, ic


PICL_AIR = dict(rmuFlag=1, rmuRef=1.716e-5, tref=273.15, suth=110.4, rcpPart=840.0, qsFlag=1, amFlag=1,
                pgFlag=1, feedbackFlag=1)

# The particle physics of the AirShock input deck
# (input_files/AirShock/rectshktb.inp:97-116): Osnes quasi-steady drag, Briney
# added mass, pressure gradient, collisions level 4 (normal and tangential
# forces, contact, rolling, hydrodynamic and undisturbed torques, Saffman and
# Magnus lift) with KSP 10 and EREST 0.7, the viscous-unsteady force, Osnes'
# quasi-steady fluctuations with the fixed box filter, periodic in y and z.
AIRSHOCK_PICL = dict(qsFlag=2, amFlag=2, pgFlag=1, collisionalFlag=4, ksp=10.0, erest=0.7, viscousUnsteady=1,
                     heatTransfer=0, feedbackFlag=1, qsFluct=2, qsFluctAdapt=0, periodicY=1, periodicZ=1)


def curtain(nx=400, ny=80, nz=80, h=0.0005, npart=4_000_000, mach=2.0, x_shock=None, x_curtain=None,
            phi=0.2, d_median=115e-6, sigma_ln=0.2, rho_p=2500.0, discr=4, jitter=0.0, seed=4,
            nparts_cap=None, bc=None, placement="random", **kw):
    """Shock-particle-curtain: hex box nx*ny*nz of spacing h, Mach `mach`
    shock at x_shock moving +x into quiescent air, and a slab of `npart`
    particles (lognormal diameters) whose thickness gives the volume fraction
    `phi` across the box section.  placement "random": uniform positions
    (particles may overlap); "lattice": a jittered lattice whose spacing
    exceeds the largest diameter, so no two particles touch at the start (a
    particle file's cloud), fitted to the section so periodic images keep
    the spacing.

    Returns (mesh, params, ic, particles, picl_params) where particles =
    (y[n,12], pconst[n,4] = (rho_p, d_p, V_p, SPL)).  ic(cofg) gives the
    pure-gas state; the PICL set-up scales it by (1 - phi).
    """
    mesh = mg.hex_box(nx, ny, nz, nx * h, ny * h, nz * h, jitter=jitter, seed=seed, bc=bc)
    params = dict(spaceDiscr=discr, spaceOrder=2, reconst=11, cfl=0.5, dtImposed=1.0e-8,
                  dtMinLimit=0.0, **AIR)
    params.update(kw)
    rho1, p1, g = 1.2, 1.0e5, 1.4
    rho2, u2, p2 = mg.normal_shock_state(mach, rho1, p1, g)
    xs = 0.1 * nx * h if x_shock is None else x_shock
    x0 = xs + 2 * h if x_curtain is None else x_curtain

    def ic(cofg):
        x = cofg[:, 0]
        post = x < xs
        rho = np.where(post, rho2, rho1)
        u = np.where(post, u2, 0.0)
        p = np.where(post, p2, p1)
        z = np.zeros_like(rho)
        return mg.prim_to_cons(rho, u, z, z, p, g).T.copy()

    rng = np.random.default_rng(seed + 1000)
    dp = d_median * np.exp(sigma_ln * rng.standard_normal(npart))
    vp = np.pi / 6.0 * dp ** 3
    area = (ny * h) * (nz * h)
    thick = vp.sum() / (phi * area)
    y = np.zeros((npart, 12))
    if placement == "lattice":
        a = (vp.mean() / phi) ** (1.0 / 3.0)
        nl = [max(1, int(ny * h // a)), max(1, int(nz * h // a))]
        per_layer = nl[0] * nl[1]
        nxl = -(-npart // per_layer)
        sp = np.array([thick / nxl, ny * h / nl[0], nz * h / nl[1]])
        if sp.min() <= dp.max():
            raise ValueError("the lattice spacing does not exceed the largest diameter")
        k = np.arange(npart)
        idx = np.stack([k // per_layer, (k % per_layer) // nl[1], k % nl[1]], axis=1)
        wob = 0.45 * (sp - dp.max()) * (2.0 * rng.random((npart, 3)) - 1.0)
        xyz = (idx + 0.5) * sp + wob
        xyz[:, 0] += x0
        y[:, :3] = xyz[rng.permutation(npart)]
    elif placement == "random":
        y[:, 0] = x0 + thick * rng.random(npart)
        y[:, 1] = ny * h * rng.random(npart)
        y[:, 2] = nz * h * rng.random(npart)
    else:
        raise ValueError(f"unknown placement {placement!r}")
    y[:, 6] = 300.0
    pconst = np.stack([np.full(npart, rho_p), dp, vp, np.ones(npart)], axis=1)
    return mesh, params, ic, (y, pconst), dict(PICL_AIR)


def wedge(nr=6, nt=8, nz=4, rin=0.02, rout=0.05, angle_deg=90.0, xangle_deg=0.0, lz=0.01, npart=240,
          swirl=200.0, near=5e-7, seed=8, periodicZ=0, discr=4, **kw):
    """Particles in a wedge about the z axis (ANGULARPERIODIC): nr x nt x nz
    hexes over radii rin..rout, polar angles xangle..xangle + angle (given in
    degrees, converted as PICL_TEMP_InitSolver.F90:271-273 does), z in [0, lz],
    slip walls, quiescent air.  Half the particles sit within `near` of the
    lower or the upper wedge face and swirl across it at `swirl` m/s (so the
    first stage carries them out), the others anywhere inside, some moving
    in z.  Returns (mesh, params, ic, particles, picl_params) like curtain."""
    pi = float(np.arccos(-1.0))
    angle, xangle = angle_deg * pi / 180.0, xangle_deg * pi / 180.0
    mesh = mg.wedge_hex(nr, nt, nz, rin, rout, angle, xangle, lz)
    params = dict(spaceDiscr=discr, spaceOrder=2, reconst=11, cfl=0.5, dtImposed=1.0e-8,
                  dtMinLimit=0.0, limiterGuard=1, **AIR)
    params.update(kw)

    def ic(cofg):
        rho = np.full(len(cofg), 1.2)
        z = np.zeros_like(rho)
        return mg.prim_to_cons(rho, z, z, z, np.full(len(cofg), 1.0e5), 1.4).T.copy()

    rng = np.random.default_rng(seed)
    r = rin + (rout - rin) * (0.1 + 0.8 * rng.random(npart))
    side = rng.integers(0, 3, npart)  # 0 lower face, 1 upper face, 2 inside
    dth = near / r
    th = np.where(side == 0, xangle + dth, np.where(side == 1, xangle + angle - dth,
                                                  xangle + angle * (0.1 + 0.8 * rng.random(npart))))
    vt = np.where(side == 0, -swirl, np.where(side == 1, swirl, swirl * (2.0 * rng.random(npart) - 1.0)))
    y = np.zeros((npart, 12))
    y[:, 0] = r * np.cos(th)
    y[:, 1] = r * np.sin(th)
    y[:, 2] = lz * (0.05 + 0.9 * rng.random(npart))
    near_z = rng.random(npart) < 0.25
    y[near_z, 2] = np.where(rng.random(near_z.sum()) < 0.5, near, lz - near)
    y[:, 3] = -vt * np.sin(th)
    y[:, 4] = vt * np.cos(th)
    y[:, 5] = np.where(near_z, np.where(y[:, 2] < 0.5 * lz, -swirl, swirl), 0.0)
    y[:, 6] = 300.0
    dp = 60e-6 * np.exp(0.2 * rng.standard_normal(npart))
    vp = np.pi / 6.0 * dp ** 3
    pconst = np.stack([np.full(npart, 2500.0), dp, vp, np.ones(npart)], axis=1)
    picl = dict(PICL_AIR, angularPeriodic=1, angAngle=angle, angXAngle=xangle, angRin=rin, angRout=rout,
                periodicZ=periodicZ)
    return mesh, params, ic, (y, pconst), picl


def ensemble(n_inst=256, cells=(5_000, 500_000), mach=(1.2, 3.0), gammas=(1.4, 1.67), seed=7):
    """The batched ensemble: n_inst independent shock tubes with cell counts
    log-uniform in `cells`, half 5-tet boxes ([0,1]^3) and half 5:1:1 hex
    boxes, shock Mach number uniform in `mach`, gamma drawn from `gammas`;
    second-order AUSM+ with the sign-safe Venkatakrishnan guard.  Returns a
    list of (key, mesh, params, ic) with key = (cell type, gamma): instances
    that share a key can share one batch."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        nc = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        m = float(rng.uniform(*mach))
        g = float(gammas[rng.integers(len(gammas))])
        sd = int(rng.integers(1 << 30))
        if i % 2 == 0:
            k = max(2, int(round((nc / 5.0) ** (1.0 / 3.0))))
            mesh, params, ic = shock_tets(n=k, discr=4, mach=m, seed=sd, gamma=g, limiterGuard=1)
            out.append((("tet", g), mesh, params, ic))
        else:
            k = max(2, int(round((nc / 5.0) ** (1.0 / 3.0))))
            mesh, params, ic = shock_hex(nx=5 * k, ny=k, nz=k, lx=0.2, ly=0.04, lz=0.04, mach=m,
                                         x_shock=0.02, jitter=0.05, seed=sd, gamma=g, limiterGuard=1)
            out.append((("hex", g), mesh, params, ic))
    return out
