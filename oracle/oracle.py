"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  Arrays are returned in the reference's Fortran layout
(AoS per cell: cv[N,5], dv[N,3], grad[N,5,3], fn[F,4], fc[F,3], f2c[F,2]).
"""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DISCR_ROE, DISCR_HLLC, DISCR_AUSM, DISCR_AUSMUP = 1, 3, 4, 5
RECONST_NONE, RECONST_VENKAT = 0, 11


class OrcParams(C.Structure):
    _fields_ = [("spaceDiscr", C.c_int), ("spaceOrder", C.c_int), ("reconst", C.c_int),
                ("patchOrder", C.c_int), ("cfl", C.c_double), ("dtImposed", C.c_double),
                ("dtMinLimit", C.c_double), ("dTimeSystem", C.c_double), ("refCp", C.c_double),
                ("refGamma", C.c_double), ("epsentr", C.c_double), ("dissFact", C.c_double),
                ("accelOn", C.c_int), ("accel", C.c_double * 3), ("limiterGuard", C.c_int),
                ("sigma", C.c_double), ("kp", C.c_double), ("ku", C.c_double), ("mInf", C.c_double)]


class OrcPiclParams(C.Structure):
    _fields_ = [("rmuFlag", C.c_int), ("rmuRef", C.c_double), ("tref", C.c_double), ("suth", C.c_double),
                ("rcpPart", C.c_double), ("qsFlag", C.c_int), ("amFlag", C.c_int), ("pgFlag", C.c_int),
                ("feedbackFlag", C.c_int), ("interpMode", C.c_int), ("projMode", C.c_int),
                ("filterWidth", C.c_double), ("filterAlpha", C.c_double), ("heatTransfer", C.c_int),
                ("periodicX", C.c_int), ("periodicY", C.c_int), ("periodicZ", C.c_int),
                ("viscousUnsteady", C.c_int), ("collisionalFlag", C.c_int), ("ksp", C.c_double),
                ("erest", C.c_double), ("neighborWidth", C.c_double), ("qsFluct", C.c_int),
                ("qsFluctAdapt", C.c_int), ("fluctHalfWidth", C.c_double), ("fluctSeed", C.c_longlong),
                ("angularPeriodic", C.c_int), ("angAngle", C.c_double), ("angXAngle", C.c_double),
                ("angRin", C.c_double), ("angRout", C.c_double)]


PICL_DEFAULTS = dict(rmuFlag=1, rmuRef=1.716e-5, tref=273.15, suth=110.4, rcpPart=840.0, qsFlag=1,
                     amFlag=1, pgFlag=1, feedbackFlag=1, interpMode=0, projMode=0, filterWidth=0.0,
                     filterAlpha=1e-2)
# rprop index (1-based, PPICLF_USER.h) of each interpolated field slot
PICL_FIELD_RPROP = [2, 6, 7, 8, 10, 11, 12, 9, 24, 5, 13, 14, 15, 16, 17, 18, 19, 37, 38, 39]


def make_picl_params(**kw):
    d = dict(PICL_DEFAULTS)
    d.update(kw)
    p = OrcPiclParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_ref():
    """Compile the reference's Briney added-mass helpers into oracle/_ref
    (make ref) where the reference tree and amdflang are at hand; elsewhere
    nothing is built and the golden vectors stand in for them."""
    ref = os.environ.get("REF", "/root/reference")
    if os.path.isdir(os.path.join(ref, "libpicl", "user_files", "Added_Mass")) and shutil.which("amdflang"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", "REF=" + ref], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, P, I, I, P, I, P, P, P, P, I, P, C.POINTER(OrcParams)]
        L.orc_last_error.restype = C.c_char_p
        L.orc_destroy.argtypes = [P]
        for n in ("orc_nfaces", "orc_nbfaces", "orc_stencil_total"):
            getattr(L, n).argtypes = [P]
            getattr(L, n).restype = I
        L.orc_get_faces.argtypes = [P, P, P, P, P]
        L.orc_get_bfaces.argtypes = [P, P, P, P]
        L.orc_get_cells.argtypes = [P, P, P]
        L.orc_get_stencils.argtypes = [P, P, P, P]
        for n in ("orc_set_cv", "orc_get_cv", "orc_get_dv", "orc_get_rhs", "orc_get_grad"):
            getattr(L, n).argtypes = [P, P]
        L.orc_timestep_inviscid.argtypes = [P, P]
        L.orc_timestep_inviscid.restype = D
        L.orc_residual.argtypes = [P]
        L.orc_step.argtypes = [P, I]
        L.orc_step.restype = I
        L.orc_time.argtypes = [P]
        L.orc_time.restype = D
        L.orc_last_dt.argtypes = [P]
        L.orc_last_dt.restype = D
        L.orc_flux_ausm.argtypes = [P, D, P, P, P]
        L.orc_rk_stage.argtypes = [P, I]
        L.orc_rk_stage.restype = I
        L.orc_set_dtmin.argtypes = [P, D]
        L.orc_set_patch_vals.argtypes = [P, I, P]
        L.orc_flux_point.argtypes = [C.POINTER(OrcParams), I, P, P, P, P]
        L.orc_picl_init.argtypes = [P, C.POINTER(OrcPiclParams), I, P, P]
        L.orc_picl_init.restype = I
        for n in ("orc_picl_npart", "orc_picl_nlost"):
            getattr(L, n).argtypes = [P]
            getattr(L, n).restype = I
        L.orc_picl_get_particles.argtypes = [P, P, P, P, P, P]
        L.orc_picl_get_rprop.argtypes = [P, P]
        L.orc_picl_get_cells.argtypes = [P, P, P, P, P]
        L.orc_picl_locate.argtypes = [P, I, P, P]
        L.orc_picl_neighbors.argtypes = [P, P, P, P]
        L.orc_picl_neighbors.restype = I
        L.orc_picl_fluct_means.argtypes = [P, I, P]
        L.orc_picl_get_nodes.argtypes = [P, P]
        L.orc_picl_qs_beta.argtypes = [I] + [C.c_double] * 7
        L.orc_picl_qs_beta.restype = C.c_double
        L.orc_picl_vu_force.argtypes = [I, P, P, P] + [C.c_double] * 6 + [P]
        L.orc_picl_vu_state.argtypes = [P, P, P]
        L.orc_picl_vu_state.restype = I
        D = C.c_double
        L.orc_picl_collision_force.argtypes = [P, P, D, D, D, P, P, D, D, D, D, D, D, P]
        L.orc_picl_collision_force2.argtypes = [P, P, D, D, D, D, D, D, D, D, D, P, P]
        for n in ("orc_am_b11_11", "orc_am_b11_22", "orc_am_b12_11", "orc_am_b12_22", "orc_am_rdf"):
            getattr(L, n).argtypes = [D, D]
            getattr(L, n).restype = D
        for n in ("orc_am_ia_analytical", "orc_am_ii_analytical", "orc_am_ia_numerical", "orc_am_ii_numerical"):
            getattr(L, n).argtypes = [D, D, D]
            getattr(L, n).restype = D
        L.orc_am_resistance_pair.argtypes = [D, D, D, D, D, P]
        L.orc_picl_torque_hydro.argtypes = [P, P, D, D, D, D, D, P]
        L.orc_picl_torque_hydro.restype = I
        L.orc_picl_lift.argtypes = [P, P, P, D, D, D, D, P]
        L.orc_fluct_uniform.argtypes = [C.c_longlong, I, I, I]
        L.orc_fluct_uniform.restype = D
        L.orc_fluct_force.argtypes = [I, P, P, P] + [D] * 4 + [P, D, D, D, P]
        L.orc_set_threads.argtypes = [I]
        L.orc_get_threads.restype = I
        L.orc_test_rk3_linear.argtypes = [D, D, P]
        L.orc_set_threads(int(os.environ.get("ROCFLUPICL_ORACLE_THREADS", "1")))
        L.orc_test_entropy_fix.argtypes = [D, D]
        L.orc_test_entropy_fix.restype = D
        L.orc_test_venkat_term.argtypes = [D, D, D, I]
        L.orc_test_venkat_term.restype = D
        _LIB = L
    return _LIB


def rk3_linear(lam, y0=1.0):
    """The oracle's Wray RK3 update on y' = lam y with dt = 1: the values
    after stages 1, 2 and 3."""
    y = np.zeros(3)
    lib().orc_test_rk3_linear(float(lam), float(y0), _p(y))
    return y


def entropy_fix(l, d):
    """The oracle's Harten-Hyman entropy fix of eigenvalue magnitude l with
    threshold d."""
    return lib().orc_test_entropy_fix(float(l), float(d))


def venkat_term(dm, d2, epsq, guard=0):
    """The oracle's Venkatakrishnan term of one face: dm = neighbour extreme
    minus the cell value, d2 = extrapolation to the face, epsq = eps^2."""
    return lib().orc_test_venkat_term(float(dm), float(d2), float(epsq), int(guard))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_params(**kw):
    p = OrcParams()
    d = dict(spaceDiscr=DISCR_AUSM, spaceOrder=1, reconst=RECONST_NONE, patchOrder=1, cfl=0.5,
             dtImposed=1e-5, dtMinLimit=1e-10, dTimeSystem=0.0, refCp=1004.5, refGamma=1.4,
             epsentr=0.05, dissFact=1.0, accelOn=0, accel=(0.0, 0.0, 0.0), limiterGuard=0,
             sigma=1.0, kp=0.25, ku=0.75, mInf=0.1)
    d.update(kw)
    for k, v in d.items():
        if k == "accel":
            for i in range(3):
                p.accel[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class Region:
    """Oracle region built from a meshgen mesh dict."""

    def __init__(self, mesh, **params):
        L = lib()
        self._keep = []
        xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
        c2v = np.ascontiguousarray(mesh["c2v"], dtype=np.int32)
        pts = mesh["patches"]
        nvf = pts[0]["bf2v"].shape[1]
        bt = np.array([p["bcType"] for p in pts], dtype=np.int32)
        bo = np.array([p["bcOpt"] for p in pts], dtype=np.int32)
        po = np.array([p["pout"] for p in pts], dtype=np.float64)
        nb = np.array([p["bf2v"].shape[0] for p in pts], dtype=np.int32)
        bf = np.ascontiguousarray(np.concatenate([p["bf2v"] for p in pts]), dtype=np.int32)
        self.params = make_params(**params)
        self.h = L.orc_create(xyz.shape[0], _p(xyz), c2v.shape[1], c2v.shape[0], _p(c2v), len(pts),
                              _p(bt), _p(bo), _p(po), _p(nb), nvf, _p(bf), C.byref(self.params))
        if not self.h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        for ip, p in enumerate(pts):
            if "vals" in p:
                v = np.zeros(6)
                v[:len(p["vals"])] = p["vals"]
                L.orc_set_patch_vals(self.h, ip, _p(v))
        self.nCells = c2v.shape[0]
        self.nFaces = L.orc_nfaces(self.h)
        self.nBFaces = L.orc_nbfaces(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def faces(self):
        F = self.nFaces
        f2c = np.zeros((F, 2), np.int32)
        f2v = np.zeros((F, 4), np.int32)
        fn = np.zeros((F, 4))
        fc = np.zeros((F, 3))
        lib().orc_get_faces(self.h, _p(f2c), _p(f2v), _p(fn), _p(fc))
        return f2c, f2v, fn, fc

    def bfaces(self):
        B = self.nBFaces
        bf2c = np.zeros(B, np.int32)
        bfn = np.zeros((B, 4))
        bfc = np.zeros((B, 3))
        lib().orc_get_bfaces(self.h, _p(bf2c), _p(bfn), _p(bfc))
        return bf2c, bfn, bfc

    def cells(self):
        vol = np.zeros(self.nCells)
        cofg = np.zeros((self.nCells, 3))
        lib().orc_get_cells(self.h, _p(vol), _p(cofg))
        return vol, cofg

    def stencils(self):
        tot = lib().orc_stencil_total(self.h)
        off = np.zeros(self.nCells + 1, np.int32)
        mem = np.zeros(max(tot, 1), np.int32)
        moms = np.zeros((self.nCells, 10))
        lib().orc_get_stencils(self.h, _p(off), _p(mem), _p(moms))
        return off, mem[:tot], moms

    def set_cv(self, cv):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        assert cv.shape == (self.nCells, 5)
        lib().orc_set_cv(self.h, _p(cv))

    def cv(self):
        a = np.zeros((self.nCells, 5))
        lib().orc_get_cv(self.h, _p(a))
        return a

    def dv(self):
        a = np.zeros((self.nCells, 3))
        lib().orc_get_dv(self.h, _p(a))
        return a

    def rhs(self):
        a = np.zeros((self.nCells, 5))
        lib().orc_get_rhs(self.h, _p(a))
        return a

    def grad(self):
        a = np.zeros((self.nCells, 5, 3))
        lib().orc_get_grad(self.h, _p(a))
        return a

    def timestep(self):
        dt = np.zeros(self.nCells)
        m = lib().orc_timestep_inviscid(self.h, _p(dt))
        return m, dt

    def residual(self):
        lib().orc_residual(self.h)
        return self.rhs()

    def step(self, n=1):
        e = lib().orc_step(self.h, n)
        if e:
            raise RuntimeError("oracle step failed: " + lib().orc_last_error().decode())

    def rk_stage(self, istage, dt):
        lib().orc_set_dtmin(self.h, dt)
        return lib().orc_rk_stage(self.h, istage)

    # ---- ppiclF coupling ------------------------------------------------
    def picl_init(self, y, pconst, **picl_params):
        """PICL_TEMP_InitSolver: y (n,12) particle state, pconst (n,4) =
        (rho_p, d_p, V_p, SPL).  cv must hold the pure-gas state."""
        y = np.ascontiguousarray(y, np.float64)
        pc = np.ascontiguousarray(pconst, np.float64)
        self._picl_params = make_picl_params(**picl_params)
        if lib().orc_picl_init(self.h, C.byref(self._picl_params), y.shape[0], _p(y), _p(pc)):
            raise RuntimeError("oracle: PICL needs an all-hex second-order region")
        self.nPart = y.shape[0]

    def picl_particles(self):
        n = self.nPart
        y, y1, yd, yc = (np.zeros((n, 12)) for _ in range(4))
        elem = np.zeros(n, np.int32)
        lib().orc_picl_get_particles(self.h, _p(y), _p(y1), _p(yd), _p(yc), _p(elem))
        return dict(y=y, y1=y1, ydot=yd, ydotc=yc, elem=elem)

    def picl_vu_state(self):
        """Viscous-unsteady bookkeeping: (nTimeBH, timeBH[5], mixt[n, 3, 5],
        plag[n, 3, 5])."""
        t = np.zeros(5)
        h = np.zeros((self.nPart, 2, 3, 5))
        n = lib().orc_picl_vu_state(self.h, _p(t), _p(h))
        return n, t, h[:, 0], h[:, 1]

    def picl_rprop(self):
        a = np.zeros((self.nPart, 48))
        lib().orc_picl_get_rprop(self.h, _p(a))
        return a

    def picl_cells(self):
        N = self.nCells
        phi, g, pf, fld = np.zeros(N), np.zeros((N, 3)), np.zeros((N, 5)), np.zeros((N, 20))
        lib().orc_picl_get_cells(self.h, _p(phi), _p(g), _p(pf), _p(fld))
        return dict(phi=phi, gradRhog=g, proFld=pf, fields=fld)

    def picl_nodes(self):
        """Projected values at the 8 nodes of every element, (nCells, 5, 8)."""
        a = np.zeros((self.nCells, 5, 8))
        lib().orc_picl_get_nodes(self.h, _p(a))
        return a

    def picl_nlost(self):
        return lib().orc_picl_nlost(self.h)

    def picl_locate(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros(xyz.shape[0], np.int32)
        lib().orc_picl_locate(self.h, xyz.shape[0], _p(xyz), _p(out))
        return out

    def picl_fluct_means(self, i):
        """Particle i's fluctuation box sums (u, v, w, u^2, v^2, w^2, count)."""
        out = np.zeros(7)
        lib().orc_picl_fluct_means(self.h, int(i), _p(out))
        return out

    def picl_neighbors(self, xp):
        xp = np.ascontiguousarray(xp, np.float64)
        ids = np.zeros(27, np.int32)
        d2 = np.zeros(27)
        n = lib().orc_picl_neighbors(self.h, _p(xp), _p(ids), _p(d2))
        return ids[:n], d2[:n]

    @property
    def time(self):
        return lib().orc_time(self.h)

    @property
    def last_dt(self):
        return lib().orc_last_dt(self.h)


def flux_ausm(n, fs, L, R):
    out = np.zeros(5)
    n = np.ascontiguousarray(n, np.float64)
    L = np.ascontiguousarray(L, np.float64)
    R = np.ascontiguousarray(R, np.float64)
    lib().orc_flux_ausm(_p(n), fs, _p(L), _p(R), _p(out))
    return out


def flux_point(discr, n, L, R, **params):
    """Face flux F(L, R) through n = (nx, ny, nz, |S|); L, R = (rho, u, v, w, p)."""
    out = np.zeros(5)
    prm = make_params(**params)
    n = np.ascontiguousarray(n, np.float64)
    L = np.ascontiguousarray(L, np.float64)
    R = np.ascontiguousarray(R, np.float64)
    lib().orc_flux_point(C.byref(prm), discr, _p(n), _p(L), _p(R), _p(out))
    return out


def vu_force(nTimeBH, timeBH, mixt, plag, dt, rmu, rhof, dp, vmag, rphip):
    """Viscous-unsteady force of one particle from its history: timeBH (5),
    mixt and plag (3 x 5, entry k of component c at [c, k])."""
    t = np.ascontiguousarray(timeBH, dtype=np.float64)
    m = np.ascontiguousarray(mixt, dtype=np.float64)
    q = np.ascontiguousarray(plag, dtype=np.float64)
    assert t.shape == (5,) and m.shape == (3, 5) and q.shape == (3, 5)
    out = np.zeros(3)
    lib().orc_picl_vu_force(int(nTimeBH), _p(t), _p(m), _p(q), float(dt), float(rmu), float(rhof), float(dp),
                            float(vmag), float(rphip), _p(out))
    return out


def collision_force(xi, vi, dpi, xj, vj, dpj, dt, ksp, erest, rhop=2500.0):
    """Collision force of particle j on particle i (COLLISIONS 1); spheres
    of density rhop, volume pi d^3 / 6."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xi, vi, xj, vj)]
    out = np.zeros(3)
    vol = lambda d: math.pi / 6.0 * d ** 3
    lib().orc_picl_collision_force(_p(a[0]), _p(a[1]), float(dpi), rhop, vol(dpi), _p(a[2]), _p(a[3]), float(dpj),
                                   rhop, vol(dpj), float(dt), float(ksp), float(erest), _p(out))
    return out


def collision_force2(xi, vi, wi, dpi, xj, vj, wj, dpj, dt, ksp, erest, rhop=2500.0):
    """COLLISIONS 2: (force, torque) of particle j on particle i, angular
    velocities wi, wj; spheres of density rhop."""
    yi, yj = np.zeros(12), np.zeros(12)
    yi[:3], yi[3:6], yi[7:10] = xi, vi, wi
    yj[:3], yj[3:6], yj[7:10] = xj, vj, wj
    f, t = np.zeros(3), np.zeros(3)
    vol = lambda d: math.pi / 6.0 * d ** 3
    lib().orc_picl_collision_force2(_p(yi), _p(yj), float(dpi), rhop, vol(dpi), float(dpj), rhop, vol(dpj),
                                    float(dt), float(ksp), float(erest), _p(f), _p(t))
    return f, t


def qs_beta(qsFlag, rmachp, rphip, rep, rmu, dp, rhof=1.0, vmag=1.0):
    """Quasi-steady coefficient beta of model qsFlag (1 Parmar, 2 Osnes,
    3 modified Parmar, 4 Gidaspow, 5 Schiller-Naumann) from SetYdot's
    capped inputs."""
    return lib().orc_picl_qs_beta(int(qsFlag), float(rmachp), float(rphip), float(rep), float(rmu), float(dp),
                                  float(rhof), float(vmag))


def am_curve(name, *args):
    """One function of Briney's added-mass model by its reference name:
    b11_11, b11_22, b12_11, b12_22, rdf (d, alpha); ia_analytical,
    ii_analytical, ia_numerical, ii_numerical (rmax, rad, alpha)."""
    return getattr(lib(), "orc_am_" + name)(*[float(a) for a in args])


def am_resistance_pair(x, y, z, alpha, rad):
    """resistance_pair: the pair's (3, 6) block [B11 | B12] rotated onto the
    pair axis (x, y, z) = x_j - x_i."""
    R = np.zeros((3, 6))
    lib().orc_am_resistance_pair(float(x), float(y), float(z), float(alpha), float(rad), _p(R))
    return R


def torque_hydro(vort, spin, rhof, dp, rmu, rmass, rhop):
    """Torque_Hydro: the torque vector, or None where ppiclF stops (Re_r >= 100)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (vort, spin)]
    t = np.zeros(3)
    rc = lib().orc_picl_torque_hydro(_p(a[0]), _p(a[1]), float(rhof), float(dp), float(rmu), float(rmass),
                                     float(rhop), _p(t))
    return None if rc else t


def lift(vort, spin, vrel, dp, rep, rmu, rhof):
    """Saffman + Magnus lift on a particle (rep capped at 0.1 beforehand)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (vort, spin, vrel)]
    out = np.zeros(3)
    lib().orc_picl_lift(_p(a[0]), _p(a[1]), _p(a[2]), float(dp), float(rep), float(rmu), float(rhof), _p(out))
    return out


def fluct_uniform(seed, call, pid, k):
    """The uniform draw k of particle pid's SetYdot call (stands in for RANDOM_NUMBER)."""
    return lib().orc_fluct_uniform(int(seed), int(call), int(pid), int(k))


def fluct_force(flag, means, F, U, rphip, rmachp, rep, vmag, vrel, rmu, dp, fac):
    """One SDE step of the fluctuating force (1 Lattanzi, 2 Osnes)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (means, F, U, vrel)]
    out = np.zeros(3)
    lib().orc_fluct_force(int(flag), _p(a[0]), _p(a[1]), _p(a[2]), float(rphip), float(rmachp), float(rep),
                          float(vmag), _p(a[3]), float(rmu), float(dp), float(fac), _p(out))
    return out


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_get_threads()
