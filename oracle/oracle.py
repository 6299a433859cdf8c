"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  Arrays are returned in the reference's Fortran layout
(AoS per cell: cv[N,5], dv[N,3], grad[N,5,3], fn[F,4], fc[F,3], f2c[F,2]).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DISCR_ROE, DISCR_HLLC, DISCR_AUSM = 1, 3, 4
RECONST_NONE, RECONST_VENKAT = 0, 11


class OrcParams(C.Structure):
    _fields_ = [("spaceDiscr", C.c_int), ("spaceOrder", C.c_int), ("reconst", C.c_int),
                ("patchOrder", C.c_int), ("cfl", C.c_double), ("dtImposed", C.c_double),
                ("dtMinLimit", C.c_double), ("dTimeSystem", C.c_double), ("refCp", C.c_double),
                ("refGamma", C.c_double), ("epsentr", C.c_double), ("dissFact", C.c_double),
                ("accelOn", C.c_int), ("accel", C.c_double * 3), ("limiterGuard", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


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
        L.orc_flux_point.argtypes = [C.POINTER(OrcParams), I, P, P, P, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_params(**kw):
    p = OrcParams()
    d = dict(spaceDiscr=DISCR_AUSM, spaceOrder=1, reconst=RECONST_NONE, patchOrder=1, cfl=0.5,
             dtImposed=1e-5, dtMinLimit=1e-10, dTimeSystem=0.0, refCp=1004.5, refGamma=1.4,
             epsentr=0.05, dissFact=1.0, accelOn=0, accel=(0.0, 0.0, 0.0), limiterGuard=0)
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
