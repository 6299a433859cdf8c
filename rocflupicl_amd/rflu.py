"""Host-side mirror of the Rocflu residual-path call surface over the HIP C ABI.

``Region`` plays the role of the reference's ``TYPE(t_region)`` for the hot
path: it owns one region resident in HBM and exposes the reference's
subroutines under their own names (see include/rflu_gpu.h for the file:line
each one replaces).  Arrays cross the boundary in the reference's Fortran
layout: cv[N,5] (cv(5,N)), dv[N,3], grad[N,5,3].
"""
import ctypes as C

import numpy as np

from . import _lib
from .meshgen import BC_OUTFLOW, BC_SLIPWALL  # noqa: F401  (re-exported)

DISCR_ROE, DISCR_HLLC, DISCR_AUSM, DISCR_AUSMUP = 1, 3, 4, 5
RECONST_NONE, RECONST_WENO_XYZ, RECONST_VENKAT = 0, 2, 11

DEFAULTS = dict(spaceDiscr=DISCR_AUSM, spaceOrder=1, reconst=RECONST_NONE, patchOrder=1,
                cfl=0.5, dtImposed=1.0e-5, dtMinLimit=1.0e-10, dTimeSystem=0.0,
                refCp=1004.5, refGamma=1.4, epsentr=0.05, dissFact=1.0, accelOn=0,
                accel=(0.0, 0.0, 0.0), limiterGuard=0, fluxPrecision=0, sigma=1.0, kp=0.25, ku=0.75,
                mInf=0.1)


PICL_DEFAULTS = dict(rmuFlag=1, rmuRef=1.716e-5, tref=273.15, suth=110.4, rcpPart=840.0, qsFlag=1,
                     amFlag=1, pgFlag=1, feedbackFlag=1, interpMode=0, projMode=0, filterWidth=0.0,
                     filterAlpha=1e-2, heatTransfer=0, periodicX=0, periodicY=0,
                     periodicZ=0, viscousUnsteady=0, collisionalFlag=0, ksp=10.0, erest=0.9,
                     neighborWidth=0.0, qsFluct=0, qsFluctAdapt=0, fluctHalfWidth=0.0, fluctSeed=0,
                     angularPeriodic=0, angAngle=0.0, angXAngle=0.0, angRin=0.0, angRout=0.0)


def angular_radians(deg):
    """A wedge angle of the deck (degrees) in radians, as
    PICL_TEMP_InitSolver.F90:271-273 converts it: deg * acos(-1) / 180."""
    return float(deg) * float(np.arccos(-1.0)) / 180.0
# rprop index (1-based, PPICLF_USER.h) of each interpolated field slot
PICL_FIELD_RPROP = [2, 6, 7, 8, 10, 11, 12, 9, 24, 5, 13, 14, 15, 16, 17, 18, 19, 37, 38, 39]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def mesh_desc(mesh):
    """rflu_mesh_desc of a mesh dict (xyz, c2v, patches) and the arrays it
    points into (keep them alive while the descriptor is used)."""
    xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
    c2v = np.ascontiguousarray(mesh["c2v"], dtype=np.int32)
    pts = mesh["patches"]
    bt = np.array([p["bcType"] for p in pts], np.int32)
    bo = np.array([p["bcOpt"] for p in pts], np.int32)
    po = np.array([p["pout"] for p in pts], np.float64)
    nb = np.array([p["bf2v"].shape[0] for p in pts], np.int32)
    bf = np.ascontiguousarray(np.concatenate([p["bf2v"] for p in pts]), np.int32)
    bv = np.zeros((len(pts), 6))
    for ip, p in enumerate(pts):
        v = np.asarray(p.get("vals", ()), np.float64)
        bv[ip, :v.size] = v
    md = _lib.MeshDesc(xyz.shape[0], _p(xyz), c2v.shape[1], c2v.shape[0], _p(c2v), len(pts),
                       _p(bt), _p(bo), _p(po), _p(nb), bf.shape[1], _p(bf), _p(bv))
    return md, [xyz, c2v, bt, bo, po, nb, bf, bv]


def make_params(**kw):
    d = dict(DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError(f"unknown parameters {sorted(unknown)}")
    d.update(kw)
    p = _lib.Params()
    for k, v in d.items():
        if k == "accel":
            for i in range(3):
                p.accel[i] = float(v[i])
        else:
            setattr(p, k, v)
    return p


class Region:
    """One Rocflu region on the GPU (RFLU_InitFlowSolver for the hot path)."""

    def __init__(self, mesh, **params):
        L = _lib.load()
        self._L = L
        self.params = make_params(**params)
        md, self._keep = mesh_desc(mesh)
        h = C.c_void_p()
        rc = L.rflu_gpu_create(C.byref(md), C.byref(self.params), C.byref(h))
        _lib.check(rc, None)
        self.h = h
        n = [C.c_int() for _ in range(4)]
        L.rflu_gpu_sizes(self.h, *[C.byref(x) for x in n])
        self.nCells, self.nFaces, self.nBFaces, self.nStencil = (x.value for x in n)

    def close(self):
        if getattr(self, "h", None):
            self._L.rflu_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        _lib.check(rc, self.h)

    # ---- mesh products -------------------------------------------------
    def mesh_arrays(self):
        F, B, N = self.nFaces, self.nBFaces, self.nCells
        f2c = np.zeros((F, 2), np.int32)
        fn = np.zeros((F + B, 4))
        fc = np.zeros((F + B, 3))
        bf2c = np.zeros(B, np.int32)
        vol = np.zeros(N)
        cofg = np.zeros((N, 3))
        self._chk(self._L.rflu_gpu_get_mesh(self.h, _p(f2c), _p(fn), _p(fc), _p(bf2c), _p(vol), _p(cofg)))
        return dict(f2c=f2c, fn=fn[:F], fc=fc[:F], bfn=fn[F:], bfc=fc[F:], bf2c=bf2c, vol=vol, cofg=cofg)

    def stencils(self):
        off = np.zeros(self.nCells + 1, np.int32)
        mem = np.zeros(max(self.nStencil, 1), np.int32)
        moms = np.zeros((self.nCells, 10))
        self._chk(self._L.rflu_gpu_get_stencils(self.h, _p(off), _p(mem), _p(moms)))
        return off, mem[: self.nStencil], moms

    # ---- state ---------------------------------------------------------
    def set_state(self, cv):
        """cv as (N,5): rho, rho u, rho v, rho w, rho E per cell."""
        cv = np.asarray(cv, dtype=np.float64)
        assert cv.shape == (self.nCells, 5)
        soa = np.ascontiguousarray(cv.T)
        self._chk(self._L.rflu_gpu_set_state(self.h, _p(soa)))

    def get_state(self):
        cv = np.zeros((5, self.nCells))
        dv = np.zeros((3, self.nCells))
        self._chk(self._L.rflu_gpu_get_state(self.h, _p(cv), _p(dv)))
        return cv.T.copy(), dv.T.copy()

    def get_rhs(self):
        r = np.zeros((5, self.nCells))
        self._chk(self._L.rflu_gpu_get_rhs(self.h, _p(r)))
        return r.T.copy()

    def get_grad(self):
        g = np.zeros((self.nCells, 5, 3))
        self._chk(self._L.rflu_gpu_get_grad(self.h, _p(g)))
        return g

    # ---- the reference's routines --------------------------------------
    def RFLU_TimeStepInviscid(self, per_cell=False):
        dt = np.zeros(self.nCells) if per_cell else None
        m = C.c_double()
        self._chk(self._L.rflu_gpu_time_step_inviscid(self.h, _p(dt) if per_cell else None, C.byref(m)))
        return (m.value, dt) if per_cell else m.value

    def RFLU_MinimumTimeStep(self):
        m = C.c_double()
        self._chk(self._L.rflu_gpu_minimum_time_step(self.h, C.byref(m)))
        return m.value

    def RKInitMP(self, istage):
        self._chk(self._L.rflu_gpu_rk_init(self.h, istage))

    def CellGradientsMP(self):
        self._chk(self._L.rflu_gpu_cell_gradients(self.h))

    def ConvectiveFluxesMP(self):
        self._chk(self._L.rflu_gpu_convective_fluxes(self.h))

    def SourceTermsMP(self):
        self._chk(self._L.rflu_gpu_source_terms(self.h))

    def IntegrateSourceTermsMP(self):
        self._chk(self._L.rflu_gpu_integrate_source_terms(self.h))

    def RKUpdateMP(self, istage):
        self._chk(self._L.rflu_gpu_rk_update(self.h, istage))

    def rk_stage(self, istage):
        self._chk(self._L.rflu_gpu_rk_stage(self.h, istage))

    def RungeKuttaMP(self):
        for s in (1, 2, 3):
            self.rk_stage(s)

    def RFLU_TimeStepping(self, nSteps, use_graph=True):
        self._chk(self._L.rflu_gpu_step(self.h, int(nSteps), 1 if use_graph else 0))

    # ---- ppiclF coupling (include/picl_gpu.h) ---------------------------
    def picl_init(self, y, pconst, **picl_params):
        """PICL_TEMP_InitSolver: attach particles.  y (n,12) in ppiclF order,
        pconst (n,4) = (rho_p, d_p, V_p, SPL); the state must hold the
        pure-gas cv (it is scaled by 1 - piclVF)."""
        d = dict(PICL_DEFAULTS)
        unknown = set(picl_params) - set(d)
        if unknown:
            raise TypeError(f"unknown PICL parameters {sorted(unknown)}")
        d.update(picl_params)
        pp = _lib.PiclParams(**d)
        y = np.ascontiguousarray(y, np.float64)
        pc = np.ascontiguousarray(pconst, np.float64)
        assert y.ndim == 2 and y.shape[1] == 12 and pc.shape == (y.shape[0], 4)
        self._chk(self._L.picl_gpu_init(self.h, C.byref(pp), y.shape[0], _p(y), _p(pc)))
        self.nPart = y.shape[0]
        self._pconst = pc.copy()

    def write_particles(self, path, cycle=0):
        """ppiclf_io_WriteParticleVTU of the cloud: y, the constant rprop
        entries (rho_p, d_p, V_p, SPL at PPICLF_R_JRHOP, JDP, JVOLP, JSPL) and
        tags (rank 0, particle number, element number + 1; 0 when lost)."""
        from . import flowio
        p = self.picl_particles()
        n = self.nPart
        rprop = np.zeros((n, flowio.LRP))
        for col, j in zip(range(4), (0, 2, 3, 21)):
            rprop[:, j] = self._pconst[:, col]
        tags = np.stack([np.zeros(n), np.arange(1, n + 1), p["elem"] + 1.0], axis=1)
        flowio.write_particle_vtu(path, p["y"], rprop, tags, time=self.time()[0], cycle=cycle)

    def PICL_TEMP_Runge(self, istage):
        """The coupling stage: call after SourceTermsMP, before RKUpdateMP."""
        self._chk(self._L.picl_gpu_runge(self.h, istage))

    def picl_sizes(self):
        n, lost = C.c_int(), C.c_int()
        self._chk(self._L.picl_gpu_sizes(self.h, C.byref(n), C.byref(lost)))
        return n.value, lost.value

    def picl_particles(self):
        n = self.nPart
        y, y1, yd, yc = (np.zeros((n, 12)) for _ in range(4))
        elem = np.zeros(n, np.int32)
        self._chk(self._L.picl_gpu_get_particles(self.h, _p(y), _p(y1), _p(yd), _p(yc), _p(elem)))
        return dict(y=y, y1=y1, ydot=yd, ydotc=yc, elem=elem)

    def picl_cells(self):
        N = self.nCells
        phi, g, pf, fld = np.zeros(N), np.zeros((N, 3)), np.zeros((N, 5)), np.zeros((N, 20))
        self._chk(self._L.picl_gpu_get_cells(self.h, _p(phi), _p(g), _p(pf), _p(fld)))
        return dict(phi=phi, gradRhog=g, proFld=pf, fields=fld)

    def picl_nodes(self):
        """Projected values at the 8 nodes of every element, (nCells, 5, 8)."""
        a = np.zeros((self.nCells, 5, 8))
        self._chk(self._L.picl_gpu_get_nodes(self.h, _p(a)))
        return a

    def picl_rot_fields(self):
        """Vorticity and its material derivative per cell at the last stage
        (collisionalFlag >= 3), (nCells, 6)."""
        a = np.zeros((self.nCells, 6))
        self._chk(self._L.picl_gpu_get_rot_fields(self.h, _p(a)))
        return a

    def picl_set_debug(self, on):
        self._chk(self._L.picl_gpu_set_debug(self.h, 1 if on else 0))

    def picl_interp(self):
        a = np.zeros((self.nPart, 20))
        self._chk(self._L.picl_gpu_get_interp(self.h, _p(a)))
        return a

    def picl_locate(self, xyz, static=False):
        """Element of each point (-1: none).  static=True stages through the
        library's static device array (picl_gpu_locate_static): no device
        allocation, at most PICL_LOCATE_STATIC_MAX points."""
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros(xyz.shape[0], np.int32)
        f = self._L.picl_gpu_locate_static if static else self._L.picl_gpu_locate
        self._chk(f(self.h, xyz.shape[0], _p(xyz), _p(out)))
        return out

    def time(self):
        t, dt = C.c_double(), C.c_double()
        self._chk(self._L.rflu_gpu_get_time(self.h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def set_time(self, t):
        self._chk(self._L.rflu_gpu_set_time(self.h, float(t)))

    # ---- flow files (RFLU_ModReadWriteFlow) ------------------------------
    def write_flow(self, case, directory="", region=1, binary=True):
        """RFLU_WriteFlowBinary / RFLU_WriteFlowASCII of the current state (the
        ASCII form with its .mixt.dva dependent-variable companion) and
        RFLU_WriteRestartInfo; returns the flow file's name."""
        from . import flowio
        t = self.time()[0]
        name = flowio.file_name(case, ".mixt.cv" if binary else ".mixt.cva", region, t, directory)
        cv, dv = self.get_state()
        if binary:
            flowio.write_flow_binary(name, cv, t)
        else:
            flowio.write_flow_ascii(name, cv, t)
            flowio.write_dv_ascii(flowio.file_name(case, ".mixt.dva", region, t, directory), dv, t)
        flowio.append_restart_info(case, t, directory)
        return name

    def read_flow(self, path):
        """RFLU_ReadFlowBinary / ASCII: state and physical time from a flow file."""
        from . import flowio
        with open(path, "rb") as f:
            binary = f.read(4) == (80).to_bytes(4, "little")
        d = (flowio.read_flow_binary if binary else flowio.read_flow_ascii)(path)
        self.set_state(d["cv"])
        self.set_time(d["time"])
        return d["time"]

    def sync(self):
        self._chk(self._L.rflu_gpu_sync(self.h))

    # ---- asynchronous stepping (several regions in flight) --------------
    def step_async(self, nSteps):
        """Enqueue nSteps graph replays on this region's stream; returns at once."""
        self._chk(self._L.rflu_gpu_step_async(self.h, int(nSteps)))

    def idle(self):
        rc = self._L.rflu_gpu_query(self.h)
        if rc < 0:
            self._chk(rc)
        return rc == 1

    def wait(self):
        """Block until the stream drains; raise on AfterUpdateMP's checks."""
        self._chk(self._L.rflu_gpu_wait(self.h))

    def autotune(self, reps=5):
        """Pick the fastest of the bit-identical gradient kernels on this
        mesh, and (without particles) whether stages 2 and 3 run fused;
        returns (variant, ms per launch)."""
        v, ms = C.c_int(), C.c_double()
        self._chk(self._L.rflu_gpu_autotune(self.h, int(reps), C.byref(v), C.byref(ms)))
        return v.value, ms.value

    @property
    def fused(self):
        """Stages 2 and 3 of each step as one fused kernel after the gradients."""
        on = C.c_int()
        self._chk(self._L.rflu_gpu_get_fusion(self.h, C.byref(on)))
        return bool(on.value)

    @fused.setter
    def fused(self, on):
        self._chk(self._L.rflu_gpu_set_fusion(self.h, 1 if on else 0))

    def device_bytes(self):
        b = C.c_size_t()
        self._chk(self._L.rflu_gpu_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_profiling(self, on):
        self._chk(self._L.rflu_gpu_set_profiling(self.h, 1 if on else 0))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int()
        self._chk(self._L.rflu_gpu_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def mem_info():
    """(free, total) device memory in bytes."""
    L = _lib.load()
    f, t = C.c_size_t(), C.c_size_t()
    _lib.check(L.rflu_gpu_mem_info(C.byref(f), C.byref(t)), None)
    return f.value, t.value


class Batch(Region):
    """Batched ensemble: independent regions of one cell type under one
    parameter set in one handle (include/rflu_gpu.h rflu_batch_*).  Every
    kernel launch covers all instances; each instance keeps its own dt and
    time, and its result equals a Region run alone, bit for bit.  Step it
    with RFLU_TimeStepping / step_async; states are per instance."""

    def __init__(self, meshes, **params):
        L = _lib.load()
        self._L = L
        self.params = make_params(**params)
        descs, self._keep = [], []
        for m in meshes:
            md, keep = mesh_desc(m)
            descs.append(md)
            self._keep += keep
        arr = (_lib.MeshDesc * len(descs))(*descs)
        self._keep.append(arr)
        h = C.c_void_p()
        _lib.check(L.rflu_batch_create(len(descs), arr, C.byref(self.params), C.byref(h)), None)
        self.h = h
        n = [C.c_int() for _ in range(4)]
        L.rflu_gpu_sizes(self.h, *[C.byref(x) for x in n])
        self.nCells, self.nFaces, self.nBFaces, self.nStencil = (x.value for x in n)
        self.nInst = len(descs)
        self.instCells = []
        for k in range(self.nInst):
            nc = C.c_int()
            self._chk(L.rflu_batch_sizes(self.h, k, C.byref(nc), None))
            self.instCells.append(nc.value)
        self.instOff = np.concatenate([[0], np.cumsum(self.instCells)]).astype(int)
        self._cofg = None

    def cofg(self, inst):
        """Cell centroids of one instance, in its own cell numbering."""
        if self._cofg is None:
            self._cofg = self.mesh_arrays()["cofg"]
        return self._cofg[self.instOff[inst]:self.instOff[inst + 1]]

    def set_state(self, inst, cv):
        cv = np.asarray(cv, dtype=np.float64)
        assert cv.shape == (self.instCells[inst], 5)
        soa = np.ascontiguousarray(cv.T)
        self._chk(self._L.rflu_batch_set_state(self.h, int(inst), _p(soa)))

    def get_state(self, inst):
        n = self.instCells[inst]
        cv, dv = np.zeros((5, n)), np.zeros((3, n))
        self._chk(self._L.rflu_batch_get_state(self.h, int(inst), _p(cv), _p(dv)))
        return cv.T.copy(), dv.T.copy()

    def time(self, inst):
        t, dt = C.c_double(), C.c_double()
        self._chk(self._L.rflu_batch_get_time(self.h, int(inst), C.byref(t), C.byref(dt)))
        return t.value, dt.value
