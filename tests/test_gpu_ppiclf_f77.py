"""ppiclF's F77 entry points (include/ppiclf_gpu.h), called through ctypes the
way rocpicl/PICL_TEMP_InitSolver.F90 and PICL_TEMP_Runge.F90 call ppiclF:
every argument by reference, one default instance bound to a Rocflu region.

Region A runs the fused coupled stage (picl_gpu_runge inside rk_stage).  The
"host" takes A's interpolation inputs of each stage (picl_gpu_get_cells),
fills all 8 nodes of every cell with them as PICL_TEMP_Runge does, hands
them to region B through InterpFieldUser and calls IntegrateParticle.  B's
particles and projection (read through GetProFldIJKEF) must equal A's: bit
for bit with the trilinear evaluation, which reads the element value, and
within 1e-12 with IDW, whose element value is the node mean sum/8 (A stores
that mean, so the host's nodes hold the mean and B takes the mean of means,
which can differ in the last bit; measured 1.2e-15).
"""
import ctypes as C

import numpy as np
import pytest

from rocflupicl_amd import _lib, configs, rflu

pytestmark = pytest.mark.gpu

H = 0.0005
NODE2V = np.array([0, 3, 4, 7, 1, 2, 5, 6])  # ppiclF node (i,j,k) -> c2v entry (PICL_TEMP_InitSolver.F90:538-597)


def ri(v):
    return C.byref(C.c_int(int(v)))


def rd(v):
    return C.byref(C.c_double(float(v)))


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def case(mono=False, **picl_over):
    xs = 0.1 * 12 * H
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=500, jitter=0.05, x_shock=xs,
                                                      x_curtain=xs - H, limiterGuard=1)
    if mono:  # Briney's added mass holds for equal spheres
        pc = pc.copy()
        pc[:, 1] = 115e-6
        pc[:, 2] = np.pi / 6.0 * 115e-6 ** 3
    picl.update(picl_over)
    return mesh, params, ic, y, pc, picl


def f77_init(L, mesh, params, ic, y, pc, picl):
    """PICL_TEMP_InitSolver's ppiclF calls on a fresh region."""
    b = rflu.Region(mesh, **params)
    b.set_state(ic(b.mesh_arrays()["cofg"]))
    d = dict(rflu.PICL_DEFAULTS)
    d.update(picl)
    pp = _lib.PiclParams(**d)
    assert L.ppiclf_gpu_bind(b.h, C.byref(pp)) == 0
    L.ppiclf_solve_initialize_(ri(0), rd(-1e4), rd(1e4), ri(0), rd(-1e4), rd(1e4), ri(0), rd(-1e4), rd(1e4),
                               ri(0), rd(0), rd(0), rd(0), rd(0))
    n = y.shape[0]
    rprop = np.zeros((n, 48))
    rprop[:, 0], rprop[:, 2], rprop[:, 3], rprop[:, 21] = pc[:, 0], pc[:, 1], pc[:, 2], pc[:, 3]
    yy = np.ascontiguousarray(y)
    L.ppiclf_solve_initparticle_(ri(2), ri(3), ri(0), ri(n), ptr(yy), ptr(rprop), rd(2 * H), rd(4 * 115e-6))
    if d["projMode"] == 1:
        L.ppiclf_solve_initgaussianfilter_(rd(d["filterWidth"]), rd(d["filterAlpha"]), ri(0))
    else:
        L.ppiclf_solve_initboxfilter_(rd(2 * H), ri(0), ri(1))
    L.ppiclf_solve_initneighborbin_(rd(4 * 115e-6))
    xyz, c2v = np.asarray(mesh["xyz"]), np.asarray(mesh["c2v"])
    g = xyz[c2v[:, NODE2V]]  # (nCells, 8, 3), node-fastest
    xg, yg, zg = (np.ascontiguousarray(g[:, :, l]) for l in range(3))
    L.ppiclf_comm_initoverlapmesh_(ri(b.nCells), ri(2), ri(2), ri(2), ptr(xg), ptr(yg), ptr(zg))
    assert L.ppiclf_gpu_last_error().decode() == ""
    b.nPart = n
    return b


def fused(mesh, params, ic, y, pc, picl):
    a = rflu.Region(mesh, **params)
    a.set_state(ic(a.mesh_arrays()["cofg"]))
    a.picl_init(y, pc, **picl)
    return a


def pro_fld(L, nCells):
    """GetProFldIJKEF over every node, element and built field: (nCells, 5, 8)."""
    out = np.zeros((nCells, 5, 8))
    v = C.c_double()
    for e in range(nCells):
        for m in range(5):
            for k in range(2):
                for j in range(2):
                    for i in range(2):
                        L.ppiclf_solve_getprofldijkef_(ri(i + 1), ri(j + 1), ri(k + 1), ri(e + 1), ri(m + 1),
                                                       C.byref(v))
                        out[e, m, i + 2 * j + 4 * k] = v.value
    return out


def run_pair(picl_over, steps=2, mono=False):
    L = _lib.load()
    L.ppiclf_gpu_set_abort(0)
    mesh, params, ic, y, pc, picl = case(mono=mono, **picl_over)
    a = fused(mesh, params, ic, y, pc, picl)
    b = f77_init(L, mesh, params, ic, y, pc, picl)
    assert np.array_equal(pro_fld(L, b.nCells), a.picl_nodes())  # InitSolve's projection
    history = []
    for _ in range(steps):
        a.RFLU_TimeStepInviscid()
        dt = a.RFLU_MinimumTimeStep()
        for s in (1, 2, 3):
            a.rk_stage(s)
            f = a.picl_cells()["fields"]
            for k, jp in enumerate(rflu.PICL_FIELD_RPROP):
                infld = np.ascontiguousarray(np.repeat(f[:, k][:, None], 8, axis=1))
                L.ppiclf_solve_interpfielduser_(ri(jp), ptr(infld))
            if picl.get("collisionalFlag", 0) >= 3:  # vorticity and D(omega)/Dt (PICL_TEMP_Runge.F90:880-892)
                r = a.picl_rot_fields()
                for k, jp in enumerate((31, 32, 33, 43, 44, 45)):
                    infld = np.ascontiguousarray(np.repeat(r[:, k][:, None], 8, axis=1))
                    L.ppiclf_solve_interpfielduser_(ri(jp), ptr(infld))
            else:
                zero = np.zeros((b.nCells, 8))
                L.ppiclf_solve_interpfielduser_(ri(31), ptr(zero))  # vorticity: read by no model that is on
            L.ppiclf_solve_integrateparticle_(ri(1), ri(0), rd(dt), rd(a.time()[0]))
            assert L.ppiclf_gpu_last_error().decode() == ""
            history.append((a.picl_particles(), b.picl_particles(), a.picl_nodes(), pro_fld(L, b.nCells)))
    L.ppiclf_gpu_reset()
    L.ppiclf_gpu_set_abort(1)
    return history


@pytest.mark.parametrize("proj", [0, 1])
def test_f77_stages_equal_the_fused_stage_trilinear(proj):
    over = dict(interpMode=1, qsFlag=5, projMode=proj, filterWidth=2 * H, filterAlpha=1e-2)
    for pa, pb, na, nb in run_pair(over):
        for k in ("y", "y1", "ydot", "ydotc", "elem"):
            assert same_bits(pa[k], pb[k]) if k != "elem" else np.array_equal(pa[k], pb[k]), k
        assert same_bits(na, nb)
    assert np.abs(pa["y"][:, 3]).max() > 0.0


def test_f77_stages_with_torque_and_lift_equal_the_fused_stage():
    """COLLISIONS 4: the host hands the vorticity and its material
    derivative over as six more fields, and the rotational state follows."""
    over = dict(interpMode=1, qsFlag=5, collisionalFlag=4, ksp=10.0, erest=0.7)
    for pa, pb, na, nb in run_pair(over):
        for k in ("y", "y1", "ydot", "ydotc", "elem"):
            assert same_bits(pa[k], pb[k]) if k != "elem" else np.array_equal(pa[k], pb[k]), k
        assert same_bits(na, nb)
    assert np.abs(pa["ydot"][:, 7:10]).max() > 0.0  # the torques act


def test_f77_stages_with_briney_and_fluctuations_equal_the_fused_stage():
    """Briney's added mass and Osnes' fluctuations through the F77 route:
    InitNeighborBin's width (4 d_p) is the fused route's default, and
    InitBoxFilter(2 dx) leaves d2chk(2) = dx, which the fused route is given."""
    over = dict(interpMode=1, qsFlag=5, amFlag=2, qsFluct=2, fluctSeed=5, fluctHalfWidth=H, collisionalFlag=2,
                ksp=10.0, erest=0.7)
    for pa, pb, na, nb in run_pair(over, mono=True):
        for k in ("y", "y1", "ydot", "ydotc", "elem"):
            assert same_bits(pa[k], pb[k]) if k != "elem" else np.array_equal(pa[k], pb[k]), k
        assert same_bits(na, nb)
    assert np.abs(pa["y"][:, 3]).max() > 0.0


def test_f77_stages_follow_the_fused_stage_idw():
    for pa, pb, na, nb in run_pair(dict(interpMode=0)):
        assert np.array_equal(pa["elem"], pb["elem"])
        for k in ("y", "ydot"):
            scale = np.maximum(np.abs(pa[k]).max(axis=0), 1e-300)
            assert (np.abs(pa[k] - pb[k]) / scale).max() <= 1e-12, k
        assert (np.abs(na - nb) / np.maximum(np.abs(na).max(), 1e-300)).max() <= 1e-12


def test_f77_errors_are_reported_as_ppiclf_reports_them():
    L = _lib.load()
    L.ppiclf_gpu_set_abort(0)
    L.ppiclf_gpu_reset()
    try:
        v = C.c_double()
        L.ppiclf_solve_getprofldijkef_(ri(1), ri(1), ri(1), ri(1), ri(1), C.byref(v))
        assert "InitOverlap must come before" in L.ppiclf_gpu_last_error().decode()
        mesh, params, ic, y, pc, picl = case()
        b = rflu.Region(mesh, **params)
        b.set_state(ic(b.mesh_arrays()["cofg"]))
        d = dict(rflu.PICL_DEFAULTS)
        pp = _lib.PiclParams(**d)
        assert L.ppiclf_gpu_bind(b.h, C.byref(pp)) == 0
        # angular periodicity with linear x periodicity (ppiclf_solve.f:75-78)
        L.ppiclf_solve_initialize_(ri(1), rd(0), rd(1), ri(0), rd(0), rd(1), ri(0), rd(0), rd(1), ri(1), rd(0),
                                   rd(0), rd(0), rd(0))
        assert "Invalid Periodicity choice" in L.ppiclf_gpu_last_error().decode()
        L.ppiclf_solve_initialize_(ri(1), rd(0), rd(1), ri(0), rd(0), rd(1), ri(0), rd(0), rd(1), ri(0), rd(0),
                                   rd(0), rd(0), rd(0))
        assert "vertex minimum and maximum" in L.ppiclf_gpu_last_error().decode()
        L.ppiclf_solve_initialize_(ri(0), rd(0), rd(1), ri(0), rd(0), rd(1), ri(0), rd(0), rd(1), ri(0), rd(0),
                                   rd(0), rd(0), rd(0))
        n = y.shape[0]
        rprop = np.zeros((n, 48))
        yy = np.ascontiguousarray(y)
        L.ppiclf_solve_initparticle_(ri(1), ri(3), ri(0), ri(n), ptr(yy), ptr(rprop), rd(1e-3), rd(1e-3))
        assert "imethod must be 2" in L.ppiclf_gpu_last_error().decode()
        L.ppiclf_solve_initparticle_(ri(2), ri(3), ri(0), ri(n), ptr(yy), ptr(rprop), rd(1e-3), rd(1e-3))
        L.ppiclf_solve_initboxfilter_(rd(1e-3), ri(0), ri(0))
        assert "multi-element box filter" in L.ppiclf_gpu_last_error().decode()
    finally:
        L.ppiclf_gpu_reset()
        L.ppiclf_gpu_set_abort(1)
