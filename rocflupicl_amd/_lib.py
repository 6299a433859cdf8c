"""ctypes binding of librflu_gpu.so (the C ABI declared in include/rflu_gpu.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the library is missing, or no HIP device is present, the calls
raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ROCFLUPICL_LIB selects another build of the same library (e.g. a tuning
# variant); the default is the in-tree build.
LIB_PATH = os.environ.get("ROCFLUPICL_LIB", os.path.join(HERE, "librflu_gpu.so"))

# Every symbol include/rflu_gpu.h declares (checked by the CPU test suite).
RFLU_SYMBOLS = [
    "rflu_gpu_create", "rflu_gpu_destroy", "rflu_gpu_last_error", "rflu_gpu_sizes",
    "rflu_gpu_get_mesh", "rflu_gpu_get_stencils", "rflu_gpu_set_state", "rflu_gpu_get_state",
    "rflu_gpu_get_rhs", "rflu_gpu_get_grad", "rflu_gpu_time_step_inviscid",
    "rflu_gpu_minimum_time_step", "rflu_gpu_rk_init", "rflu_gpu_cell_gradients",
    "rflu_gpu_convective_fluxes", "rflu_gpu_source_terms", "rflu_gpu_integrate_source_terms",
    "rflu_gpu_rk_update", "rflu_gpu_rk_stage", "rflu_gpu_step", "rflu_gpu_get_time",
    "rflu_gpu_sync", "rflu_gpu_set_profiling", "rflu_gpu_kernel_time", "rflu_gpu_bandwidth_probe",
    "rflu_gpu_step_async", "rflu_gpu_query", "rflu_gpu_wait", "rflu_gpu_device_bytes", "rflu_gpu_mem_info",
    "rflu_batch_create", "rflu_batch_sizes", "rflu_batch_set_state", "rflu_batch_get_state",
    "rflu_batch_get_time", "rflu_gpu_autotune", "rflu_gpu_set_time", "rflu_gpu_set_fusion",
    "rflu_gpu_get_fusion",
]

# Every symbol include/picl_gpu.h declares.
PICL_SYMBOLS = [
    "picl_gpu_init", "picl_gpu_runge", "picl_gpu_sizes", "picl_gpu_get_particles", "picl_gpu_get_cells",
    "picl_gpu_get_interp", "picl_gpu_set_debug", "picl_gpu_locate", "picl_gpu_get_nodes",
    "picl_gpu_locate_static", "picl_gpu_get_rot_fields",
]

# Every symbol include/ppiclf_gpu.h declares (ppiclF's F77 entry points and
# the set-up of their default instance).
PPICLF_SYMBOLS = [
    "ppiclf_gpu_bind", "ppiclf_gpu_reset", "ppiclf_gpu_set_abort", "ppiclf_gpu_last_error", "ppiclf_gpu_handle",
    "ppiclf_solve_initialize_", "ppiclf_solve_initparticle_", "ppiclf_solve_initboxfilter_",
    "ppiclf_solve_initgaussianfilter_", "ppiclf_solve_initneighborbin_", "ppiclf_comm_initoverlapmesh_",
    "ppiclf_solve_interpfielduser_", "ppiclf_solve_integrateparticle_", "ppiclf_solve_getprofldijkef_",
]

# picl_gpu.h: PICL_LOCATE_STATIC_MAX, the largest n of picl_gpu_locate_static
PICL_LOCATE_STATIC_MAX = 262144

ERRORS = {1: "EINVAL", 2: "ENODEV", 3: "EHIP", 4: "EMESH", 5: "ENAN", 6: "ENEGATIVE"}


class MeshDesc(C.Structure):
    _fields_ = [("nVert", C.c_int), ("xyz", C.c_void_p), ("nvc", C.c_int), ("nCells", C.c_int),
                ("c2v", C.c_void_p), ("nPatches", C.c_int), ("bcType", C.c_void_p),
                ("bcOpt", C.c_void_p), ("pout", C.c_void_p), ("nBFaces", C.c_void_p),
                ("nvf", C.c_int), ("bf2v", C.c_void_p), ("bcVals", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("spaceDiscr", C.c_int), ("spaceOrder", C.c_int), ("reconst", C.c_int),
                ("patchOrder", C.c_int), ("cfl", C.c_double), ("dtImposed", C.c_double),
                ("dtMinLimit", C.c_double), ("dTimeSystem", C.c_double), ("refCp", C.c_double),
                ("refGamma", C.c_double), ("epsentr", C.c_double), ("dissFact", C.c_double),
                ("accelOn", C.c_int), ("accel", C.c_double * 3), ("limiterGuard", C.c_int),
                ("fluxPrecision", C.c_int), ("sigma", C.c_double), ("kp", C.c_double), ("ku", C.c_double),
                ("mInf", C.c_double)]


class PiclParams(C.Structure):
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


_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    PP = C.POINTER(C.c_void_p)
    L.rflu_gpu_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params), PP]
    L.rflu_gpu_destroy.argtypes = [P]
    L.rflu_gpu_last_error.argtypes = [P]
    L.rflu_gpu_last_error.restype = C.c_char_p
    L.rflu_gpu_sizes.argtypes = [P, P, P, P, P]
    L.rflu_gpu_get_mesh.argtypes = [P, P, P, P, P, P, P]
    L.rflu_gpu_get_stencils.argtypes = [P, P, P, P]
    L.rflu_gpu_set_state.argtypes = [P, P]
    L.rflu_gpu_get_state.argtypes = [P, P, P]
    L.rflu_gpu_get_rhs.argtypes = [P, P]
    L.rflu_gpu_get_grad.argtypes = [P, P]
    L.rflu_gpu_time_step_inviscid.argtypes = [P, P, P]
    L.rflu_gpu_minimum_time_step.argtypes = [P, P]
    for n in ("rflu_gpu_rk_init", "rflu_gpu_rk_update", "rflu_gpu_rk_stage"):
        getattr(L, n).argtypes = [P, I]
    for n in ("rflu_gpu_cell_gradients", "rflu_gpu_convective_fluxes", "rflu_gpu_source_terms",
              "rflu_gpu_integrate_source_terms", "rflu_gpu_sync"):
        getattr(L, n).argtypes = [P]
    L.rflu_gpu_step.argtypes = [P, I, I]
    L.rflu_gpu_get_time.argtypes = [P, P, P]
    L.rflu_gpu_set_profiling.argtypes = [P, I]
    L.rflu_gpu_kernel_time.argtypes = [P, C.c_char_p, P, P]
    L.rflu_gpu_bandwidth_probe.argtypes = [C.c_size_t, I, P]
    L.rflu_gpu_step_async.argtypes = [P, I]
    L.rflu_gpu_query.argtypes = [P]
    L.rflu_gpu_wait.argtypes = [P]
    L.rflu_gpu_device_bytes.argtypes = [P, P]
    L.rflu_gpu_mem_info.argtypes = [P, P]
    L.rflu_batch_create.argtypes = [I, C.POINTER(MeshDesc), C.POINTER(Params), PP]
    L.rflu_batch_sizes.argtypes = [P, I, P, P]
    L.rflu_batch_set_state.argtypes = [P, I, P]
    L.rflu_batch_get_state.argtypes = [P, I, P, P]
    L.rflu_batch_get_time.argtypes = [P, I, P, P]
    L.rflu_gpu_autotune.argtypes = [P, I, P, P]
    L.rflu_gpu_set_time.argtypes = [P, C.c_double]
    L.picl_gpu_init.argtypes = [P, C.POINTER(PiclParams), I, P, P]
    L.picl_gpu_runge.argtypes = [P, I]
    L.picl_gpu_sizes.argtypes = [P, P, P]
    L.picl_gpu_get_particles.argtypes = [P, P, P, P, P, P]
    L.picl_gpu_get_cells.argtypes = [P, P, P, P, P]
    L.picl_gpu_get_interp.argtypes = [P, P]
    L.picl_gpu_get_nodes.argtypes = [P, P]
    L.picl_gpu_get_rot_fields.argtypes = [P, P]
    L.picl_gpu_set_debug.argtypes = [P, I]
    L.picl_gpu_locate.argtypes = [P, I, P, P]
    L.picl_gpu_locate_static.argtypes = [P, I, P, P]
    for n in RFLU_SYMBOLS + PICL_SYMBOLS:
        if n != "rflu_gpu_last_error":
            getattr(L, n).restype = I
    # ppiclF's F77 entry points: every argument by reference, no result
    nargs = {"ppiclf_solve_initialize_": 14, "ppiclf_solve_initparticle_": 8, "ppiclf_solve_initboxfilter_": 3,
             "ppiclf_solve_initgaussianfilter_": 3, "ppiclf_solve_initneighborbin_": 1,
             "ppiclf_comm_initoverlapmesh_": 7, "ppiclf_solve_interpfielduser_": 2,
             "ppiclf_solve_integrateparticle_": 4, "ppiclf_solve_getprofldijkef_": 6}
    for n, k in nargs.items():
        getattr(L, n).argtypes = [P] * k
        getattr(L, n).restype = None
    L.ppiclf_gpu_bind.argtypes = [P, C.POINTER(PiclParams)]
    L.ppiclf_gpu_bind.restype = I
    L.ppiclf_gpu_reset.restype = I
    L.ppiclf_gpu_set_abort.argtypes = [I]
    L.ppiclf_gpu_set_abort.restype = None
    L.ppiclf_gpu_last_error.restype = C.c_char_p
    L.ppiclf_gpu_handle.restype = P
    _lib = L
    return L


def exported_symbols():
    """Names the shared object exports (read with nm; no GPU needed)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


class GpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(rc, handle=None):
    if rc != 0:
        L = load()
        raise GpuError(rc, L.rflu_gpu_last_error(handle).decode())
