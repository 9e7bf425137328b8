"""Python binding of the MI355X hot-path library (misc_amd/libastro_hip.so).

The functions keep the names, argument meaning and index conventions of the
reference routines they replace (Working/AStrO_solvers.f90,
Working/AStrO_bookKeeping.f90): 1-based column values with 0 for an unused
slot, (0:n) range arrays of n + 1 entries, rows of a skyline profile holding
columns i-len+1..i with the diagonal last.  Errors raise AstroError with the
library's status and message.  There is no fallback: if the library is not
built, importing the binding raises.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libastro_hip.so")

STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "zero pivot", 4: "no convergence",
          5: "device wait timeout", 6: "unsupported"}

# Exported symbols (include/astro_abi.h); checked by the CPU tests.
EXPORTS = [
    "astro_last_error", "astro_version", "astro_getSparseLDLFact", "astro_solveSparseLDLFact", "astro_gMRes",
    "astro_convertToLTri", "astro_ctx_create", "astro_ctx_destroy", "astro_ctx_scalar", "astro_ctx_scalar_f",
    "astro_ctx_get_i32", "astro_ctx_get_f64", "astro_ctx_set_f64", "astro_ctx_assemble_elastic",
    "astro_ctx_scale_mpc", "astro_ctx_build_profile", "astro_ctx_factor", "astro_ctx_solve",
    "astro_ctx_static_solve", "astro_ctx_assemble_thermal", "astro_ctx_thermal_solve", "astro_ctx_precond_apply",
    "astro_ctx_operator_apply", "astro_ctx_sync", "astro_ctx_phase_ms", "astro_prep_create", "astro_prep_destroy",
    "astro_prep_scalar", "astro_prep_get_i32", "astro_prep_get_f64", "astro_ctx_set_timing",
    "astro_ctx_kernel_stat", "astro_ctx_reset", "astro_ctx_true_residual", "astro_ctx_dynamic_init",
    "astro_ctx_dynamic_step", "astro_ctx_dynamic_run", "astro_ctx_set_design", "astro_ctx_dRdD",
    "astro_ctx_compliance_gradient", "astro_ctx_cs_batch_gradient", "astro_ctx_cs_batch_autotune",
    "astro_ctx_cs_batch_tile",
    "astro_ctx_elastic_soln_load", "astro_ctx_thermal_soln_load", "astro_ctx_element_stress",
    "astro_ctx_thermal_dynamic_init", "astro_ctx_thermal_dynamic_run",
    "astro_ctx_true_residual_noalloc", "astro_gnn_create", "astro_gnn_destroy", "astro_gnn_num_params",
    "astro_gnn_num_fields", "astro_gnn_get_params", "astro_gnn_set_params", "astro_ctx_gnn_collect",
    "astro_gnn_train", "astro_gnn_loss_grad", "astro_gnn_field_size", "astro_gnn_get_field",
    "astro_gnn_field_blocks", "astro_ctx_gnn_attach", "astro_ctx_gnn_detach", "astro_ctx_gnn_get",
    "astro_ctx_set_design_vars", "astro_ctx_dRdD_vars", "astro_ctx_dRtdD_vars", "astro_ctx_objective",
    "astro_ctx_objective_gradient", "astro_ctx_element_flux", "astro_ctx_shell_internal_force",
    "astro_ctx_dRdT_product", "astro_ctx_dynamic_objective_gradient", "astro_ctx_nonlinear_solve",
    "astro_ctx_nonlinear_tangent",
]

# Design-variable categories and objective terms (include/astro_abi.h).
DV_CATEGORY = {"modulus": 1, "poissonRatio": 2, "shearModulus": 3, "density": 4, "thermalCond": 5, "nodeCoord": 6,
               "nodeLoad": 7}
OBJ_CATEGORY = {"displacement": 1, "temperature": 2, "stress": 3, "strain": 4, "strainEnergy": 5, "mass": 6,
                "volume": 7}
OBJ_OPERATOR = {"powerNorm": 1, "volumeIntegral": 2, "volumeAverage": 3}


class AstroError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_i32 = C.c_int32
_lib = None


class _Model(C.Structure):
    _fields_ = [("numNodes", C.c_int32), ("numEls", C.c_int32), ("nodeList", C.c_void_p),
                ("elementList", C.c_void_p), ("elementType", C.c_void_p), ("eprops", C.c_void_p),
                ("density", C.c_void_p), ("cond", C.c_void_p), ("numMpcTerms", C.c_int32), ("mpcEqn", C.c_void_p),
                ("mpcDof", C.c_void_p), ("mpcCoef", C.c_void_p), ("mpcIsSet", C.c_void_p),
                ("mpcNodeStart", C.c_void_p), ("mpcNodes", C.c_void_p), ("numMPC", C.c_int32),
                ("mpcRHS", C.c_void_p), ("numLoads", C.c_int32), ("loadNode", C.c_void_p),
                ("loadVal", C.c_void_p), ("numLayers", C.c_int32), ("layers", C.c_void_p), ("zOffset", C.c_void_p),
                ("orient", C.c_void_p), ("specHeat", C.c_void_p), ("thermExp", C.c_void_p)]


class _Design(C.Structure):
    _fields_ = [("numDV", C.c_int32), ("category", C.c_void_p), ("component", C.c_void_p),
                ("elToDRange", C.c_void_p), ("elToD", C.c_void_p), ("elToCoef", C.c_void_p),
                ("ndToDRange", C.c_void_p), ("ndToD", C.c_void_p), ("ndToCoef", C.c_void_p), ("h", C.c_double)]


class _ObjTerm(C.Structure):
    _fields_ = [("category", C.c_int32), ("op", C.c_int32), ("component", C.c_int32), ("coef", C.c_double),
                ("expo", C.c_double), ("numItems", C.c_int32), ("items", C.c_void_p), ("numTgt", C.c_int32),
                ("tgt", C.c_void_p)]


@dataclass
class ObjectiveTerm:
    """One objective term as the reference's objective file states it
    (Working/objectiveTemplate.yaml; Working/AStrO_objective.f90:9-470):
    category ('displacement', 'temperature', 'stress', 'strain',
    'strainEnergy', 'mass', 'volume'), operator ('powerNorm',
    'volumeIntegral', 'volumeAverage'), 1-based component, coefficient,
    exponent, the 1-based node or element ids of its set, and the target
    (None: 'none'; a scalar: 'single'; one value per item: 'vector')."""

    category: str
    items: object
    operator: str = "powerNorm"
    component: int = 1
    coef: float = 1.0
    expo: float = 1.0
    target: object = None


@dataclass
class DesignVariables:
    """Design variables as readDesignVarInput leaves them
    (Working/AStrO_input.f90:1099-1408): per variable its category (a key of
    DV_CATEGORY) and 1-based component, the element map (elToDRange of
    numEls + 1 offsets, 1-based variable ids elToD, coefficients elToCoef)
    and the node map (ndToDRange, ndToD, ndToCoef)."""

    category: list
    component: list
    elToDRange: object = None
    elToD: object = None
    elToCoef: object = None
    ndToDRange: object = None
    ndToD: object = None
    ndToCoef: object = None

    @property
    def num(self):
        return len(self.category)


def _marshal_terms(terms):
    keep = []

    def ptr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    arr = (_ObjTerm * len(terms))()
    for t, T in zip(arr, terms):
        items = np.ascontiguousarray(T.items, dtype=np.int32).reshape(-1)
        t.category = OBJ_CATEGORY[T.category]
        t.op = OBJ_OPERATOR[T.operator]
        t.component = int(T.component)
        t.coef = float(T.coef)
        t.expo = float(T.expo)
        t.numItems = int(items.size)
        t.items = ptr(items, np.int32)
        if T.target is None:
            t.numTgt = 0
        else:
            tg = np.atleast_1d(np.asarray(T.target, dtype=np.float64))
            t.numTgt = int(tg.size)
            t.tgt = ptr(tg, np.float64)
    return arr, keep


class _Options(C.Structure):
    _fields_ = [("maxBW", C.c_int32), ("literalMaxBW", C.c_int32), ("thermLiteralD8", C.c_int32),
                ("numVecs", C.c_int32), ("rtol", C.c_double), ("maxCycles", C.c_int32),
                ("coupledProfile", C.c_int32), ("precision", C.c_int32),
                ("stepGraphs", C.c_int32)]


def load_library(path: str = LIB_PATH):
    """Load the library; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    L.astro_last_error.restype = C.c_char_p
    L.astro_version.restype = C.c_char_p
    L.astro_getSparseLDLFact.argtypes = [_dp, _i64, _ip, _i32]
    L.astro_solveSparseLDLFact.argtypes = [_dp, _dp, _dp, _i64, _ip, _i32, _dp]
    L.astro_gMRes.argtypes = [_dp, _dp, _ip, _ip, _i64, _i32, _dp, _ip, _ip, _i64, _i32, _dp, _ip, _i64, _dp, _i32,
                              _i32, C.c_double, C.c_double]
    L.astro_convertToLTri.argtypes = [_dp, _ip, _i64, _dp, _ip, _ip, _i64, _i32, _dp, _ip, _ip, _i64, _i32]
    L.astro_ctx_create.argtypes = [C.POINTER(_Model), C.POINTER(_Options), C.POINTER(C.c_void_p)]
    L.astro_ctx_destroy.argtypes = [C.c_void_p]
    L.astro_ctx_scalar.restype = C.c_int64
    L.astro_ctx_scalar.argtypes = [C.c_void_p, C.c_char_p]
    L.astro_ctx_scalar_f.restype = C.c_double
    L.astro_ctx_scalar_f.argtypes = [C.c_void_p, C.c_char_p]
    L.astro_ctx_get_i32.argtypes = [C.c_void_p, C.c_char_p, _ip, _i64]
    L.astro_ctx_get_f64.argtypes = [C.c_void_p, C.c_char_p, _dp, _i64]
    L.astro_ctx_set_f64.argtypes = [C.c_void_p, C.c_char_p, _dp, _i64]
    L.astro_ctx_assemble_elastic.argtypes = [C.c_void_p, C.c_double]
    for f in ("astro_ctx_scale_mpc", "astro_ctx_build_profile", "astro_ctx_factor", "astro_ctx_solve",
              "astro_ctx_static_solve", "astro_ctx_assemble_thermal", "astro_ctx_thermal_solve", "astro_ctx_sync",
              "astro_ctx_reset", "astro_ctx_elastic_soln_load", "astro_ctx_thermal_soln_load"):
        getattr(L, f).argtypes = [C.c_void_p]
    L.astro_ctx_dynamic_init.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
    L.astro_ctx_dynamic_step.argtypes = [C.c_void_p, C.c_double]
    L.astro_ctx_dynamic_run.argtypes = [C.c_void_p, C.c_int, _dp]
    L.astro_ctx_set_design.argtypes = [C.c_void_p, C.c_int, _ip, C.c_void_p, C.c_int, C.c_double]
    L.astro_ctx_dRdD.argtypes = [C.c_void_p, _dp]
    L.astro_ctx_compliance_gradient.argtypes = [C.c_void_p, _dp]
    L.astro_ctx_cs_batch_gradient.argtypes = [C.c_void_p, _dp, C.c_void_p]
    L.astro_ctx_cs_batch_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    L.astro_ctx_cs_batch_tile.argtypes = [C.c_void_p, C.c_int32]
    L.astro_ctx_element_stress.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.astro_ctx_element_flux.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.astro_ctx_shell_internal_force.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, _dp, _dp]
    L.astro_ctx_dRdT_product.argtypes = [C.c_void_p, _dp, _dp]
    L.astro_ctx_nonlinear_tangent.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    L.astro_ctx_nonlinear_solve.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                                  C.POINTER(C.c_int32)]
    L.astro_ctx_dynamic_objective_gradient.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, C.c_void_p, C.c_void_p,
                                                       C.c_int32, C.c_void_p, _dp, _dp]
    L.astro_ctx_thermal_dynamic_init.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.astro_ctx_thermal_dynamic_run.argtypes = [C.c_void_p, C.c_int]
    L.astro_ctx_precond_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.astro_ctx_operator_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.astro_prep_create.argtypes = [C.POINTER(_Model), C.POINTER(_Options), C.POINTER(C.c_void_p)]
    L.astro_prep_destroy.argtypes = [C.c_void_p]
    L.astro_prep_scalar.restype = C.c_int64
    L.astro_prep_scalar.argtypes = [C.c_void_p, C.c_char_p]
    L.astro_prep_get_i32.argtypes = [C.c_void_p, C.c_char_p, _ip, _i64]
    L.astro_prep_get_f64.argtypes = [C.c_void_p, C.c_char_p, _dp, _i64]
    L.astro_ctx_true_residual.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.astro_ctx_true_residual_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    L.astro_ctx_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.astro_ctx_kernel_stat.restype = C.c_double
    L.astro_ctx_kernel_stat.argtypes = [C.c_void_p, C.c_char_p]
    L.astro_ctx_phase_ms.restype = C.c_double
    L.astro_ctx_phase_ms.argtypes = [C.c_void_p, C.c_char_p]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.astro_gnn_create.argtypes = [C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]
    L.astro_gnn_destroy.argtypes = [C.c_void_p]
    L.astro_gnn_num_params.restype = C.c_int64
    L.astro_gnn_num_params.argtypes = [C.c_void_p]
    L.astro_gnn_num_fields.argtypes = [C.c_void_p]
    L.astro_gnn_get_params.argtypes = [C.c_void_p, _fp, _i64]
    L.astro_gnn_set_params.argtypes = [C.c_void_p, _fp, _i64]
    L.astro_ctx_gnn_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_int32]
    L.astro_gnn_train.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_void_p]
    L.astro_gnn_loss_grad.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    L.astro_gnn_field_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.astro_gnn_get_field.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _fp, _dp, _dp, C.POINTER(C.c_double)]
    L.astro_gnn_field_blocks.argtypes = [C.c_void_p, C.c_int, _dp]
    L.astro_ctx_gnn_attach.argtypes = [C.c_void_p, C.c_void_p]
    L.astro_ctx_gnn_detach.argtypes = [C.c_void_p]
    L.astro_ctx_gnn_get.argtypes = [C.c_void_p, C.c_char_p, _dp, _i64]
    L.astro_ctx_set_design_vars.argtypes = [C.c_void_p, C.POINTER(_Design)]
    L.astro_ctx_dRdD_vars.argtypes = [C.c_void_p, _dp]
    L.astro_ctx_dRtdD_vars.argtypes = [C.c_void_p, _dp]
    L.astro_ctx_objective.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, _dp]
    L.astro_ctx_objective_gradient.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, _dp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise AstroError(rc, load_library().astro_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------------------
# Reference routine replacements (host arrays, synchronous)

def getSparseLDLFact(ALTri, ARange, n):
    """In-place LDL^T of a row-skyline profile (Working/AStrO_solvers.f90:10-78).
    Returns the factored array (a new array; the input is not modified)."""
    A = _f64(ALTri).copy()
    _check(load_library().astro_getSparseLDLFact(A, A.size, _i(ARange), int(n)))
    return A


def solveSparseLDLFact(Lmat, LRange, rhs):
    """x = (L D L^T)^-1 rhs (Working/AStrO_solvers.f90:83-126); returns (soln, zVec)."""
    rhs = _f64(rhs)
    n = rhs.size
    x = np.zeros(n)
    z = np.zeros(n)
    L = _f64(Lmat)
    _check(load_library().astro_solveSparseLDLFact(x, rhs, L, L.size, _i(LRange), n, z))
    return x, z


def gMRes(xVec, Amat, Acols, ARange, Cmat, Ccols, CRange, Pmat, PRange, bVec, numVecs, maxIt, rtol=1e-12,
          atol=0.0, allow_no_converge=False):
    """Preconditioned GMRES(numVecs) (Working/AStrO_solvers.f90:128-297).  xVec is
    the initial guess; returns the solution.  rtol <= 0 and atol <= 0 select the
    reference's absolute 1e-12 stop.  allow_no_converge: return the last
    iterate when the maxIt/numVecs + 1 restart cycles run out (the reference
    returns it without a word) instead of raising."""
    x = _f64(xVec).copy()
    A, Ac, Ar = _f64(Amat), _i(Acols), _i(ARange)
    cdim = len(CRange) - 1 if CRange is not None and len(CRange) > 1 else 0
    Cm = _f64(Cmat) if cdim else np.zeros(1)
    Cc = _i(Ccols) if cdim else np.zeros(1, np.int32)
    Cr = _i(CRange) if cdim else np.zeros(1, np.int32)
    P = _f64(Pmat)
    rc = load_library().astro_gMRes(x, A, Ac, Ar, A.size, x.size, Cm, Cc, Cr, Cm.size, cdim, P, _i(PRange), P.size,
                                    _f64(bVec), int(numVecs), int(maxIt), float(rtol), float(atol))
    if not (allow_no_converge and rc == 4):
        _check(rc)
    return x


def convertToLTri(lTRange, mainMat, mainCols, mainRange, cMat=None, cCols=None, cRange=None):
    """Profile from CSR plus penalty C^T C (Working/AStrO_bookKeeping.f90:754-792)."""
    lTRange = _i(lTRange)
    n = lTRange.size - 1
    lt = np.zeros(int(lTRange[-1]))
    M = _f64(mainMat)
    cdim = len(cRange) - 1 if cRange is not None and len(cRange) > 1 else 0
    Cm = _f64(cMat) if cdim else np.zeros(1)
    Cc = _i(cCols) if cdim else np.zeros(1, np.int32)
    Cr = _i(cRange) if cdim else np.zeros(1, np.int32)
    _check(load_library().astro_convertToLTri(lt, lTRange, lt.size, M, _i(mainCols), _i(mainRange), M.size, n, Cm,
                                              Cc, Cr, Cm.size, cdim))
    return lt


# ---------------------------------------------------------------------------
@dataclass
class Options:
    maxBW: int = 0
    literalMaxBW: bool = True
    thermLiteralD8: bool = True
    numVecs: int = 0
    rtol: float = 1e-12
    maxCycles: int = 1000
    coupledProfile: bool = False
    precision: int = 0  # 0: FP64 factor; 1: FP32-stored factor, FP32 matrix-core updates, FP64 refinement
    stepGraphs: bool = False  # dynamic_run replays one HIP graph per Newmark step


def _marshal(model, o: "Options"):
    keep = []

    def ptr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    M = _Model()
    M.numNodes = model.num_nodes
    M.numEls = model.num_els
    M.nodeList = ptr(model.nodes.reshape(-1), np.float64)
    M.elementList = ptr(model.elements.reshape(-1), np.int32)
    M.elementType = ptr(model.etype, np.int32)
    M.eprops = ptr(model.eprops.reshape(-1), np.float64)
    M.density = ptr(model.density, np.float64)
    M.cond = ptr(model.cond.reshape(-1), np.float64)
    if model.mpc:
        M.numMpcTerms = len(model.mpc)
        M.mpcEqn = ptr([t.eqn for t in model.mpc], np.int32)
        M.mpcDof = ptr([t.dof for t in model.mpc], np.int32)
        M.mpcCoef = ptr([t.coef for t in model.mpc], np.float64)
        M.mpcIsSet = ptr([1 if t.is_set else 0 for t in model.mpc], np.int32)
        starts = np.zeros(len(model.mpc) + 1, np.int32)
        starts[1:] = np.cumsum([len(t.nodes) for t in model.mpc])
        M.mpcNodeStart = ptr(starts, np.int32)
        M.mpcNodes = ptr(np.concatenate([t.nodes for t in model.mpc]), np.int32)
        M.numMPC = len(model.mpc_rhs)
        M.mpcRHS = ptr(model.mpc_rhs, np.float64)
    M.numLoads = int(model.load_nodes.size)
    if M.numLoads:
        M.loadNode = ptr(model.load_nodes, np.int32)
        M.loadVal = ptr(model.load_vals.reshape(-1), np.float64)
    if getattr(model, "layers", None) is not None:
        M.numLayers = int(model.layers.shape[1])
        M.layers = ptr(model.layers.reshape(-1), np.float64)
        M.zOffset = ptr(model.z_offset, np.float64)
        M.orient = ptr(model.orient.reshape(-1), np.float64)
    if getattr(model, "spec_heat", None) is not None:
        M.specHeat = ptr(model.spec_heat, np.float64)
    if getattr(model, "therm_exp", None) is not None:
        M.thermExp = ptr(np.asarray(model.therm_exp).reshape(-1), np.float64)
    O = _Options(int(o.maxBW), 1 if o.literalMaxBW else 0, 1 if o.thermLiteralD8 else 0, int(o.numVecs),
                 float(o.rtol), int(o.maxCycles), 1 if o.coupledProfile else 0, int(o.precision),
                 1 if o.stepGraphs else 0)
    return M, O, keep


class Prep:
    """Host-only analysisPrep of the library (no device needed)."""

    def __init__(self, model, options: Optional[Options] = None):
        L = load_library()
        self.model = model
        M, O, keep = _marshal(model, options or Options())
        h = C.c_void_p()
        _check(L.astro_prep_create(C.byref(M), C.byref(O), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load_library().astro_prep_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def scalar(self, name):
        return int(load_library().astro_prep_scalar(self.h, name.encode()))

    def geti(self, name):
        n = self.scalar
        N = self.model.num_nodes
        sizes = {"currentRank": N, "nDofIndex": 6 * N, "elMatRange": n("elMatDim") + 1,
                 "elMatLTRange": n("elMatDim") + 1, "elMatCols": n("elMatSize"), "elMPCMatRange": n("elMPCDim") + 1,
                 "elMPCMatCols": n("elMPCSize"), "thermMatRange": N + 1, "thermMatLTRange": N + 1,
                 "thermMatCols": n("thermMatSize"), "thermMPCMatRange": n("thermMPCDim") + 1}
        out = np.zeros(sizes[name], np.int32)
        _check(load_library().astro_prep_get_i32(self.h, name.encode(), out, out.size))
        return out

    def getd(self, name):
        sizes = {"elMPCMat": self.scalar("elMPCSize"), "elMPCRHS": self.scalar("elMPCDim")}
        out = np.zeros(sizes[name])
        _check(load_library().astro_prep_get_f64(self.h, name.encode(), out, out.size))
        return out


class Context:
    """Device-resident model: the reference's global state of one job in HBM."""

    def __init__(self, model, options: Optional[Options] = None):
        L = load_library()
        self.model = model
        M, O, keep = _marshal(model, options or Options())
        h = C.c_void_p()
        _check(L.astro_ctx_create(C.byref(M), C.byref(O), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load_library().astro_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalar(self, name):
        return int(load_library().astro_ctx_scalar(self.h, name.encode()))

    def scalar_f(self, name):
        return float(load_library().astro_ctx_scalar_f(self.h, name.encode()))

    def _len(self, name):
        n = self.scalar
        sizes = {
            "currentRank": self.model.num_nodes, "nDofIndex": 6 * self.model.num_nodes,
            "elMatRange": n("elMatDim") + 1, "elMatLTRange": n("elMatDim") + 1, "elMatCols": n("elMatSize"),
            "elasticMat": n("elMatSize"), "elMatLT": n("elMatLTSize"), "elasticLoad": n("elMatDim"),
            "nodeDisp": n("elMatDim"), "delDisp": n("elMatDim"), "nodeVel": n("elMatDim"), "nodeAcc": n("elMatDim"),
            "prevDisp": n("elMatDim"), "thermMatRange": self.model.num_nodes + 1,
            "thermMatLTRange": self.model.num_nodes + 1, "thermMatCols": n("thermMatSize"),
            "thermMat": n("thermMatSize"), "thermMatLT": n("thermMatLTSize"), "thermalLoad": self.model.num_nodes,
            "nodeTemp": self.model.num_nodes, "nodeTdot": self.model.num_nodes, "elMPCMatRange": n("elMPCDim") + 1,
            "elMPCMatCols": n("elMPCSize"), "elMPCMat": n("elMPCSize"), "elMPCRHS": n("elMPCDim"),
        }
        return sizes[name]

    def geti(self, name, count=None):
        count = self._len(name) if count is None else count
        out = np.zeros(count, np.int32)
        _check(load_library().astro_ctx_get_i32(self.h, name.encode(), out, count))
        return out

    def getd(self, name, count=None):
        count = self._len(name) if count is None else count
        out = np.zeros(count)
        _check(load_library().astro_ctx_get_f64(self.h, name.encode(), out, count))
        return out

    def getd_sized(self, name, count):
        out = np.zeros(count)
        _check(load_library().astro_ctx_get_f64(self.h, name.encode(), out, count))
        return out

    def setd(self, name, values):
        v = _f64(values)
        _check(load_library().astro_ctx_set_f64(self.h, name.encode(), v, v.size))

    def assemble(self, mass_factor=0.0):
        _check(load_library().astro_ctx_assemble_elastic(self.h, float(mass_factor)))

    def scale_mpc(self):
        _check(load_library().astro_ctx_scale_mpc(self.h))

    def build_profile(self):
        _check(load_library().astro_ctx_build_profile(self.h))

    def factor(self):
        _check(load_library().astro_ctx_factor(self.h))

    def solve(self):
        _check(load_library().astro_ctx_solve(self.h))

    def reset(self):
        _check(load_library().astro_ctx_reset(self.h))

    def static_solve(self):
        _check(load_library().astro_ctx_static_solve(self.h))

    def elastic_soln_load(self):
        """getElasticSolnLoad(0) on a zeroed load: "elasticLoad" = -R(u) of the
        current nodeDisp (Working/AStrO_commandFunctions.f90:260-351)."""
        _check(load_library().astro_ctx_elastic_soln_load(self.h))

    def thermal_soln_load(self):
        """getThermalSolnLoad(0) on a zeroed load: "thermalLoad" = -K T."""
        _check(load_library().astro_ctx_thermal_soln_load(self.h))

    def node_values(self, name, nodes):
        """Values of a DOF vector ("nodeDisp", "elasticLoad", ...) at the given
        1-based nodes as (len(nodes), 6), 0 where a node has no such DOF, the
        layout of writeNodeResults (Working/AStrO_output.f90:60-141)."""
        v = self.getd(name)
        nd = self.geti("nDofIndex").reshape(-1, 6)
        out = np.zeros((len(nodes), 6))
        for r, node in enumerate(nodes):
            for c in range(6):
                d = nd[node - 1, c]
                if d:
                    out[r, c] = v[d - 1]
        return out

    def reaction_force(self, nodes):
        """writeNodeResults' reactionForce field at `nodes`: -elasticLoad after
        getElasticSolnLoad(0) (Working/AStrO_runJob.f90:228-233,
        Working/AStrO_output.f90:122-141)."""
        self.elastic_soln_load()
        return -self.node_values("elasticLoad", nodes)

    def displacement_objective(self, nodes, component, coef=1.0, exponent=1.0, target=0.0):
        """The 'displacement' objective term: coef * sum over the node set of
        (u_component - target)^exponent (Working/AStrO_objective.f90:147-158);
        component is 1-based."""
        u = self.node_values("nodeDisp", nodes)[:, component - 1]
        return float(coef * np.sum((u - target) ** exponent))

    def dynamic_init(self, beta: float = 0.25, gamma: float = 0.5, dt: float = 1e-5):
        """Newmark set-up: tangent K + M/(beta dt^2), penalty scaling, profile,
        factor; the state starts at rest."""
        _check(load_library().astro_ctx_dynamic_init(self.h, float(beta), float(gamma), float(dt)))

    def dynamic_step(self, load_scale: float):
        _check(load_library().astro_ctx_dynamic_step(self.h, float(load_scale)))

    def dynamic_run(self, load_scales):
        f = np.ascontiguousarray(load_scales, dtype=np.float64)
        _check(load_library().astro_ctx_dynamic_run(self.h, int(f.size), f))

    def set_design(self, el_dv, el_coef=None, dv_layer: int = 0, h: float = 1e-30):
        """Design variables for complex-step sensitivities (shell4 layer thickness)."""
        dv = np.ascontiguousarray(el_dv, dtype=np.int32)
        self._coef = None if el_coef is None else np.ascontiguousarray(el_coef, dtype=np.float64)
        self.n_dv = int(dv.max()) + 1
        _check(load_library().astro_ctx_set_design(self.h, self.n_dv, dv,
                                                   None if self._coef is None else self._coef.ctypes.data,
                                                   int(dv_layer), float(h)))

    def dRdD(self):
        n = self.scalar("elMatDim")
        out = np.zeros(n * self.n_dv)
        _check(load_library().astro_ctx_dRdD(self.h, out))
        return out.reshape(self.n_dv, n).T

    def compliance_gradient(self):
        g = np.zeros(self.n_dv)
        _check(load_library().astro_ctx_compliance_gradient(self.h, g))
        return g

    def set_design_vars(self, dv: "DesignVariables", h: float = 1e-30):
        """Design variables of the brick8 path (material constants, density,
        conductivity, node coordinates, node loads); derivatives by complex
        step h at the model's current properties."""
        keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        D = _Design()
        D.numDV = dv.num
        D.category = ptr([DV_CATEGORY[c] if isinstance(c, str) else int(c) for c in dv.category], np.int32)
        D.component = ptr(dv.component, np.int32)
        D.elToDRange = ptr(dv.elToDRange, np.int32)
        D.elToD = ptr(dv.elToD, np.int32)
        D.elToCoef = ptr(dv.elToCoef, np.float64)
        D.ndToDRange = ptr(dv.ndToDRange, np.int32)
        D.ndToD = ptr(dv.ndToD, np.int32)
        D.ndToCoef = ptr(dv.ndToCoef, np.float64)
        D.h = float(h)
        _check(load_library().astro_ctx_set_design_vars(self.h, C.byref(D)))
        self.n_dv = dv.num

    def dRdD_vars(self):
        """dR/dd (elMatDim x numDV) of the linear static residual at nodeDisp."""
        n = self.scalar("elMatDim")
        out = np.zeros(n * self.n_dv)
        _check(load_library().astro_ctx_dRdD_vars(self.h, out))
        return out.reshape(self.n_dv, n).T

    def dRtdD_vars(self):
        """dRt/dd (numNodes x numDV, rank order) of the thermal residual at nodeTemp."""
        n = self.model.num_nodes
        out = np.zeros(n * self.n_dv)
        _check(load_library().astro_ctx_dRtdD_vars(self.h, out))
        return out.reshape(self.n_dv, n).T

    def objective(self, terms):
        """Value of each ObjectiveTerm at the current state."""
        arr, keep = _marshal_terms(terms)
        out = np.zeros(len(terms))
        _check(load_library().astro_ctx_objective(self.h, len(terms), C.cast(arr, C.c_void_p), out))
        return out

    def objective_gradient(self, terms):
        """dL/dd of the summed terms by the adjoint (after static_solve and,
        for temperature terms, thermal_solve)."""
        arr, keep = _marshal_terms(terms)
        out = np.zeros(self.n_dv)
        _check(load_library().astro_ctx_objective_gradient(self.h, len(terms), C.cast(arr, C.c_void_p), out))
        return out

    def dynamic_objective_gradient(self, terms, U, A, load_scale=None, windows=None):
        """(J, dJ/dd) of a Newmark objective, J = sum over steps 1..nSteps of
        the terms active at i dt, from the forward's states U, A ((nSteps + 1)
        x elMatDim, row 0 the start).  windows: (numTerms, 2) activeTime
        windows or None."""
        U = np.ascontiguousarray(U, np.float64)
        A = np.ascontiguousarray(A, np.float64)
        ns = U.shape[0] - 1
        if A.shape != U.shape or ns < 1:
            raise ValueError("U and A must hold nSteps + 1 rows of elMatDim values")
        ls = None if load_scale is None else np.ascontiguousarray(load_scale, np.float64)
        if ls is not None and ls.size != ns:
            raise ValueError("load_scale must hold nSteps values")
        w = None if windows is None else np.ascontiguousarray(windows, np.float64).reshape(-1)
        arr, keep = _marshal_terms(terms)
        out = np.zeros(self.n_dv)
        J = np.zeros(1)
        _check(load_library().astro_ctx_dynamic_objective_gradient(
            self.h, ns, U.reshape(-1), A.reshape(-1), None if ls is None else ls.ctypes.data,
            None if w is None else w.ctypes.data, len(terms), C.cast(arr, C.c_void_p), J, out))
        return float(J[0]), out

    def thermal_dynamic_init(self, gamma: float = 0.5, dt: float = 1.0):
        """Transient conduction set-up: tangent K + C/(gamma dt), penalty
        scaling, profile, factor; T = Tdot = 0."""
        _check(load_library().astro_ctx_thermal_dynamic_init(self.h, float(gamma), float(dt)))

    def thermal_dynamic_run(self, n_steps: int):
        _check(load_library().astro_ctx_thermal_dynamic_run(self.h, int(n_steps)))

    def element_stress(self, point=(0.0, 0.0, 0.0)):
        """Stress and strain (numEls x 6, Voigt 11, 22, 33, 12, 13, 23) of every
        brick8 at the natural point for the current nodeDisp."""
        E = self.model.num_els
        st, sn = np.zeros(6 * E), np.zeros(6 * E)
        _check(load_library().astro_ctx_element_stress(self.h, np.ascontiguousarray(point, np.float64), st, sn))
        return st.reshape(E, 6), sn.reshape(E, 6)

    def element_flux(self, point=(0.0, 0.0, 0.0)):
        """Heat flux and temperature gradient (numEls x 3 each) of every brick8
        at the natural point for the current nodeTemp."""
        E = self.model.num_els
        fl, gr = np.zeros(3 * E), np.zeros(3 * E)
        _check(load_library().astro_ctx_element_flux(self.h, np.ascontiguousarray(point, np.float64), fl, gr))
        return fl.reshape(E, 3), gr.reshape(E, 3)

    def nonlinear_solve(self, max_it=100, tol=1e-10, int_iters=16):
        """Newton static solve with the internal force at u and its tangent
        (co-rotational shell4, Green-Lagrange brick8); returns the steps
        taken."""
        it = C.c_int32(0)
        _check(load_library().astro_ctx_nonlinear_solve(self.h, int(max_it), float(tol), int(int_iters),
                                                         C.byref(it)))
        return it.value

    def dRdT_product(self, lam):
        """(dR/dT)^T lam (numNodes, by rank) for a displacement-sized lam of a
        brick8 model with thermal expansion."""
        lam = np.ascontiguousarray(lam, np.float64)
        out = np.zeros(self.model.num_nodes)
        _check(load_library().astro_ctx_dRdT_product(self.h, lam, out))
        return out

    def nonlinear_tangent(self, u=None, iters=16):
        """Assemble the tangent at the displacement u (elMatDim; None: the
        current solution) into the CSR values (``getd("elasticMat")``):
        co-rotational for shell4, Green-Lagrange for brick8."""
        up = None
        if u is not None:
            u = np.ascontiguousarray(u, np.float64)
            if u.shape != (self.scalar("elMatDim"),):
                raise ValueError("u must hold elMatDim values")
            up = u.ctypes.data
        _check(load_library().astro_ctx_nonlinear_tangent(self.h, up, int(iters)))

    def shell_internal_force(self, u=None, iters=16):
        """Co-rotational internal force (elMatDim) and per-element strain
        energy (numEls) of a shell4 model at the displacement u (elMatDim;
        None: the current solution)."""
        n = self.scalar("elMatDim")
        E = self.model.num_els
        f, en = np.zeros(n), np.zeros(E)
        up = None
        if u is not None:
            u = np.ascontiguousarray(u, np.float64)
            if u.shape != (n,):
                raise ValueError(f"u must hold {n} values")
            up = u.ctypes.data
        _check(load_library().astro_ctx_shell_internal_force(self.h, up, int(iters), f, en))
        return f, en

    def cs_batch_gradient(self):
        """The compliance gradient by direct complex-step solves of all design
        variables in one launch; returns (gradient, per-instance compliance)."""
        g = np.zeros(self.n_dv)
        cm = np.zeros(self.n_dv)
        _check(load_library().astro_ctx_cs_batch_gradient(self.h, g, cm.ctypes.data))
        return g, cm

    def cs_batch_autotune(self):
        """Times every tile shape of the batched kernel; keeps the fastest.
        Returns {RB * 100 + CB: ms} (ms < 0: does not fit)."""
        t = np.zeros(8, np.int32)
        ms = np.zeros(8, np.float32)
        _check(load_library().astro_ctx_cs_batch_autotune(self.h, t.ctypes.data, ms.ctypes.data, 8))
        return {int(a): float(b) for a, b in zip(t, ms) if a > 0}

    def cs_batch_tile(self, tile: int):
        """Selects the batched kernel's tile RB * 100 + CB."""
        _check(load_library().astro_ctx_cs_batch_tile(self.h, int(tile)))

    def assemble_thermal(self):
        _check(load_library().astro_ctx_assemble_thermal(self.h))

    def thermal_solve(self):
        _check(load_library().astro_ctx_thermal_solve(self.h))

    def sync(self):
        _check(load_library().astro_ctx_sync(self.h))

    def true_residual(self):
        r = C.c_double(0)
        _check(load_library().astro_ctx_true_residual(self.h, C.byref(r)))
        return r.value

    def operator_apply(self, d_x: int, d_y: int):
        """y = (A + C^T C) x on device vectors at addresses d_x, d_y: the
        operator GMRES applies (node-block form for brick8 models)."""
        _check(load_library().astro_ctx_operator_apply(self.h, C.c_void_p(d_x), C.c_void_p(d_y)))

    def precond_apply(self, d_b: int, d_x: int):
        """x = P^-1 b on device vectors at addresses d_b, d_x (distinct)."""
        _check(load_library().astro_ctx_precond_apply(self.h, C.c_void_p(d_b), C.c_void_p(d_x)))

    def true_residual_noalloc(self, d_u: int, d_r: int):
        """|b - (A + C^T C) u| / |b| for a device displacement at address d_u,
        writing the residual to the device buffer at d_r; allocates nothing
        on the device (see include/astro_abi.h)."""
        r = C.c_double(0)
        _check(load_library().astro_ctx_true_residual_noalloc(self.h, C.c_void_p(d_u), C.c_void_p(d_r), C.byref(r)))
        return r.value

    # ---- learned block preconditioner
    def gnn_collect(self, gnn: "Gnn", n_rhs: int, seed: int = 0, rtol: float = 1e-10, pairs: str = "solution"):
        """Add one training field: n_rhs pairs of the current (factored)
        operator, "solution" (x random, b = A x) or "solve" (b random, x by
        GMRES to rtol), with the residual its LDL^T stage leaves."""
        kinds = {"solution": 0, "solve": 1}
        if pairs not in kinds:
            raise ValueError("pairs must be 'solution' or 'solve'")
        _check(load_library().astro_ctx_gnn_collect(self.h, gnn.h, int(n_rhs), int(seed), float(rtol),
                                                    kinds[pairs]))

    def gnn_attach(self, gnn: "Gnn"):
        """Blocks of the network for this operator; GMRES then runs with the
        two-stage preconditioner, the plain LDL^T as judge and fallback."""
        _check(load_library().astro_ctx_gnn_attach(self.h, gnn.h))

    def gnn_detach(self):
        _check(load_library().astro_ctx_gnn_detach(self.h))

    def gnn_get(self, name):
        N = self.model.num_nodes
        out = np.zeros(N * 9)
        _check(load_library().astro_ctx_gnn_get(self.h, name.encode(), out, out.size))
        return out.reshape(N, 9)

    def set_timing(self, on: bool):
        _check(load_library().astro_ctx_set_timing(self.h, 1 if on else 0))

    def kernel_stat(self, name):
        return float(load_library().astro_ctx_kernel_stat(self.h, name.encode()))

    def phase_ms(self, phase):
        return float(load_library().astro_ctx_phase_ms(self.h, phase.encode()))


class Gnn:
    """The learned block preconditioner's graph network (FP32 weights on the
    device): solver-generated training fields, Adam training with the
    library's HIP kernels, parameters as a flat float32 vector."""

    def __init__(self, hidden: int = 64, layers: int = 3, seed: int = 0):
        h = C.c_void_p()
        _check(load_library().astro_gnn_create(int(hidden), int(layers), int(seed), C.byref(h)))
        self.h = h
        self.hidden, self.layers = hidden, layers

    def close(self):
        if getattr(self, "h", None):
            load_library().astro_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return int(load_library().astro_gnn_num_params(self.h))

    @property
    def num_fields(self):
        return int(load_library().astro_gnn_num_fields(self.h))

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(load_library().astro_gnn_get_params(self.h, p, p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(load_library().astro_gnn_set_params(self.h, p, p.size))

    def train(self, epochs: int, lr: float = 1e-3, seed: int = 0):
        """Adam over the collected fields; returns the mean field loss per epoch."""
        out = np.zeros(max(1, epochs))
        _check(load_library().astro_gnn_train(self.h, int(epochs), float(lr), int(seed), out.ctypes.data))
        return out[:epochs]

    def loss_grad(self, field: int = 0):
        loss = C.c_double(0)
        g = np.zeros(self.num_params, np.float32)
        _check(load_library().astro_gnn_loss_grad(self.h, int(field), C.byref(loss), g.ctypes.data))
        return loss.value, g

    def _size(self, field):
        N, na, nd, ns = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        _check(load_library().astro_gnn_field_size(self.h, int(field), C.byref(N), C.byref(na), C.byref(nd),
                                                   C.byref(ns)))
        return N.value, na.value, nd.value, ns.value

    def field(self, field: int = 0):
        """range, adj, X, G, R1 (samples x DOFs) and eTr of a collected field."""
        N, na, nd, ns = self._size(field)
        rng = np.zeros(N + 1, np.int32)
        adj = np.zeros(max(1, na), np.int32)
        X = np.zeros(N * 9, np.float32)
        G, R1 = np.zeros(N * 9), np.zeros(max(1, nd * ns))
        eTr = C.c_double(0)
        _check(load_library().astro_gnn_get_field(self.h, int(field), rng, adj, X, G, R1, C.byref(eTr)))
        return dict(range=rng, adj=adj[:na], X=X.reshape(N, 9), G=G.reshape(N, 3, 3),
                    R1=R1[:nd * ns].reshape(ns, nd), eTr=eTr.value)

    def field_blocks(self, field: int = 0):
        N, _, _, _ = self._size(field)
        B = np.zeros(N * 9)
        _check(load_library().astro_gnn_field_blocks(self.h, int(field), B))
        return B.reshape(N, 3, 3)
