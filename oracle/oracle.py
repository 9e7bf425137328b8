"""ctypes front end of the CPU oracle (oracle/liborc.so) and of the literal
reference driver (oracle/_ref/astro_ref_driver).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package misc_amd.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_DRIVER = os.path.join(HERE, "_ref", "astro_ref_driver")
# the reference host with its solver module replaced by the ISO_C_BINDING
# module of misc_amd/fortran (links misc_amd/libastro_hip.so): the drop-in
DROPIN_DRIVER = os.path.join(HERE, "_ref", "hip", "astro_ref_driver_hip")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False) -> None:
    """Compile liborc.so with oracle/Makefile (g++ only)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "liborc"], cwd=HERE, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_set_properties.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_mpc.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _ip, _ip, _ip, C.c_int, _dp]
        L.orc_set_nodal_loads.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.orc_analysis_prep.argtypes = [C.c_void_p, C.c_int]
        L.orc_assemble_elastic.argtypes = [C.c_void_p, C.c_double]
        L.orc_assemble_thermal.argtypes = [C.c_void_p, C.c_int]
        for f in ("orc_scale_elastic_mpc", "orc_scale_thermal_mpc", "orc_build_elastic_profile",
                  "orc_build_thermal_profile", "orc_elastic_rhs", "orc_thermal_rhs"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_factor_elastic.argtypes = [C.c_void_p]
        L.orc_factor_thermal.argtypes = [C.c_void_p]
        L.orc_solve_elastic.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.orc_solve_thermal.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.orc_gmres_mode.restype = C.c_int
        L.orc_gmres_mode.argtypes = [_dp, _dp, _ip, _ip, C.c_int, _dp, _ip, _ip, C.c_int, _dp, _ip, _dp, C.c_int,
                                     C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_geti.restype = C.POINTER(C.c_int)
        L.orc_geti.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_longlong)]
        L.orc_getd.restype = C.POINTER(C.c_double)
        L.orc_getd.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_longlong)]
        L.orc_geti_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_getd_scalar.restype = C.c_double
        L.orc_getd_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_solver_max_bw.argtypes = [C.c_int]
        L.orc_material_stiffness.argtypes = [_dp, _dp]
        L.orc_brick8_stiffness.argtypes = [_dp, _dp, _dp]
        L.orc_brick8_mass.argtypes = [_dp, C.c_double, _dp]
        L.orc_brick8_internal_force.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_brick8_stress.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_element_stress.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_shell_frame.argtypes = [_dp, _dp, _dp]
        L.orc_shell_abd.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.orc_shell4_stiffness.argtypes = [_dp, _dp, _dp, C.c_void_p, _dp]
        L.orc_set_shell_sections.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_set_coupled_profile.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_design.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int]
        L.orc_shell4_dstiffness_cs.argtypes = [_dp, _dp, C.c_int, _dp, C.c_double, C.c_int, C.c_double, C.c_void_p, _dp]
        L.orc_dRdD.argtypes = [C.c_void_p, C.c_double, _dp]
        L.orc_compliance_gradient.argtypes = [C.c_void_p, C.c_double, _dp]
        L.orc_dyn_init.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_dyn_residual_load.argtypes = [C.c_void_p, C.c_double]
        L.orc_dyn_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double]
        L.orc_brick8_conductivity.argtypes = [_dp, _dp, _dp]
        L.orc_brick8_capacitance.argtypes = [_dp, C.c_double, _dp]
        L.orc_set_spec_heat.argtypes = [C.c_void_p, _dp]
        L.orc_thermal_dyn_init.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.orc_thermal_dyn_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.orc_brick8_volume.restype = C.c_double
        L.orc_brick8_volume.argtypes = [_dp]
        L.orc_convert_ltri.argtypes = [_dp, _ip, C.c_longlong, _dp, _ip, _ip, C.c_int, _dp, _ip, _ip, C.c_int]
        L.orc_ldl_factor.restype = C.c_int
        L.orc_ldl_factor.argtypes = [_dp, C.c_longlong, _ip, C.c_int]
        L.orc_ldl_solve.argtypes = [_dp, _dp, _dp, _ip, C.c_int, _dp]
        L.orc_gmres.restype = C.c_int
        L.orc_gmres.argtypes = [_dp, _dp, _ip, _ip, C.c_int, _dp, _ip, _ip, C.c_int, _dp, _ip, _dp, C.c_int,
                                C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_apply_operator.argtypes = [_dp, _dp, _dp, _ip, _ip, C.c_int, _dp, _ip, _ip, C.c_int]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------------------
# Raw kernels
def material_stiffness(eprops9):
    C6 = np.zeros(36)
    lib().orc_material_stiffness(_f64(eprops9), C6)
    return C6.reshape(6, 6, order="F")


def brick8_stiffness(xyz38, Cmat):
    Ke = np.zeros(576)
    lib().orc_brick8_stiffness(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")),
                               _f64(np.asarray(Cmat).ravel(order="F")), Ke)
    return Ke.reshape(24, 24, order="F")


def brick8_mass(xyz38, rho):
    Me = np.zeros(576)
    lib().orc_brick8_mass(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")), float(rho), Me)
    return Me.reshape(24, 24, order="F")


def brick8_internal_force(xyz38, Cmat, ue24):
    """Green-Lagrange internal force of one brick8 at element DOFs ue
    (component-major)."""
    R = np.zeros(24)
    lib().orc_brick8_internal_force(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")),
                                    _f64(np.asarray(Cmat).ravel(order="F")), _f64(ue24), R)
    return R


def brick8_stress(xyz38, Cmat, ue24, s3):
    """(stress, strain) of one brick8 at the natural point s3 for element DOFs
    ue (component-major), Voigt (11, 22, 33, 12, 13, 23)."""
    st, sn = np.zeros(6), np.zeros(6)
    lib().orc_brick8_stress(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")),
                            _f64(np.asarray(Cmat).ravel(order="F")), _f64(ue24), _f64(s3), st, sn)
    return st, sn


def shell_frame(orient6, xyz34):
    """Rows = local axes of a shell4 element (section orientation aligned to the
    element normal)."""
    a = np.zeros(9)
    lib().orc_shell_frame(_f64(orient6), _f64(np.asarray(xyz34).reshape(3, 4).ravel(order="F")), a)
    return a.reshape(3, 3, order="F")


def shell_abd(layers, z_off=0.0):
    """9x9 ABD of a layup; layers: (nL, 6) E1, E2, nu12, G12, thickness, angle."""
    L = np.asarray(layers, np.float64).reshape(-1, 6)
    A = np.zeros(81)
    lib().orc_shell_abd(L.shape[0], _f64(L.ravel()), float(z_off), A)
    return A.reshape(9, 9, order="F")


def shell4_stiffness(xyz34, alpha33, abd99):
    """(K32 before condensation, Kc 24x24 condensed) of one shell4."""
    K32 = np.zeros(32 * 32)
    Kc = np.zeros(24 * 24)
    lib().orc_shell4_stiffness(_f64(np.asarray(xyz34).reshape(3, 4).ravel(order="F")),
                               _f64(np.asarray(alpha33).ravel(order="F")), _f64(np.asarray(abd99).ravel(order="F")),
                               K32.ctypes.data, Kc)
    return K32.reshape(32, 32, order="F"), Kc.reshape(24, 24, order="F")


def shell4_dstiffness_cs(xyz34, alpha33, layers, layer, h=1e-30, z_off=0.0):
    """Complex-step d/dt (thickness of `layer`) of a shell4: (dK32, dKc)."""
    L = np.asarray(layers, np.float64).reshape(-1, 6)
    K32 = np.zeros(32 * 32)
    Kc = np.zeros(24 * 24)
    lib().orc_shell4_dstiffness_cs(_f64(np.asarray(xyz34).reshape(3, 4).ravel(order="F")),
                                   _f64(np.asarray(alpha33).ravel(order="F")), L.shape[0], _f64(L.ravel()),
                                   float(z_off), int(layer), float(h), K32.ctypes.data, Kc)
    return K32.reshape(32, 32, order="F"), Kc.reshape(24, 24, order="F")


def brick8_conductivity(xyz38, k33):
    Kt = np.zeros(64)
    lib().orc_brick8_conductivity(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")),
                                  _f64(np.asarray(k33).ravel(order="F")), Kt)
    return Kt.reshape(8, 8, order="F")


def brick8_volume(xyz38):
    return lib().orc_brick8_volume(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")))


def ldl_factor(A, ARange, n):
    A = _f64(A).copy()
    bad = lib().orc_ldl_factor(A, A.size, _i32(ARange), int(n))
    return A, bad


def ldl_solve(L, LRange, rhs):
    n = len(rhs)
    x = np.zeros(n)
    z = np.zeros(n)
    lib().orc_ldl_solve(x, _f64(rhs), _f64(L), _i32(LRange), n, z)
    return x


def convert_ltri(lt_range, mat, cols, rng, n, cmat, ccols, crange, cdim):
    lt = np.zeros(int(lt_range[-1]))
    cm = _f64(cmat) if cdim > 0 else np.zeros(1)
    cc = _i32(ccols) if cdim > 0 else np.zeros(1, np.int32)
    cr = _i32(crange) if cdim > 0 else np.zeros(1, np.int32)
    lib().orc_convert_ltri(lt, _i32(lt_range), lt.size, _f64(mat), _i32(cols), _i32(rng), int(n), cm, cc, cr,
                           int(cdim))
    return lt


def gmres(x0, A, Acols, ARange, Cm, Ccols, CRange, cdim, P, PRange, b, num_vecs, max_it, mode=0, rtol=0.0):
    x = _f64(x0).copy()
    res = C.c_double(0)
    steps = C.c_int(0)
    cyc = lib().orc_gmres_mode(x, _f64(A), _i32(Acols), _i32(ARange), len(b), _f64(Cm), _i32(Ccols),
                               _i32(CRange), int(cdim), _f64(P), _i32(PRange), _f64(b), int(num_vecs), int(max_it),
                               int(mode), float(rtol), C.byref(res), C.byref(steps))
    return x, cyc, res.value, steps.value


def apply_operator(x, A, Acols, ARange, Cm, Ccols, CRange, cdim):
    y = np.zeros(len(x))
    lib().orc_apply_operator(y, _f64(x), _f64(A), _i32(Acols), _i32(ARange), len(x), _f64(Cm), _i32(Ccols),
                             _i32(CRange), int(cdim))
    return y


def solver_max_bw(num_nodes):
    return lib().orc_solver_max_bw(int(num_nodes))


# ---------------------------------------------------------------------------
class OracleModel:
    """A model loaded into the oracle: the reference's global state for one job."""

    def __init__(self, model):
        L = lib()
        self.model = model
        self._nodes = _f64(model.nodes.reshape(-1))
        self._els = _i32(model.elements.reshape(-1))
        self._et = _i32(model.etype)
        self.h = L.orc_create(model.num_nodes, self._nodes, model.num_els, self._els, self._et)
        ep, den, cond = _f64(model.eprops.reshape(-1)), _f64(model.density), _f64(model.cond.reshape(-1))
        L.orc_set_properties(self.h, ep.ctypes.data, den.ctypes.data, cond.ctypes.data)
        if getattr(model, "spec_heat", None) is not None:
            L.orc_set_spec_heat(self.h, _f64(model.spec_heat))
        if model.mpc:
            eqn = _i32([t.eqn for t in model.mpc])
            dof = _i32([t.dof for t in model.mpc])
            coef = _f64([t.coef for t in model.mpc])
            iss = _i32([1 if t.is_set else 0 for t in model.mpc])
            starts = np.zeros(len(model.mpc) + 1, np.int32)
            starts[1:] = np.cumsum([len(t.nodes) for t in model.mpc])
            nodes = _i32(np.concatenate([t.nodes for t in model.mpc]))
            L.orc_set_mpc(self.h, len(model.mpc), eqn, dof, coef, iss, starts, nodes, len(model.mpc_rhs),
                          _f64(model.mpc_rhs))
        if model.load_nodes.size:
            L.orc_set_nodal_loads(self.h, int(model.load_nodes.size), _i32(model.load_nodes),
                                  _f64(model.load_vals.reshape(-1)))
        if model.layers is not None:
            lay = _f64(model.layers.reshape(-1))
            L.orc_set_shell_sections(self.h, int(model.layers.shape[1]), lay, _f64(model.z_offset),
                                     _f64(model.orient.reshape(-1)))

    def __del__(self):
        try:
            if self.h:
                lib().orc_free(self.h)
        except Exception:
            pass

    def prep(self, max_bw: int = 0, coupled_profile: bool = False):
        lib().orc_set_coupled_profile(self.h, 1 if coupled_profile else 0)
        lib().orc_analysis_prep(self.h, int(max_bw))
        return self

    def geti(self, name):
        n = C.c_longlong(0)
        p = lib().orc_geti(self.h, name.encode(), C.byref(n))
        if n.value < 0:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int32)

    def getd(self, name):
        n = C.c_longlong(0)
        p = lib().orc_getd(self.h, name.encode(), C.byref(n))
        if n.value < 0:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0)

    def setd(self, name, values):
        """Overwrite one of the model's double arrays (same length)."""
        n = C.c_longlong(0)
        p = lib().orc_getd(self.h, name.encode(), C.byref(n))
        v = np.asarray(values, np.float64)
        if n.value != v.size:
            raise ValueError(f"{name}: length {n.value}, got {v.size}")
        np.ctypeslib.as_array(p, shape=(n.value,))[:] = v
        return self

    def scalar(self, name):
        if name == "gmresRes":
            return lib().orc_getd_scalar(self.h, name.encode())
        return lib().orc_geti_scalar(self.h, name.encode())

    def static_elastic(self, num_vecs: int = 0, stop_after: Optional[str] = None, mode: int = 0,
                       rtol: float = 1e-12):
        """solve() + stepSolve() for a linear static elastic job
        (Working/AStrO_commandFunctions.f90:751-782)."""
        L = lib()
        L.orc_assemble_elastic(self.h, 0.0)
        if stop_after == "assemble":
            return self
        L.orc_scale_elastic_mpc(self.h)
        L.orc_build_elastic_profile(self.h)
        if stop_after == "profile":
            return self
        L.orc_factor_elastic(self.h)
        if stop_after == "factor":
            return self
        L.orc_solve_elastic(self.h, int(num_vecs), int(mode), float(rtol))
        return self

    def set_design(self, el_dv, el_coef=None, dv_layer=0):
        el_dv = _i32(el_dv)
        coef = _f64(np.ones(len(el_dv)) if el_coef is None else el_coef)
        self.n_dv = int(el_dv.max()) + 1
        lib().orc_set_design(self.h, self.n_dv, el_dv, coef, int(dv_layer))
        return self

    def dRdD(self, h=1e-30):
        n = self.scalar("elMatDim")
        out = np.zeros(n * self.n_dv)
        lib().orc_dRdD(self.h, float(h), out)
        return out.reshape(self.n_dv, n).T

    def compliance_gradient(self, h=1e-30):
        g = np.zeros(self.n_dv)
        lib().orc_compliance_gradient(self.h, float(h), g)
        return g

    def dynamic_init(self, beta: float = 0.25, gamma: float = 0.5, dt: float = 1e-5):
        """solve()'s dynamic set-up: tangent K + M/(beta dt^2), penalty scaling,
        profile and factor; the state starts at rest."""
        L = lib()
        L.orc_dyn_init(self.h, float(beta), float(gamma), float(dt))
        bad = L.orc_factor_elastic(self.h)
        if bad:
            raise RuntimeError(f"zero pivot at row {bad}")
        return self

    def element_stress(self, s3=(0.0, 0.0, 0.0)):
        """(stress, strain), numEls x 6 each, at the natural point s3 for the
        current nodeDisp."""
        E = self.scalar("numEls")
        st, sn = np.zeros(6 * E), np.zeros(6 * E)
        lib().orc_element_stress(self.h, _f64(s3), st, sn)
        return st.reshape(E, 6), sn.reshape(E, 6)

    def dynamic_step(self, load_scale: float, num_vecs: int = 0, mode: int = 1, rtol: float = 1e-12):
        lib().orc_dyn_step(self.h, float(load_scale), int(num_vecs), int(mode), float(rtol))
        return self

    def static_thermal(self, literal_d8: bool = True, num_vecs: int = 0, mode: int = 0, rtol: float = 1e-12):
        L = lib()
        L.orc_assemble_thermal(self.h, 1 if literal_d8 else 0)
        L.orc_scale_thermal_mpc(self.h)
        L.orc_build_thermal_profile(self.h)
        L.orc_factor_thermal(self.h)
        L.orc_solve_thermal(self.h, int(num_vecs), int(mode), float(rtol))
        return self


    def thermal_dynamic_init(self, gamma: float = 0.5, dt: float = 1.0, literal_d8: bool = True):
        """The transient thermal set-up: K + C/(gamma dt) into the thermal CSR,
        the capacitance C on the same columns (capMat), penalty scaling,
        profile and factor; T = Tdot = 0."""
        lib().orc_thermal_dyn_init(self.h, float(gamma), float(dt), 1 if literal_d8 else 0)
        return self

    def thermal_dynamic_step(self, num_vecs: int = 0, mode: int = 1, rtol: float = 1e-12):
        lib().orc_thermal_dyn_step(self.h, int(num_vecs), int(mode), float(rtol))
        return self


def brick8_capacitance(xyz38, rhoc):
    M = np.zeros(64)
    lib().orc_brick8_capacitance(_f64(np.asarray(xyz38).reshape(3, 8).ravel(order="F")), float(rhoc), M)
    return M.reshape(8, 8, order="F")


# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def run_reference(model, mode: str = "elastic", workdir: Optional[str] = None,
                  driver: Optional[str] = None, gmres: Optional[tuple] = None,
                  env: Optional[Dict[str, str]] = None, design=None, objective: Optional[str] = None,
                  write_results: bool = False, temperatures=None) -> Dict[str, np.ndarray]:
    """Write `model` in the reference's format, run the literal reference driver
    on it and return every dumped array (1-based values as the reference holds
    them).  gmres = (numVecs, maxIt, solverMaxBW) overrides the restart length,
    the iteration cap and the profile bandwidth of the job (0 keeps the
    reference's rule).  design: design variables for
    misc_amd.refformat.write_design_vars, read by the driver through
    ASTRO_DV_FILE; objective: the text of an objective file (ASTRO_OBJ_FILE);
    write_results: run the reference's writers after the solve (their files
    come back as text under the keys 'nodes.yaml', 'elements.yaml',
    'objective.yaml', 'dvars.yaml'); temperatures: nodal temperatures by rank
    for mode 'thermoelastic' (ASTRO_TEMP_FILE: the elastic step with these
    temperatures in its residual)."""
    from misc_amd.refformat import write_model

    exe = driver or REF_DRIVER
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} not built (make -C oracle ref / ref-hip)")
    own = workdir is None
    wd = workdir or tempfile.mkdtemp(prefix="astro_ref_")
    mf = os.path.join(wd, "model.yaml")
    write_model(model, mf)
    # The reference reads heap memory it never initialised (the (0) entry of
    # elMPCMatRange / thermMPCMatRange, Working/AStrO_bookKeeping.f90:516-529,
    # 656-674, read during buildElasticMPC itself) and works only where fresh
    # allocations come back zero-filled.  glibc's MALLOC_PERTURB_=255 zero-fills
    # every allocation served by the main allocator; its per-thread cache
    # bypasses that, so the cache is switched off as well.  (This matters once
    # another runtime, e.g. HIP in the drop-in build, has used the heap.)
    tun = os.environ.get("GLIBC_TUNABLES", "")
    tun = (tun + ":" if tun else "") + "glibc.malloc.tcache_count=0"
    env = dict(os.environ, MALLOC_PERTURB_="255", GLIBC_TUNABLES=tun, **(env or {}))
    if design:
        from misc_amd.refformat import write_design_vars

        df = os.path.join(wd, "design.yaml")
        write_design_vars(df, design)
        env["ASTRO_DV_FILE"] = df
    if objective:
        of = os.path.join(wd, "objective_in.yaml")
        with open(of, "w") as f:
            f.write(objective)
        env["ASTRO_OBJ_FILE"] = of
    if write_results:
        env["ASTRO_WRITE_PREFIX"] = os.path.join(wd, "w_")
    if temperatures is not None:
        tf = os.path.join(wd, "temps.f64")
        np.ascontiguousarray(temperatures, "<f8").tofile(tf)
        env["ASTRO_TEMP_FILE"] = tf
    extra = [str(int(v)) for v in gmres] if gmres else []
    subprocess.run([exe, mf, os.path.join(wd, "o_"), mode] + extra, cwd=wd, check=True, env=env)
    out = {}
    for fn in os.listdir(wd):
        if fn.startswith("w_"):
            with open(os.path.join(wd, fn)) as f:
                out[fn[2:]] = f.read()
            continue
        if fn.startswith("SolutionHistory"):
            with open(os.path.join(wd, fn), "rb") as f:
                out[fn] = f.read()
            continue
        if not fn.startswith("o_"):
            continue
        key, ext = os.path.splitext(fn[2:])
        if ext == ".i32":
            out[key] = np.fromfile(os.path.join(wd, fn), dtype="<i4")
        elif ext == ".f64":
            out[key] = np.fromfile(os.path.join(wd, fn), dtype="<f8")
        elif ext == ".txt" and key == "times":
            t = {}
            with open(os.path.join(wd, fn)) as f:
                for ln in f:
                    a, b = ln.split()
                    t[a] = float(b)
            out["times"] = t
    if own:
        import shutil
        shutil.rmtree(wd, ignore_errors=True)
    return out
