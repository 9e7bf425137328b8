"""Drop-in: the reference's own Fortran host (readModelInput, analysisPrep,
scaleElasticMPC, convertToLTri) with its solver module replaced by
misc_amd/fortran/astro_solvers_hip.f90 and its element assembly
(getElasticSolnLoad) by misc_amd/fortran/astro_elastic_hip.f90, both calling
libastro_hip.so through ISO_C_BINDING.  In the same process the host's own
getElasticSolnLoad runs on the same arrays for comparison; the factor and the
solution are compared with the oracle."""

import os

import numpy as np
import pytest

from helpers import csr_sparse, entry_rel, golden, rel_norm
from misc_amd.meshes import cantilever_block
from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(O.DROPIN_DRIVER), reason="drop-in driver not built here")]


@pytest.mark.parametrize("dims", [(6, 5, 3), (12, 10, 6)])
def test_reference_host_with_gpu_solvers(dims):
    m = cantilever_block(*dims)
    print("drop-in", dims, flush=True)
    # the relative stop, so that the result is the converged solution
    r = O.run_reference(m, "elastic", driver=O.DROPIN_DRIVER, env={"ASTRO_GMRES_RTOL": "1e-12"})
    om = O.OracleModel(m).prep().static_elastic(stop_after="profile")
    # GPU assembly against the host's own routine on the same arrays
    np.testing.assert_array_equal(r["elMatCols"], r["elMatCols_host"])
    Kh = r["elasticMat_host"]
    e_gpu = entry_rel(r["elasticMat"], Kh, floor=1e-3 * np.abs(Kh).max())
    # the solution load at the converged u: -K u with the host's own matrix
    # (the host's getElasticSolnLoad(0) gathers element displacements by
    # local node number, so its own dump is not that product)
    n = len(r["nodeDisp"])
    K = csr_sparse(Kh, r["elMatCols_host"], r["elMatRange"], n)
    Rk = -(K @ r["nodeDisp"])
    e_load = np.max(np.abs(r["solnLoad"] - Rk)) / np.max(np.abs(Rk))
    print(f"GPU vs host assembly: csr {e_gpu:.3e}, solution load vs -K u {e_load:.3e} "
          f"(host routine's own: max {np.abs(r['solnLoad_host']).max():.3e} vs {np.abs(Rk).max():.3e})", flush=True)
    assert e_gpu < 1e-12 and e_load < 1e-12
    np.testing.assert_array_equal(r["elMatCols"], om.geti("elMatCols"))
    Kr = om.getd("elasticMat")
    e_csr = entry_rel(r["elasticMat"], Kr, floor=1e-3 * np.abs(Kr).max())
    e_prof = entry_rel(r["elMatLT"], om.getd("elMatLT"), floor=1e-3 * np.abs(Kr).max())
    O.lib().orc_factor_elastic(om.h)
    e_fact = entry_rel(r["elMatLTfact"], om.getd("elMatLT"), floor=1e-6)
    print(f"csr {e_csr:.3e} profile {e_prof:.3e} factor {e_fact:.3e}", flush=True)
    assert e_csr < 1e-12 and e_prof < 1e-12
    assert e_fact < 1e-8
    om2 = O.OracleModel(m).prep().static_elastic(mode=1, rtol=1e-13)
    assert rel_norm(r["nodeDisp"], om2.getd("nodeDisp")) < 1e-10


GMRES_PIN = golden("gmres")


@pytest.mark.parametrize("dims,ref", GMRES_PIN, ids=[str(d) for d, _ in GMRES_PIN])
def test_reference_host_reproduces_reference_displacements(dims, ref):
    """The same reference host and inputs, once with its own Fortran solvers
    (the committed fixture) and once with the GPU solver module, both under
    the reference's own GMRES rule (absolute 1e-12, maxIt / numVecs + 1
    restart cycles) stopped before convergence: the same nodeDisp."""
    nv, it, bw = (int(v) for v in ref["gmres"])
    print("drop-in gmres pin", dims, flush=True)
    r = O.run_reference(cantilever_block(*dims), "elastic", driver=O.DROPIN_DRIVER, gmres=(nv, it, bw))
    assert rel_norm(r["nodeDisp"], ref["nodeDisp"]) <= 1e-10
