"""Writes tests/golden/ref_writers_{elastic,thermal}.npz: the reference's own
writers (writeNodeResults, writeElementResults, writeObjective after
getObjective, writeDVarValues, and writeBinarySolution for a state of
distinct values; Working/AStrO_output.f90) run by the
instrumented driver (oracle/ref_driver.f90, ASTRO_WRITE_PREFIX) after a solve,
with the input files and the arrays they printed.

Run where the reference sources are (oracle/Makefile ref); the tests read the
committed fixtures only.
"""

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from misc_amd.meshes import cantilever_block, thermal_block  # noqa: E402
from misc_amd.refformat import write_design_vars, write_model  # noqa: E402
from oracle import oracle as O  # noqa: E402

OBJECTIVE = """objectiveTerms:
    - category: displacement
      operator: powerNorm
      component: 3
      coefficient: 2.0
      exponent: 1.0
      nodeSet: topn
    - category: displacement
      operator: powerNorm
      component: 1
      coefficient: 3.0e9
      exponent: 2.0
      targetValue: 1.0e-7
      nodeSet: topn
    - category: volume
      operator: volumeIntegral
      coefficient: 1.0
      exponent: 1.0
      elementSet: allel
    - category: mass
      operator: volumeIntegral
      coefficient: 1.0e-3
      exponent: 2.0
      targetValue: 1.0
      elementSet: allel
"""


def elastic_case():
    m = cantilever_block(2, 2, 2, e_field=np.random.default_rng(4).uniform(150e9, 250e9, 8))
    m.density = np.random.default_rng(5).uniform(2000, 3000, 8)
    m.element_sets = {"allel": np.arange(1, 9)}
    m.node_sets["topn"] = np.arange(19, 28)
    design = [dict(category="elastic", subCategory="modulus", component=1, elementSet="allel", coefficients=1e9),
              dict(category="nodeLoad", subCategory="elastic", component=3, nodeSet="topn", coefficients=10.0,
                   activeTime=[0.0, 2.0])]
    return m, design


def thermal_case():
    m = thermal_block(2, 2, 2)
    rng = np.random.default_rng(6)
    m.cond[:, 0] = rng.uniform(20, 80, 8)
    m.cond[:, 1] = rng.uniform(20, 80, 8)
    m.cond[:, 2] = rng.uniform(20, 80, 8)
    m.cond[:, 4] = rng.uniform(-5, 5, 8)
    m.element_sets = {"allel": np.arange(1, 9)}
    return m


def texts(m, design=None):
    with tempfile.TemporaryDirectory() as wd:
        write_model(m, os.path.join(wd, "m.yaml"))
        model_txt = open(os.path.join(wd, "m.yaml")).read()
        dv_txt = ""
        if design:
            write_design_vars(os.path.join(wd, "d.yaml"), design)
            dv_txt = open(os.path.join(wd, "d.yaml")).read()
    return model_txt, dv_txt


def main():
    m, design = elastic_case()
    r = O.run_reference(m, "elastic", objective=OBJECTIVE, design=design, write_results=True)
    model_txt, dv_txt = texts(m, design)
    np.savez_compressed(os.path.join(HERE, "ref_writers_elastic.npz"), model=model_txt, design=dv_txt, objective=OBJECTIVE,
             nodes=r["nodes.yaml"], elements=r["elements.yaml"], objective_out=r["objective.yaml"],
             dvars=r["dvars.yaml"], nodeDisp=r["nodeDisp"], elasticLoad=r["elasticLoad"],
             nDofIndex=r["nDofIndex"], currentRank=r["currentRank"], objVal=r["objVal"],
             history3=np.frombuffer(r[[k for k in r if k.startswith("SolutionHistory")][0]], np.uint8),
             history3_name=[k for k in r if k.startswith("SolutionHistory")][0])
    mt = thermal_case()
    r = O.run_reference(mt, "thermal", write_results=True)
    model_txt, _ = texts(mt)
    np.savez_compressed(os.path.join(HERE, "ref_writers_thermal.npz"), model=model_txt, nodes=r["nodes.yaml"],
             elements=r["elements.yaml"], nodeTemp=r["nodeTemp"], thermalLoad=r["thermalLoadSoln"],
             currentRank=r["currentRank"])


if __name__ == "__main__":
    main()
