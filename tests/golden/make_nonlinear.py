"""Writes tests/golden/ref_nonlinear_*.npz from the literal reference.

Runs oracle/_ref/astro_ref_driver in its 'nonlinear' mode: the reference's
own Fortran modules (flang-compiled from /root/reference) scale the penalty
rows from the tangent at rest and run stepSolve's elastic loop with
nonLinearGeom (Working/AStrO_commandFunctions.f90:704-747): at most 20 Newton
steps, each assembling the tangent and the Green-Lagrange residual at u
(getElasticSolnLoad(1)), factoring and solving.  Stored: the load, the
displacement the reference ends with, its step count, its linear solution
and its DOF numbering.  The model is rebuilt from tests/nonlinear_cases.py.

The reference's Newton loop does not converge: every element gathers the
solution of nodes 1..8 instead of its own nodes (r_getElementSolution, D4),
here the clamped nodes, so the residual never sees u and each of the 20
steps adds the linear solution again.  The fixture records that.

    python tests/golden/make_nonlinear.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from nonlinear_cases import CASES, column_model  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    for name, P in CASES:
        m = column_model(P)
        ref = O.run_reference(m, "nonlinear")
        lin = O.run_reference(m, "elastic")
        out = dict(load=np.array(P), nodeDisp=ref["nodeDisp"], nlSteps=ref["nlSteps"], linearDisp=lin["nodeDisp"],
                   currentRank=ref["currentRank"], nDofIndex=ref["nDofIndex"])
        fn = os.path.join(HERE, "ref_nonlinear_%s.npz" % name)
        np.savez_compressed(fn, **out)
        u, ul = ref["nodeDisp"], lin["nodeDisp"]
        print(fn, "steps", ref["nlSteps"], "max|u|", np.abs(u).max(), "max|u - u_lin|", np.abs(u - ul).max())


if __name__ == "__main__":
    main()
