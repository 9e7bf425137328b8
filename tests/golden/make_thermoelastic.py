"""Writes tests/golden/ref_thermoelastic_*.npz from the literal reference.

Runs oracle/_ref/astro_ref_driver in its 'thermoelastic' mode: the
reference's own Fortran modules (flang-compiled from /root/reference) assemble
and factor the elastic tangent as solve() does, take the nodal temperatures
from a file (the temperatures a thermal step would leave; the reference's own
thermal step drops a coupling per row, D8), and run stepSolve's elastic step,
whose residual holds the thermal strain (stress = C (strain - T alpha),
Working/AStrO_r_elementEqns.f90:2433-2447).  Stored: the model parameters, the
temperatures, and the right-hand side the reference built (elasticLoad =
F - R(0, T)).  Only data is stored; the model is rebuilt from the parameters.

    python tests/golden/make_thermoelastic.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from misc_amd.meshes import thermoelastic_block, uniform_e_field  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [((4, 3, 2), 11), ((3, 3, 3), 12)]


def case_model(dims, seed):
    """The fixture's model: per-element moduli in [150, 250] GPa and an
    anisotropic expansion per element (all six Voigt components)."""
    nx, ny, nz = dims
    E = nx * ny * nz
    rng = np.random.default_rng(seed)
    alpha = rng.uniform(0.5e-5, 2.0e-5, (E, 6))
    alpha[:, 3:] = rng.uniform(-0.3e-5, 0.3e-5, (E, 3))
    m = thermoelastic_block(nx, ny, nz, alpha=alpha, e_field=uniform_e_field(E, 150e9, 250e9, seed), pressure=1e6)
    T = rng.uniform(0.0, 100.0, m.num_nodes)  # by rank
    return m, T


def main():
    for dims, seed in CASES:
        m, T = case_model(dims, seed)
        ref = O.run_reference(m, "thermoelastic", temperatures=T)
        out = dict(dims=np.array(dims), seed=np.array(seed), temperatures=T, elasticLoad=ref["elasticLoad"],
                   currentRank=ref["currentRank"], nDofIndex=ref["nDofIndex"])
        fn = os.path.join(HERE, "ref_thermoelastic_%dx%dx%d.npz" % dims)
        np.savez_compressed(fn, **out)
        print(fn)


if __name__ == "__main__":
    main()
