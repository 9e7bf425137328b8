"""The Newmark objective gradient of the CPU oracle (oracle/newmark_adjoint.py,
the reference's getTotaldLdD dynamic branch with augmentdLdu and
solveDynamicAdjoint) against central finite differences of the whole
forward-then-sum pipeline, for material, density, coordinate and load
variables, with a term active on part of the steps."""

import numpy as np

from design_cases import elastic_case
from misc_amd.astro import ObjectiveTerm
from oracle import design_oracle as DO
from oracle.newmark_adjoint import NewmarkAdjointOracle

DT, STEPS = 2e-5, 6


def case():
    m, dv, top = elastic_case()
    terms = [ObjectiveTerm("displacement", top, component=3, coef=1e6, expo=2.0),
             ObjectiveTerm("strainEnergy", np.arange(1, m.num_els + 1), operator="volumeIntegral")]
    windows = [(0.0, 1.0), (3.5 * DT, 1.0)]  # the second term from step 4 on
    scales = np.sin(np.pi * np.arange(1, STEPS + 1) / (STEPS + 1))
    return m, dv, terms, windows, scales


def test_newmark_gradient_matches_finite_differences():
    m, dv, terms, windows, scales = case()
    o = NewmarkAdjointOracle(m, dv, dt=DT)
    U, V, A = o.forward(scales)
    J = o.objective(terms, U, windows)
    G = o.gradient(terms, U, A, scales, windows)
    assert J > 0
    for k in range(len(dv.category)):
        d = 1e-3  # in units of the variable: each property moves by coef * d
        Jp = NewmarkAdjointOracle(DO.perturbed(m, dv, k, d), dv, dt=DT)
        Jm = NewmarkAdjointOracle(DO.perturbed(m, dv, k, -d), dv, dt=DT)
        fd = (Jp.objective(terms, Jp.forward(scales)[0], windows) -
              Jm.objective(terms, Jm.forward(scales)[0], windows)) / (2 * d)
        # the central difference carries O(d^2) truncation and the rounding of
        # J ~ 50 over 2 d: about 1e-6 of the smaller gradients
        assert abs(fd - G[k]) <= 2e-5 * abs(G[k]), (k, fd, G[k])
