"""Pins of the design / objective / adjoint oracle (CPU only).

The oracle's gradient (adjoint + complex steps) is checked against central
finite differences of the whole pipeline: move one design variable, rebuild
the model, re-solve, re-evaluate the objective.  Closed forms pin the
objective values and the homogeneity of the compliance: with every modulus
scaled by (1 + d) the compliance scales by 1 / (1 + d), and with every
coordinate scaled by (1 + d) it scales the same way (K_e grows with the
element length in 3-D), so the relative-change gradients sum to -C."""

import numpy as np
import pytest

from misc_amd.astro import DesignVariables, ObjectiveTerm
from misc_amd.meshes import cantilever_block
from oracle import design_oracle as DO
from tests.design_cases import elastic_case, elastic_terms, thermal_case


def _fd(model, dv, terms, k, delta, thermal=False):
    vals = []
    for sgn in (1, -1):
        o = DO.DesignOracle(DO.perturbed(model, dv, k, sgn * delta), dv)
        if thermal:
            o.solve_thermal()
        else:
            o.solve()
        vals.append(o.objective(terms).sum())
    return (vals[0] - vals[1]) / (2 * delta)


def test_gradient_matches_central_differences_elastic():
    m, dv, top = elastic_case()
    terms = elastic_terms(m, top)
    o = DO.DesignOracle(m, dv)
    o.solve()
    g = o.gradient(terms)
    # steps: 1e-5 of each variable's natural scale (coefficients carry units)
    deltas = [1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-5]
    for k in range(len(dv.category)):
        fd = _fd(m, dv, terms, k, deltas[k])
        scale = max(abs(fd), abs(g[k]), 1e-12 * np.abs(g).max())
        assert abs(fd - g[k]) <= 2e-6 * scale + 1e-9 * np.abs(g).max(), (k, fd, g[k])


def test_gradient_matches_central_differences_thermal():
    m, dv, terms = thermal_case()
    o = DO.DesignOracle(m, dv)
    o.solve_thermal()
    g = o.gradient(terms)
    assert np.abs(g).max() > 0
    # The penalty weight of the temperature constraints is 100 max |K_t|, which
    # moves with the conductivity: the differences see that move, the adjoint
    # (like the reference's dRt/dd, element terms only) holds the weight fixed.
    # Its effect is a few 1e-6 here, so the bar is 2e-5: a wrong sign, index or
    # term is O(1).
    for k in range(len(dv.category)):
        fd = _fd(m, dv, terms, k, 1e-3, thermal=True)
        assert abs(fd - g[k]) <= 2e-5 * max(abs(fd), abs(g[k])) + 1e-10 * np.abs(g).max(), (k, fd, g[k])


def _load_compliance_case():
    m = cantilever_block(3, 2, 2, e_field=np.random.default_rng(1).uniform(150e9, 250e9, 12), pressure=0.0)
    top = np.arange(4 * 3 * 2 + 1, 4 * 3 * 3 + 1, dtype=np.int32)
    m.load_nodes = top
    m.load_vals = np.zeros((top.size, 7))
    m.load_vals[:, 2] = -1e6
    compliance = ObjectiveTerm("displacement", top, component=3, coef=-1e6, expo=1.0)
    return m, top, compliance


def test_compliance_is_homogeneous_in_the_moduli():
    m, top, comp = _load_compliance_case()
    E = m.num_els
    rng = np.arange(E + 1, dtype=np.int32) * 6
    ids = np.tile(np.arange(1, 7, dtype=np.int32), E)
    coef = np.concatenate([m.eprops[e, [0, 1, 2, 6, 7, 8]] for e in range(E)])
    dv = DesignVariables(category=["modulus"] * 3 + ["shearModulus"] * 3, component=[1, 2, 3, 1, 2, 3],
                         elToDRange=rng, elToD=ids, elToCoef=coef)
    o = DO.DesignOracle(m, dv)
    o.solve()
    C = o.objective([comp])[0]
    g = o.gradient([comp])
    assert C > 0
    assert abs(g.sum() + C) <= 1e-9 * C


def test_compliance_is_homogeneous_in_the_coordinates():
    m, top, comp = _load_compliance_case()
    N = m.num_nodes
    rng = np.arange(N + 1, dtype=np.int32) * 3
    ids = np.tile(np.arange(1, 4, dtype=np.int32), N)
    coef = m.nodes.reshape(-1).copy()
    dv = DesignVariables(category=["nodeCoord"] * 3, component=[1, 2, 3], ndToDRange=rng, ndToD=ids, ndToCoef=coef)
    o = DO.DesignOracle(m, dv)
    o.solve()
    C = o.objective([comp])[0]
    g = o.gradient([comp])
    assert abs(g.sum() + C) <= 1e-9 * C
    # the volume scales as (1 + d)^3: each axis contributes V
    alls = np.arange(1, m.num_els + 1, dtype=np.int32)
    vol = ObjectiveTerm("volume", alls, operator="volumeIntegral")
    gv = o.gradient([vol])
    V = o.objective([vol])[0]
    assert abs(V - 1.0 * 1.0 * 0.5) <= 1e-14
    assert np.allclose(gv, V, rtol=1e-12)


def test_objective_closed_forms():
    m, dv, top = elastic_case()
    o = DO.DesignOracle(m, dv)
    u = o.solve()
    E = m.num_els
    alls = np.arange(1, E + 1, dtype=np.int32)
    vals = o.objective([ObjectiveTerm("volume", alls, operator="volumeIntegral"),
                        ObjectiveTerm("mass", alls, operator="volumeIntegral"),
                        ObjectiveTerm("strainEnergy", alls, operator="volumeIntegral")])
    hv = 0.5 / E  # equal cubes-of-a-block: lx * ly * lz / E
    assert abs(vals[0] - 0.5) <= 1e-14
    assert abs(vals[1] - hv * m.density.sum()) <= 1e-12 * vals[1]
    # small strains: the Green-Lagrange energy is u.K u / 2 to O(strain)
    uKu = u @ (o.A @ u)  # the penalty rows vanish at the clamped solution to rounding
    assert abs(vals[2] - 0.5 * uKu) <= 1e-4 * vals[2]
    # stress at the centre = C times the strain there
    st = o.objective([ObjectiveTerm("stress", [1], component=c, coef=1.0) for c in range(1, 7)])
    sn = o.objective([ObjectiveTerm("strain", [1], component=c, coef=1.0) for c in range(1, 7)])
    assert np.allclose(st, DO.stiffness(m.eprops[0]) @ sn, rtol=1e-12, atol=1e-12 * np.abs(st).max())


def test_dRdD_columns_match_differences_of_K_u():
    m, dv, top = elastic_case()
    o = DO.DesignOracle(m, dv)
    u = o.solve()
    D = o.dRdD()
    for k in (0, 1, 4, 6):
        rs = []
        for sgn in (1, -1):
            p = DO.DesignOracle(DO.perturbed(m, dv, k, sgn * 1e-4), dv)
            p.solve()
            rs.append(p.A @ u)  # (K(d) + C^T C) u at the same u
        fd = (rs[0] - rs[1]) / 2e-4
        # the penalty scaling follows max |K|, which moves with d: compare away from the clamped rows
        rows = np.setdiff1d(np.arange(o.n), _clamped_dofs(o, m))
        assert np.allclose(D[rows, k], fd[rows], rtol=1e-6, atol=1e-6 * np.abs(D[:, k]).max()), k


def _clamped_dofs(o, m):
    bottom = m.node_sets["bottom"]
    return np.array([o.nDof[n - 1, c] - 1 for n in bottom for c in range(3)])
