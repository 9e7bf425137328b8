"""Pins of the CPU oracle of the learned preconditioner's network
(oracle/gnn_oracle.py) on the CPU: its hand-derived gradient against central
finite differences, its loss against the energy norm of the remaining error
computed from e1 = A^-1 r1 directly, and the blocks at Y = 0 against the
inverse of the diagonal block (the block-Jacobi limit)."""

import numpy as np
import scipy.sparse as sp

from oracle import gnn_oracle as GO


def random_field(N=30, seed=0, samples=5):
    rng = np.random.default_rng(seed)
    # symmetric random graph, every node with at least one neighbour
    Adj = rng.random((N, N)) < 0.15
    Adj = np.triu(Adj, 1)
    Adj = Adj | Adj.T
    for i in range(N):
        if not Adj[i].any():
            j = (i + 1) % N
            Adj[i, j] = Adj[j, i] = True
    rg = np.zeros(N + 1, np.int32)
    adj = []
    for i in range(N):
        nb = np.nonzero(Adj[i])[0]
        adj.extend(nb)
        rg[i + 1] = len(adj)
    # an SPD operator with 3 x 3 node blocks on the graph (component-major DOFs)
    n = 3 * N
    nd = np.stack([np.arange(N), N + np.arange(N), 2 * N + np.arange(N)], axis=1)
    M = np.zeros((n, n))
    for i in range(N):
        for j in list(np.nonzero(Adj[i])[0]) + [i]:
            M[np.ix_(nd[i], nd[j])] += 0.3 * rng.standard_normal((3, 3))
    Adense = M @ M.T + 2.0 * np.eye(n)
    A = sp.csr_matrix(Adense)
    X = rng.standard_normal((N, GO.F))
    D = np.stack([Adense[np.ix_(nd[i], nd[i])] for i in range(N)])
    G = np.transpose(np.linalg.inv(np.linalg.cholesky(D)), (0, 2, 1))
    R1 = rng.standard_normal((samples, n))
    E1 = np.linalg.solve(Adense, R1.T).T
    fld = dict(range=rg, adj=np.asarray(adj, np.int32), X=X, G=G, A=A, node_dofs=nd, R1=R1,
               eTr=float(np.sum(E1 * R1)))
    return fld, E1, Adense, D


def test_gradient_matches_central_differences():
    for H, L, seed in [(8, 2, 1), (16, 3, 2)]:
        fld, _, _, _ = random_field(seed=seed)
        _, n = GO.layout(H, L)
        p = np.random.default_rng(seed).standard_normal(n) * 0.3
        l0, g = GO.loss_grad(p, fld, H, L)
        rng = np.random.default_rng(seed + 10)
        for _ in range(6):
            d = rng.standard_normal(n)
            h = 1e-6
            fd = (GO.loss(p + h * d, fld, H, L) - GO.loss(p - h * d, fld, H, L)) / (2 * h)
            assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd))


def test_loss_is_energy_norm_of_remaining_error():
    fld, E1, Ad, _ = random_field(seed=4)
    H, L = 8, 2
    _, n = GO.layout(H, L)
    p = np.random.default_rng(4).standard_normal(n) * 0.3
    B, _ = GO.forward(p, fld["X"], fld["range"], fld["adj"], fld["G"], H, L)
    num = den = 0.0
    for r, e in zip(fld["R1"], E1):
        d = e - GO.block_apply(B, fld["node_dofs"], r)
        num += d @ Ad @ d
        den += e @ Ad @ e
    assert abs(GO.loss(p, fld, H, L) - num / den) < 1e-12


def test_zero_correction_is_block_jacobi():
    fld, _, _, D = random_field(seed=5)
    H, L = 8, 2
    _, n = GO.layout(H, L)
    B, _ = GO.forward(np.zeros(n), fld["X"], fld["range"], fld["adj"], fld["G"], H, L)
    np.testing.assert_allclose(B, np.linalg.inv(D), rtol=1e-12, atol=1e-14)
