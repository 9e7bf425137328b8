"""CPU oracle of the learned block preconditioner's graph network (numpy,
float64): the same parameter layout, forward pass, loss and hand-derived
backward pass as the library's HIP kernels (misc_amd/csrc/astro_gnn.hip),
plus the node features of an assembled operator.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the device's
features, blocks, loss and gradients; never by the product package.  The
backward pass here is itself pinned by central finite differences
(tests/test_gnn_oracle.py).

The network has no counterpart in the reference (whose Newmark loop,
Working/AStrO_commandFunctions.f90:781-815, uses the skyline LDL^T alone).
"""

from __future__ import annotations

import numpy as np

F = 9     # node features
OUT = 6   # symmetric 3 x 3 correction


def layout(H, L):
    """Offsets of the flat parameter vector: W_in (H x F), b_in (H), then per
    layer W_s (H x H), W_n (H x H), b (H), then W_out (6 x H), b_out (6)."""
    off = {}
    o = 0
    off["Win"] = o
    o += H * F
    off["bin"] = o
    o += H
    for l in range(1, L + 1):
        off[f"Ws{l}"] = o
        o += H * H
        off[f"Wn{l}"] = o
        o += H * H
        off[f"b{l}"] = o
        o += H
    off["Wout"] = o
    o += OUT * H
    off["bout"] = o
    o += OUT
    return off, o


def unpack(p, H, L):
    off, n = layout(H, L)
    assert p.size == n
    P = {"Win": p[off["Win"]:off["Win"] + H * F].reshape(H, F), "bin": p[off["bin"]:off["bin"] + H],
         "Wout": p[off["Wout"]:off["Wout"] + OUT * H].reshape(OUT, H), "bout": p[off["bout"]:off["bout"] + OUT]}
    for l in range(1, L + 1):
        P[f"Ws{l}"] = p[off[f"Ws{l}"]:off[f"Ws{l}"] + H * H].reshape(H, H)
        P[f"Wn{l}"] = p[off[f"Wn{l}"]:off[f"Wn{l}"] + H * H].reshape(H, H)
        P[f"b{l}"] = p[off[f"b{l}"]:off[f"b{l}"] + H]
    return P


def mean_agg(h, rng, adj):
    """a_i = mean over the neighbours j of i of h_j (0 for an isolated node)."""
    N = len(rng) - 1
    deg = np.diff(rng)
    rows = np.repeat(np.arange(N), deg)
    s = np.zeros_like(h)
    np.add.at(s, rows, h[adj])
    with np.errstate(divide="ignore"):
        inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    return s * inv[:, None], inv, rows


def sym(y):
    """(N, 6) -> (N, 3, 3): [[y0, y3, y4], [y3, y1, y5], [y4, y5, y2]]."""
    Y = np.zeros((y.shape[0], 3, 3))
    Y[:, 0, 0], Y[:, 1, 1], Y[:, 2, 2] = y[:, 0], y[:, 1], y[:, 2]
    Y[:, 0, 1] = Y[:, 1, 0] = y[:, 3]
    Y[:, 0, 2] = Y[:, 2, 0] = y[:, 4]
    Y[:, 1, 2] = Y[:, 2, 1] = y[:, 5]
    return Y


def forward(p, X, rng, adj, G, H=64, L=3):
    """Blocks B_i = G_i (I + Y_i) G_i^T and the cached activations."""
    P = unpack(np.asarray(p, np.float64), H, L)
    hs = [np.maximum(X @ P["Win"].T + P["bin"], 0.0)]
    aggs = [None]
    for l in range(1, L + 1):
        a, _, _ = mean_agg(hs[-1], rng, adj)
        aggs.append(a)
        hs.append(np.maximum(hs[-1] @ P[f"Ws{l}"].T + a @ P[f"Wn{l}"].T + P[f"b{l}"], 0.0))
    y = hs[-1] @ P["Wout"].T + P["bout"]
    M = np.eye(3)[None] + sym(y)
    B = G @ M @ np.transpose(G, (0, 2, 1))
    return B, dict(P=P, hs=hs, aggs=aggs, y=y)


def block_apply(B, node_dofs, r):
    """q = B r node by node: q[dofs of i] = B_i r[dofs of i]."""
    q = np.zeros_like(r)
    q[node_dofs] = np.einsum("nab,nb->na", B, r[node_dofs])
    return q


def loss(p, fld, H=64, L=3):
    """Energy norm of the error the block stage leaves, over the field's
    pairs (r1, e1 = A^-1 r1): sum (e1 - B r1)^T A (e1 - B r1) / sum e1^T A e1
    = (eTr + sum_s (-2 r1^T q + q^T A q)) / eTr with q = B r1."""
    return loss_grad(p, fld, H, L)[0]


def loss_grad(p, fld, H=64, L=3):
    """The loss and its gradient with respect to the flat parameter vector:
    dl/dB_i = sum_s -2 (r - A q)_i r_i^T / eTr, then dl/dM = G^T (dl/dB) G,
    M = I + Y, and the network's backward pass."""
    X, rng, adj, G = fld["X"], fld["range"], fld["adj"], fld["G"]
    A, nd = fld["A"], fld["node_dofs"]
    B, c = forward(p, X, rng, adj, G, H, L)
    P, hs, aggs = c["P"], c["hs"], c["aggs"]
    acc = 0.0
    gB = np.zeros_like(B)
    for r in fld["R1"]:
        q = block_apply(B, nd, r)
        w = A @ q
        acc += -2.0 * (r @ q) + q @ w
        gB += -2.0 * np.einsum("na,nb->nab", (r - w)[nd], r[nd])
    lval = float((fld["eTr"] + acc) / fld["eTr"])
    gM = np.transpose(G, (0, 2, 1)) @ (gB / fld["eTr"]) @ G
    dy = np.stack([gM[:, 0, 0], gM[:, 1, 1], gM[:, 2, 2], gM[:, 0, 1] + gM[:, 1, 0], gM[:, 0, 2] + gM[:, 2, 0],
                   gM[:, 1, 2] + gM[:, 2, 1]], axis=1)
    g = {"Wout": dy.T @ hs[L], "bout": dy.sum(0)}
    dh = dy @ P["Wout"]
    N = len(rng) - 1
    deg = np.diff(rng)
    rows = np.repeat(np.arange(N), deg)
    invdeg = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    for l in range(L, 0, -1):
        dz = dh * (hs[l] > 0)
        g[f"Ws{l}"] = dz.T @ hs[l - 1]
        g[f"Wn{l}"] = dz.T @ aggs[l]
        g[f"b{l}"] = dz.sum(0)
        u = dz @ P[f"Wn{l}"]
        dprev = dz @ P[f"Ws{l}"]
        # transpose of the mean aggregation: h_j receives u_i / deg_i from each i with j in N(i)
        np.add.at(dprev, adj, u[rows] * invdeg[rows][:, None])
        dh = dprev
    dz = dh * (hs[0] > 0)
    g["Win"] = dz.T @ X
    g["bin"] = dz.sum(0)
    off, n = layout(H, L)
    flat = np.zeros(n)
    for k, v in g.items():
        flat[off[k]:off[k] + v.size] = v.ravel()
    return lval, flat


def features(K, C, node_dofs, constrained):
    """Node features of A = K + C^T C (scipy sparse), as k_gnn_features forms
    them before centring: (D11, D22, D33, D12, D13, D23) / s, log10 s, the
    constraint flag, |A(node, other nodes)|_F / |D|_F / 8, and G = L^-T of
    D = L L^T."""
    A = (K + C.T @ C).tocsr()
    N = node_dofs.shape[0]
    Xf = np.zeros((N, F))
    G = np.zeros((N, 3, 3))
    Ad = A.toarray() if A.shape[0] <= 4000 else None
    for i in range(N):
        d = node_dofs[i]
        D = (Ad[np.ix_(d, d)] if Ad is not None else A[d][:, d].toarray())
        s = np.trace(D) / 3
        rows = Ad[d] if Ad is not None else A[d].toarray()
        mask = np.ones(A.shape[0], bool)
        mask[d] = False
        coup = np.sum(rows[:, mask] ** 2)
        Xf[i] = [D[0, 0] / s, D[1, 1] / s, D[2, 2] / s, D[0, 1] / s, D[0, 2] / s, D[1, 2] / s, np.log10(s),
                 1.0 if constrained[i] else 0.0, np.sqrt(coup / np.sum(D ** 2)) / 8]
        Lc = np.linalg.cholesky(D)
        G[i] = np.linalg.inv(Lc).T
    return Xf, G
