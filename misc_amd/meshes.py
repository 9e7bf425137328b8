"""Synthetic structured models for the workloads of the hot path.

A :class:`Model` holds exactly what the reference's model file carries for the
linear-static, thermal and Newmark cases this package solves
(Working/InputTemplate.yaml): node coordinates, element connectivity and type,
per-element material constants, multi-point constraints and nodal loads.
Numbering follows the reference's conventions: node and element ids are
1-based, connectivity rows hold 8 node ids padded with 0
(Working/AStrO_globalData.f90 ``elementList(8, numEls)``), element type codes
are those of Working/AStrO_input.f90:429-452 (8 = brick8, 41 = shell4).

Node id = i + NX*(j + NY*k) + 1 with x fastest; elements are ordered x, y, z.
Brick connectivity (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k) then the same
four at k+1, which gives detJ > 0 with the shape functions of
Working/AStrO_r_elementEqns.f90:27-35.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

BRICK8 = 8
SHELL4 = 41


@dataclass
class MPCTerm:
    """One ``[eqn, node|set, dof, coef]`` line of the constraints block."""

    eqn: int
    nodes: np.ndarray  # 1-based node ids
    dof: int
    coef: float
    is_set: bool
    set_name: str = ""


@dataclass
class Model:
    name: str
    nodes: np.ndarray            # (N, 3) float64
    elements: np.ndarray         # (E, 8) int32, 1-based, 0-padded
    etype: np.ndarray            # (E,) int32
    eprops: np.ndarray           # (E, 9) E1,E2,E3,nu12,nu13,nu23,G12,G13,G23
    density: np.ndarray          # (E,)
    cond: np.ndarray             # (E, 6) k11,k22,k33,k12,k13,k23
    mpc: List[MPCTerm] = field(default_factory=list)
    mpc_rhs: List[float] = field(default_factory=list)
    load_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    load_vals: np.ndarray = field(default_factory=lambda: np.zeros((0, 7)))
    node_sets: dict = field(default_factory=dict)
    element_sets: dict = field(default_factory=dict)  # name -> 1-based element ids
    colors: Optional[np.ndarray] = None  # (E,) element colour, conflict-free scatter
    # shell sections (shell elements only): per element nL layers of
    # (E1, E2, nu12, G12, thickness, angle in radians), z offset, and the two
    # section orientation vectors (x axis, a vector in the x-y plane)
    layers: Optional[np.ndarray] = None   # (E, nL, 6)
    z_offset: Optional[np.ndarray] = None  # (E,)
    orient: Optional[np.ndarray] = None   # (E, 6)
    spec_heat: Optional[np.ndarray] = None  # (E,) specific heat (transient thermal)
    therm_exp: Optional[np.ndarray] = None  # (E, 6) thermal expansion, Voigt 11 22 33 12 13 23

    @property
    def num_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def num_els(self) -> int:
        return int(self.elements.shape[0])


def iso_props(E, nu, G=None):
    """Engineering constants of an isotropic material, G = E / (2 (1 + nu))
    as Working/AStrO_input.f90:807-809 derives it when only E and nu are given."""
    if G is None:
        G = E / (2.0 * (1.0 + nu))
    return np.array([E, E, E, nu, nu, nu, G, G, G], dtype=np.float64)


def brick_block(nx, ny, nz, lx, ly, lz):
    """Nodes and brick8 connectivity of an nx*ny*nz structured block."""
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    i = np.arange(NX)
    j = np.arange(NY)
    k = np.arange(NZ)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    nodes = np.stack([I.ravel() * (lx / nx), J.ravel() * (ly / ny), K.ravel() * (lz / nz)], axis=1)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()

    def nid(a, b, c):
        return a + NX * (b + NY * c) + 1

    conn = np.stack(
        [nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
         nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1), nid(ei, ej + 1, ek + 1)],
        axis=1).astype(np.int32)
    colors = ((ei & 1) | ((ej & 1) << 1) | ((ek & 1) << 2)).astype(np.int32)
    return nodes, conn, colors


def face_pressure_loads(nx, ny, nz, lx, ly, lz, pressure):
    """Consistent nodal forces of a uniform pressure on the top face z = lz of a
    brick block with bilinear faces: p * hx * hy / 4 from every face touching the
    node (interior 1, edge 1/2, corner 1/4 of p * hx * hy), pointing -z."""
    NX, NY = nx + 1, ny + 1
    hx, hy = lx / nx, ly / ny
    wx = np.full(NX, 1.0)
    wx[0] = wx[-1] = 0.5
    wy = np.full(NY, 1.0)
    wy[0] = wy[-1] = 0.5
    W = np.outer(wy, wx).ravel()  # j slow, i fast
    ids = (np.arange(NX * NY) + NX * NY * nz + 1).astype(np.int32)
    vals = np.zeros((NX * NY, 7))
    vals[:, 2] = -pressure * hx * hy * W
    return ids, vals


def clamp_set_mpc(node_ids, dofs, set_name, first_eqn=1):
    """One constraint equation per DOF naming a node set: the reference expands
    it into one single-DOF row per set member
    (Working/AStrO_bookKeeping.f90:585-601)."""
    terms = []
    for q, d in enumerate(dofs):
        terms.append(MPCTerm(eqn=first_eqn + q, nodes=np.asarray(node_ids, np.int32), dof=d, coef=1.0,
                             is_set=True, set_name=set_name))
    return terms


def cantilever_block(nx, ny, nz, lx=1.0, ly=1.0, lz=0.5, E=200e9, nu=0.3, pressure=1e6,
                     e_field=None, name=None):
    """Linear-static brick8 block clamped on z = 0 with a uniform top-face pressure
    written as consistent nodal loads.  ``e_field`` optionally gives one Young's
    modulus per element (same nu, G = E / (2 (1 + nu)))."""
    nodes, conn, colors = brick_block(nx, ny, nz, lx, ly, lz)
    E_el = np.full(conn.shape[0], float(E)) if e_field is None else np.asarray(e_field, np.float64)
    eprops = np.stack([iso_props(e, nu) for e in E_el]) if e_field is not None else \
        np.tile(iso_props(E, nu), (conn.shape[0], 1))
    NX, NY = nx + 1, ny + 1
    bottom = np.arange(1, NX * NY + 1, dtype=np.int32)
    m = Model(name=name or f"brick{nx}x{ny}x{nz}", nodes=nodes, elements=conn,
              etype=np.full(conn.shape[0], BRICK8, np.int32), eprops=eprops,
              density=np.zeros(conn.shape[0]), cond=np.zeros((conn.shape[0], 6)), colors=colors)
    m.mpc = clamp_set_mpc(bottom, [1, 2, 3], "bottom")
    m.mpc_rhs = [0.0, 0.0, 0.0]
    m.node_sets["bottom"] = bottom
    ids, vals = face_pressure_loads(nx, ny, nz, lx, ly, lz, pressure)
    m.load_nodes, m.load_vals = ids, vals
    return m


def thermal_block(nx, ny, nz, lx=1.0, ly=1.0, lz=0.5, k=50.0, t_top=100.0, t_bottom=0.0, rho=0.0, c=0.0,
                  name=None):
    """Conduction on a brick8 block: temperature fixed on the top and bottom
    faces through dof-7 constraint equations on node sets; density rho and
    specific heat c give the capacitance of the transient path."""
    nodes, conn, colors = brick_block(nx, ny, nz, lx, ly, lz)
    E = conn.shape[0]
    cond = np.tile(np.array([k, k, k, 0.0, 0.0, 0.0]), (E, 1))
    NX, NY = nx + 1, ny + 1
    bottom = np.arange(1, NX * NY + 1, dtype=np.int32)
    top = bottom + NX * NY * nz
    m = Model(name=name or f"thermal{nx}x{ny}x{nz}", nodes=nodes, elements=conn,
              etype=np.full(E, BRICK8, np.int32), eprops=np.tile(iso_props(200e9, 0.3), (E, 1)),
              density=np.full(E, float(rho)), cond=cond, colors=colors, spec_heat=np.full(E, float(c)))
    m.mpc = [MPCTerm(eqn=1, nodes=top, dof=7, coef=1.0, is_set=True, set_name="top"),
             MPCTerm(eqn=2, nodes=bottom, dof=7, coef=1.0, is_set=True, set_name="bottom")]
    m.mpc_rhs = [t_top, t_bottom]
    m.node_sets["top"] = top
    m.node_sets["bottom"] = bottom
    return m


def thermoelastic_block(nx, ny, nz, lx=1.0, ly=1.0, lz=0.5, E=200e9, nu=0.3, k=50.0, alpha=1.2e-5,
                        t_top=100.0, t_bottom=0.0, pressure=0.0, e_field=None, name=None):
    """The cantilever block (clamped on z = 0, top pressure as nodal loads)
    with conduction and thermal expansion: temperature fixed on the top and
    bottom faces (dof-7 equations after the three clamp equations), k on
    the diagonal, and isotropic expansion alpha (alpha may also be an (E, 6)
    array of Voigt components)."""
    m = cantilever_block(nx, ny, nz, lx, ly, lz, E, nu, pressure, e_field, name or f"thermoel{nx}x{ny}x{nz}")
    Ne = m.num_els
    m.cond = np.tile(np.array([k, k, k, 0.0, 0.0, 0.0]), (Ne, 1))
    a = np.asarray(alpha, np.float64)
    m.therm_exp = np.tile([a, a, a, 0.0, 0.0, 0.0], (Ne, 1)) if a.ndim == 0 else a.reshape(Ne, 6).copy()
    NX, NY = nx + 1, ny + 1
    bottom = np.arange(1, NX * NY + 1, dtype=np.int32)
    top = bottom + NX * NY * nz
    m.mpc = m.mpc + [MPCTerm(eqn=4, nodes=top, dof=7, coef=1.0, is_set=True, set_name="top"),
                     MPCTerm(eqn=5, nodes=bottom, dof=7, coef=1.0, is_set=True, set_name="bottom")]
    m.mpc_rhs = m.mpc_rhs + [t_top, t_bottom]
    m.node_sets["top"] = top
    return m


def uniform_e_field(num_els, lo, hi, seed):
    """Young's modulus per element drawn i.i.d. uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=num_els)


def lognormal_e_field(nx, ny, nz, median, log_sigma, corr_len, seed):
    """Young's modulus per element, log E = log(median) + log_sigma * g.

    g is white noise on the element grid, filtered by a Gaussian kernel of
    standard deviation corr_len / 2 elements (periodic edges), then shifted
    and scaled to zero mean and unit variance.  With that kernel the field's
    correlation function is exp(-r^2 / corr_len^2), so corr_len is the
    correlation length in elements.  Element order is x fastest, then y, z,
    as brick_block numbers them."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nz, ny, nx))
    g = gaussian_filter(w, sigma=corr_len / 2.0, mode="wrap")
    g = (g - g.mean()) / g.std()
    return np.exp(np.log(median) + log_sigma * g).ravel()


def top_patch_loads(nx, ny, nz, lx, ly, lz, px, py, total):
    """Consistent nodal forces of a uniform traction over a px x py-element
    patch centred on the top face z = lz, summing to `total` in -z: each
    patch element adds a quarter of its share to each of its four top nodes."""
    NX, NY = nx + 1, ny + 1
    i0, j0 = (nx - px) // 2, (ny - py) // 2
    share = total / (px * py)
    f = np.zeros((NY, NX))
    for j in range(j0, j0 + py):
        for i in range(i0, i0 + px):
            f[j:j + 2, i:i + 2] += 0.25 * share
    jj, ii = np.nonzero(f)
    ids = (ii + NX * (jj + NY * nz) + 1).astype(np.int32)
    vals = np.zeros((ids.size, 7))
    vals[:, 2] = -f[jj, ii]
    return ids, vals


def impulse_block(nx=48, ny=48, nz=24, lx=1.0, ly=1.0, lz=0.5, median_E=70e9, log_sigma=0.5, corr_len=8.0,
                  nu=0.3, rho=2700.0, patch=8, total=1e3, seed=0, name=None):
    """The Newmark workload: brick8 block with a log-normal Young's modulus
    field, density rho, clamped on z = 0, and a load of `total` newtons
    (at load factor 1) on a patch x patch-element area centred on the top
    face.  The time history is applied through per-step load factors
    (see half_sine)."""
    nodes, conn, colors = brick_block(nx, ny, nz, lx, ly, lz)
    E_el = lognormal_e_field(nx, ny, nz, median_E, log_sigma, corr_len, seed)
    E = conn.shape[0]
    NX, NY = nx + 1, ny + 1
    bottom = np.arange(1, NX * NY + 1, dtype=np.int32)
    m = Model(name=name or f"impulse{nx}x{ny}x{nz}", nodes=nodes, elements=conn,
              etype=np.full(E, BRICK8, np.int32), eprops=np.stack([iso_props(e, nu) for e in E_el]),
              density=np.full(E, float(rho)), cond=np.zeros((E, 6)), colors=colors)
    m.mpc = clamp_set_mpc(bottom, [1, 2, 3], "bottom")
    m.mpc_rhs = [0.0, 0.0, 0.0]
    m.node_sets["bottom"] = bottom
    m.load_nodes, m.load_vals = top_patch_loads(nx, ny, nz, lx, ly, lz, min(patch, nx), min(patch, ny), total)
    return m


def half_sine(n_steps, dt, duration):
    """Load factor at the end of each step t_k = k dt (k = 1..n_steps): a
    half-sine pulse sin(pi t / duration) for t <= duration, zero after."""
    t = dt * np.arange(1, n_steps + 1)
    return np.where(t <= duration, np.sin(np.pi * t / duration), 0.0)


def shell_plate(nx, ny, lx=1.0, ly=0.1, E=70e9, nu=0.33, thickness=0.002, t_field=None, tip_load=1.0, name=None):
    """Cantilever plate of nx x ny shell4 elements in the x-y plane, clamped on
    x = 0 (DOFs 1-6 of every root node through node-set constraint equations),
    tip_load newtons in +z on every node of x = lx.  One isotropic layer per
    element, G = E / (2 (1 + nu)); t_field optionally gives the thickness per
    element.  Section orientation [1,0,0], [0,1,0], zero offset.  Node id =
    i + NX j + 1, elements x fastest with connectivity (i,j), (i+1,j),
    (i+1,j+1), (i,j+1)."""
    NX, NY = nx + 1, ny + 1
    J, I = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    nodes = np.stack([I.ravel() * (lx / nx), J.ravel() * (ly / ny), np.zeros(NX * NY)], axis=1)
    ej, ei = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej = ei.ravel(), ej.ravel()

    def nid(a, b):
        return a + NX * b + 1

    conn = np.zeros((ei.size, 8), np.int32)
    conn[:, 0] = nid(ei, ej)
    conn[:, 1] = nid(ei + 1, ej)
    conn[:, 2] = nid(ei + 1, ej + 1)
    conn[:, 3] = nid(ei, ej + 1)
    nE = conn.shape[0]
    colors = ((ei & 1) | ((ej & 1) << 1)).astype(np.int32)
    t_el = np.full(nE, float(thickness)) if t_field is None else np.asarray(t_field, np.float64)
    G = E / (2.0 * (1.0 + nu))
    layers = np.zeros((nE, 1, 6))
    layers[:, 0, :] = [E, E, nu, G, 0.0, 0.0]
    layers[:, 0, 4] = t_el
    m = Model(name=name or f"plate{nx}x{ny}", nodes=np.ascontiguousarray(nodes), elements=conn,
              etype=np.full(nE, SHELL4, np.int32), eprops=np.tile(iso_props(E, nu), (nE, 1)),
              density=np.zeros(nE), cond=np.zeros((nE, 6)), colors=colors, layers=layers,
              z_offset=np.zeros(nE), orient=np.tile([1.0, 0.0, 0.0, 0.0, 1.0, 0.0], (nE, 1)))
    root = (np.arange(NY) * NX + 1).astype(np.int32)
    tip = root + nx
    m.mpc = clamp_set_mpc(root, [1, 2, 3, 4, 5, 6], "root")
    m.mpc_rhs = [0.0] * 6
    m.node_sets["root"] = root
    m.node_sets["tip"] = tip
    vals = np.zeros((tip.size, 7))
    vals[:, 2] = tip_load
    m.load_nodes, m.load_vals = tip, vals
    return m


def strip_plate(nx=128, ny=16, lx=1.0, ly=0.1, E=70e9, nu=0.33, t0=0.002, t_spread=0.1, strip=2, seed=0,
                tip_load=1.0, name=None):
    """The design-sensitivity workload: cantilever shell plate (see
    shell_plate) whose element thickness is t0 (1 + U(-t_spread, t_spread))
    from seed, with one design variable per `strip` element columns (the
    thickness of the strip's single layer).  Returns (model, elDV) with elDV
    the 0-based design variable of each element."""
    rng = np.random.default_rng(seed)
    t = t0 * (1.0 + rng.uniform(-t_spread, t_spread, size=nx * ny))
    m = shell_plate(nx, ny, lx=lx, ly=ly, E=E, nu=nu, t_field=t, tip_load=tip_load,
                    name=name or f"strips{nx}x{ny}")
    ei = np.tile(np.arange(nx), ny)
    return m, (ei // strip).astype(np.int32)
