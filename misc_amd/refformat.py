"""Writer and reader for the reference's YAML-like model file and command script.

The format is the one parsed by ``readModelInput`` (Working/AStrO_input.f90:27-1097)
and laid out in Working/InputTemplate.yaml; the command script is the one read
by Working/AStrO_runJob.f90:44-376 (Working/ScriptCommands.txt).  The
reference's parser is line-oriented with fixed rules that the writer honours:

* a section header (``nodes:``, ``loads:`` ...) is alone on its line, because
  the parser looks for the header followed by blanks (findSubstrings,
  Working/AStrO_input.f90:7-25);
* ``sets`` precede ``sections``, which look their element set up by name;
* list items are ``- [a, b, ...]`` on one line of at most 241 characters;
* reals are written with 17 significant digits so the list-directed read
  recovers the same double.

Materials are emitted one per distinct property row, with one section and one
element set per material, so per-element moduli survive the round trip.
"""

from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

from .meshes import Model, MPCTerm, BRICK8, SHELL4

_TYPE_NAME = {BRICK8: "brick8", SHELL4: "shell4"}


def _r(x: float) -> str:
    return repr(float(x))


def write_model(model: Model, path: str) -> None:
    lines: List[str] = []
    lines.append("nodes:")
    for i, (x, y, z) in enumerate(model.nodes, start=1):
        lines.append(f"    - [{i}, {_r(x)}, {_r(y)}, {_r(z)}]")
    lines.append("elements:")
    for t in np.unique(model.etype):
        idx = np.nonzero(model.etype == t)[0]
        nn = 8 if t == BRICK8 else 4
        lines.append(f"    - type: {_TYPE_NAME[int(t)]}")
        lines.append("      connectivity:")
        for e in idx:
            ids = ", ".join(str(int(v)) for v in model.elements[e, :nn])
            lines.append(f"          - [{e + 1}, {ids}]")
    # one material / section / element set per distinct property row
    tex = getattr(model, "therm_exp", None)
    tex = np.zeros((model.num_els, 6)) if tex is None else np.asarray(tex, np.float64).reshape(-1, 6)
    key = np.concatenate([model.eprops, model.density[:, None], model.cond, tex], axis=1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.ravel()
    lines.append("sets:")
    if model.node_sets:
        lines.append("    node:")
        for name, ids in model.node_sets.items():
            lines.append(f"        - name: {name}")
            lines.append("          labels:")
            lines.extend(_label_lines(np.asarray(ids)))
    lines.append("    element:")
    for q in range(uniq.shape[0]):
        lines.append(f"        - name: es{q + 1}")
        lines.append("          labels:")
        lines.extend(_label_lines(np.nonzero(inv == q)[0] + 1))
    for name, ids in getattr(model, "element_sets", {}).items():
        lines.append(f"        - name: {name}")
        lines.append("          labels:")
        lines.extend(_label_lines(np.asarray(ids)))
    lines.append("sections:")
    for q in range(uniq.shape[0]):
        lines.append("    - type: solid")
        lines.append(f"      material: mat{q + 1}")
        lines.append(f"      elementSet: es{q + 1}")
    lines.append("materials:")
    for q in range(uniq.shape[0]):
        ep = uniq[q, :9]
        lines.append(f"    - name: mat{q + 1}")
        lines.append(f"      density: {_r(uniq[q, 9])}")
        lines.append("      elastic:")
        lines.append(f"          E: [{_r(ep[0])}, {_r(ep[1])}, {_r(ep[2])}]")
        lines.append(f"          nu: [{_r(ep[3])}, {_r(ep[4])}, {_r(ep[5])}]")
        lines.append(f"          G: [{_r(ep[6])}, {_r(ep[7])}, {_r(ep[8])}]")
        k = uniq[q, 10:16]
        x = uniq[q, 16:22]
        if np.any(k != 0) or np.any(x != 0):
            lines.append("      thermal:")
            lines.append("          conductivity: [" + ", ".join(_r(v) for v in k) + "]")
            if np.any(x != 0):
                lines.append("          expansion: [" + ", ".join(_r(v) for v in x) + "]")
    if model.mpc:
        lines.append("constraints:")
        lines.append("    matrix:")
        for t in model.mpc:
            who = t.set_name if t.is_set else str(int(t.nodes[0]))
            lines.append(f"        - [{t.eqn}, {who}, {t.dof}, {_r(t.coef)}]")
        lines.append("    rhs:")
        for v in model.mpc_rhs:
            lines.append(f"        - {_r(v)}")
    if model.load_nodes.size:
        lines.append("loads:")
        lines.append("    - type: nodal")
        lines.append("      activeTime: 0.0")
        lines.append("      setLoads:")
        for nd, v in zip(model.load_nodes, model.load_vals):
            lines.append(f"          - [{int(nd)}, " + ", ".join(_r(x) for x in v) + "]")
    for ln in lines:
        if len(ln) > 241:
            raise ValueError("line longer than the reference parser reads: " + ln[:60])
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _label_lines(ids: np.ndarray) -> List[str]:
    """Compress sorted ids into ``[a, b]`` runs (inclusive), else single ids."""
    ids = np.sort(np.asarray(ids, dtype=np.int64))
    out = []
    if ids.size == 0:
        return out
    start = prev = int(ids[0])
    for v in ids[1:]:
        v = int(v)
        if v == prev + 1:
            prev = v
            continue
        out.append(_run(start, prev))
        start = prev = v
    out.append(_run(start, prev))
    return out


def _run(a: int, b: int) -> str:
    if a == b:
        return f"              - {a}"
    return f"              - [{a}, {b}]"


def write_design_vars(path: str, variables) -> None:
    """Design-variable file read by readDesignVarInput
    (Working/AStrO_input.f90:1099-1408; layout of Working/DVarInputTemplate.yaml).
    Each variable is a dict with category, subCategory (optional), component,
    elementSet or nodeSet (a set name of the model file), coefficients (a
    scalar for every member, or one value per member in set order) and
    activeTime (optional, a scalar or a [start, end] pair)."""
    lines = ["designVariables:"]
    for v in variables:
        lines.append(f"    - category: {v['category']}")
        if v.get("subCategory"):
            lines.append(f"      subCategory: {v['subCategory']}")
        lines.append(f"      component: {int(v.get('component', 1))}")
        if v.get("activeTime") is not None:
            at = v["activeTime"]
            if np.ndim(at):
                lines.append(f"      activeTime: [{_r(at[0])}, {_r(at[1])}]")
            else:
                lines.append(f"      activeTime: {_r(at)}")
        if v.get("elementSet"):
            lines.append(f"      elementSet: {v['elementSet']}")
        else:
            lines.append(f"      nodeSet: {v['nodeSet']}")
        c = v.get("coefficients", 1.0)
        if np.ndim(c):
            lines.append("      coefficients:")
            lines.extend(f"          - {_r(x)}" for x in c)
        else:
            lines.append(f"      coefficients: {_r(c)}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_static_script(path: str, model_file: str) -> None:
    """Command script: read the model, run one linear-static solve
    (Working/ScriptCommands.txt:1-16)."""
    with open(path, "w") as f:
        f.write("*readModelInput:\n")
        f.write(f"    fileName: {model_file}\n")
        f.write("*solve:\n")
        f.write("    elastic: yes\n")
        f.write("    nonLinearGeom: no\n")
        f.write("    dynamic: no\n")
