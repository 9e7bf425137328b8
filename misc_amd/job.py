"""The reference's job host on the device path: its input files, its command
script and its result writers, driving the library through misc_amd.astro.

Mirrors Working/AStrO_runJob.f90:44-376 (command loop), the readers of
Working/AStrO_input.f90 (readModelInput :27-1097, readDesignVarInput
:1099-1408, readObjectiveInput :1410-1576, readDesignVarValues :1578-1612)
and the writers of Working/AStrO_output.f90 (writeNodeResults :10-152,
writeElementResults :154-637, writeDVarValues :639-679, writeObjective
:681-758).  Commands: *readModelInput, *readDesignVarInput,
*readObjectiveInput, *readDesignVarValues, *solve (elastic / thermal,
static or dynamic), *calcObjective, *calcObjGradient, *writeNodeResults,
*writeElementResults, *writeObjective, *writeDesignVarValues.  The log goes
to jobLogFile.txt in the working directory, as the reference's does.

Where the reference's own runner cannot do what its script says, the
intended behaviour is implemented and named: *writeNodeResults is matched
(the reference compares fileLine(i1:16), which never equals the command,
D12); 'thermal: yes' enables the thermal solve (the reference only ever sets
solveThermal = 0, D13); *calcObjective, *calcObjGradient, *writeObjective,
*writeDesignVarValues and *writeElementResults are dispatched (the
reference's loop has no branch for them, D13); element displacements are
gathered by global node (D4).  The solver maxBW is the 64-bit form of the
reference's rule and the thermal scatter keeps every column (D21, D8).

Numbers are written with the reference's G18.12 edit descriptor and its
list-directed records (a leading blank, character items joined, a blank
before a number), as flang prints them.

Scope: brick8 and shell4 models (one element type per model), materials by
engineering constants, nodal loads, constraint equations; design variables
and objective terms of the brick8 path (misc_amd.astro.DesignVariables,
ObjectiveTerm); element results of brick8 elements."""

from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import astro
from .meshes import BRICK8, SHELL4, Model, MPCTerm

TYPE_CODES = {"brick8": BRICK8, "shell4": SHELL4}
_UNSUPPORTED_TYPES = ("tet4", "tet10", "wedge6", "brickIM", "shell3", "beam2", "frcFld")


class JobError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# Fortran output editing as the reference's records use it.
def g18(x: float) -> str:
    """Gw.d edit of a double with w = 18, d = 12 (Fortran 2008 13.7.5):
    F editing with d - k decimals and four trailing blanks when the value
    rounded to 12 digits lies in [0.1, 1e12), else E editing 0.dddE+ee
    (a 3-digit exponent drops the E; a minus sign displaces the optional
    leading zero, as flang prints it)."""
    x = float(x)
    if x == 0.0:  # a negative zero keeps its sign
        return ("-0.00000000000" if math.copysign(1.0, x) < 0 else " 0.00000000000") + "    "
    m = "%.11e" % abs(x)
    e = int(m.split("e")[1])
    k = e + 1  # 10^(k-1) <= |x| rounded < 10^k
    neg = x < 0
    if 0 <= k <= 12:
        s = "%#.*f" % (12 - k, abs(x))
        if neg:
            if s.startswith("0."):
                s = s[1:]
            s = "-" + s
        return s.rjust(14) + "    "
    digits = m.split("e")[0].replace(".", "")
    ex = k
    if abs(ex) <= 99:
        tail = "E%+03d" % ex
    else:
        tail = "%+04d" % ex
    body = "." + digits + tail
    return ("-" + body) if neg else ("0" + body)


def ld_real(x: float) -> str:
    """A double in a list-directed record as flang writes it: the shortest
    digits that read back to the same value, F form for 0.1 <= |x| < 1e15
    (no leading zero below 1), else d.ddE+ee."""
    x = float(x)
    if x == 0.0:
        return "0."
    r = repr(abs(x))
    mant, _, ex = r.partition("e")
    digits = mant.replace(".", "").lstrip("0")
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    e10 = (int(ex) if ex else 0) + (len(ip.lstrip("0")) - 1 if ip.strip("0") else -(len(fp) - len(fp.lstrip("0")) + 1))
    digits = digits.rstrip("0") or "0"
    sign = "-" if x < 0 else ""
    if -1 <= e10 < 15:
        if e10 >= 0:
            ipart = digits[:e10 + 1].ljust(e10 + 1, "0")
            fpart = digits[e10 + 1:]
            s = ipart + "." + fpart
        else:
            s = "." + digits
        return sign + s
    return sign + digits[0] + "." + digits[1:] + "E%+03d" % e10


def ld_line(*items) -> str:
    """One list-directed output statement as flang writes it: records of at
    most 79 characters, the first starting with a blank; adjacent character
    items joined, a blank between an item and a number.  A character item
    that does not fit fills the record and continues on the next one after a
    blank; an item that meets a full record starts the next one directly; a
    number that does not fit moves to the next record whole."""
    recs = [" "]
    prev_char = True
    for q, it in enumerate(items):
        is_char = isinstance(it, str)
        if is_char:
            text = it if (q == 0 or prev_char) else " " + it
            while text:
                room = 79 - len(recs[-1])
                if room <= 0:
                    recs.append("")
                    continue
                recs[-1] += text[:room]
                text = text[room:]
                if text:
                    recs.append(" ")
        else:
            text = " " + (str(int(it)) if isinstance(it, (int, np.integer)) else ld_real(it))
            if len(recs[-1]) + len(text) > 79:
                recs.append("")
            recs[-1] += text
        prev_char = is_char
    return "\n".join(recs)


# ---------------------------------------------------------------------------
# The YAML-like input language (Working/InputTemplate.yaml): indentation
# nests blocks, "key: value" lines, "- item" lists whose items are values,
# flow lists "[a, b]" or blocks.  Comments start with '#'.
_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?$")


def _value(tok: str):
    tok = tok.strip()
    if tok.startswith("[") and tok.endswith("]"):
        inner = tok[1:-1].strip()
        return [_value(t) for t in inner.split(",")] if inner else []
    if _NUM.match(tok):
        t = tok.replace("d", "e").replace("D", "e")
        if re.match(r"^[+-]?\d+$", t):
            return int(t)
        return float(t)
    return tok


def parse_blocks(text: str):
    """Nested dicts / lists of the input language."""
    lines = []
    for raw in text.splitlines():
        raw = raw.replace("\t", "    ")
        s = raw.split("#", 1)[0].rstrip()
        if not s.strip():
            continue
        lines.append((len(s) - len(s.lstrip(" ")), s.strip()))
    pos = 0

    def block(indent):
        nonlocal pos
        node = None
        while pos < len(lines):
            ind, s = lines[pos]
            if ind < indent:
                break
            if s.startswith("- ") or s == "-":
                if node is None:
                    node = []
                if not isinstance(node, list):
                    break
                item = s[1:].strip()
                pos += 1
                if ":" in item and not item.startswith("["):
                    # a mapping item: its first key on this line, the rest below
                    k, _, v = item.partition(":")
                    d = {}
                    sub_ind = ind + 2
                    if v.strip():
                        d[k.strip()] = _value(v)
                    else:
                        d[k.strip()] = block(sub_ind + 1)
                    more = block(sub_ind)
                    if isinstance(more, dict):
                        d.update(more)
                    node.append(d)
                else:
                    node.append(_value(item))
                continue
            if node is None:
                node = {}
            if not isinstance(node, dict):
                break
            k, _, v = s.partition(":")
            pos += 1
            if v.strip():
                node[k.strip()] = _value(v)
            else:
                node[k.strip()] = block(ind + 1)
        return node

    return block(0) or {}


def _labels(spec, count):
    """A set's labels: 'all', ids, or [a, b(, step)] runs (inclusive)."""
    if spec == "all":
        return np.arange(1, count + 1, dtype=np.int32)
    out = []
    for it in spec if isinstance(spec, list) else [spec]:
        if isinstance(it, list):
            a, b = int(it[0]), int(it[1])
            st = int(it[2]) if len(it) > 2 else 1
            out.extend(range(a, b + 1, st))
        else:
            out.append(int(it))
    return np.asarray(out, dtype=np.int32)


# ---------------------------------------------------------------------------
@dataclass
class ModelInfo:
    """What the reference keeps beside the arrays the library takes: set
    names, element sections, material names and strengths, load blocks with
    their activeTime windows, and the initial state."""

    node_sets: Dict[str, np.ndarray] = field(default_factory=dict)
    element_sets: Dict[str, np.ndarray] = field(default_factory=dict)
    el_section: Optional[np.ndarray] = None
    el_material: Optional[np.ndarray] = None
    materials: List[dict] = field(default_factory=list)
    load_blocks: List[dict] = field(default_factory=list)
    initial: Dict[str, np.ndarray] = field(default_factory=dict)


def read_model(path: str):
    """readModelInput: the model file into a misc_amd.meshes.Model plus the
    ModelInfo the writers and the other readers use."""
    with open(path) as f:
        tree = parse_blocks(f.read())
    nodes_in = tree.get("nodes") or []
    if not nodes_in:
        raise JobError("model file has no nodes")
    lab = np.array([int(r[0]) for r in nodes_in])
    N = int(lab.max())
    nodes = np.zeros((N, 3))
    nodes[lab - 1] = np.array([[float(v) for v in r[1:4]] for r in nodes_in])
    conns, types, elab = [], [], []
    for blk in tree.get("elements") or []:
        t = blk.get("type")
        if t not in TYPE_CODES:
            raise JobError(f"element type {t} is not on the device path" if t in _UNSUPPORTED_TYPES
                           else f"unknown element type {t}")
        nn = 8 if TYPE_CODES[t] == BRICK8 else 4
        for r in blk.get("connectivity") or []:
            elab.append(int(r[0]))
            c = [int(v) for v in r[1:1 + nn]] + [0] * (8 - nn)
            conns.append(c)
            types.append(TYPE_CODES[t])
    E = int(max(elab))
    order = np.argsort(elab)
    if sorted(elab) != list(range(1, E + 1)):
        raise JobError("element labels must run 1..numEls")
    elements = np.asarray(conns, np.int32)[order]
    etype = np.asarray(types, np.int32)[order]
    info = ModelInfo()
    sets = tree.get("sets") or {}
    for s in (sets.get("node") or []):
        info.node_sets[str(s["name"])] = _labels(s.get("labels", []), N)
    for s in (sets.get("element") or []):
        info.element_sets[str(s["name"])] = _labels(s.get("labels", []), E)
    mats = []
    for mdef in tree.get("materials") or []:
        el = mdef.get("elastic") or {}
        if "stiffness" in el:
            raise JobError("materials given by a stiffness matrix are not on the device path")
        Ev = el.get("E", [0.0, 0.0, 0.0])
        nu = el.get("nu")
        G = el.get("G")
        Ev = [float(v) for v in (Ev if isinstance(Ev, list) else [Ev] * 3)]
        if nu is None and G is not None:
            G = [float(v) for v in (G if isinstance(G, list) else [G] * 3)]
            # nu from E and G (Working/AStrO_input.f90:811 writes 0.5 E / G + 1, D14)
            nu = [Ev[0] / (2.0 * G[0]) - 1.0, Ev[0] / (2.0 * G[1]) - 1.0, Ev[1] / (2.0 * G[2]) - 1.0]
        nu = [float(v) for v in (nu if isinstance(nu, list) else [nu if nu is not None else 0.0] * 3)]
        if G is None:
            G = [Ev[0] / (2.0 * (1.0 + nu[0])), Ev[0] / (2.0 * (1.0 + nu[1])), Ev[1] / (2.0 * (1.0 + nu[2]))]
        G = [float(v) for v in (G if isinstance(G, list) else [G] * 3)]
        th = mdef.get("thermal") or {}
        k = th.get("conductivity", [0.0] * 6)
        k = [float(v) for v in (k if isinstance(k, list) else [k, k, k, 0.0, 0.0, 0.0])] + [0.0] * 6
        # expansion: six components, or one value for 11, 22, 33 (Working/AStrO_input.f90:751-760)
        x = th.get("expansion", [0.0] * 6)
        x = [float(v) for v in (x if isinstance(x, list) else [x, x, x, 0.0, 0.0, 0.0])] + [0.0] * 6
        fl = mdef.get("failure") or {}
        ms = fl.get("maxStress") or {}
        mn = fl.get("maxStrain") or {}
        mats.append(dict(name=str(mdef.get("name")), eprops=np.array(Ev + nu + G), density=float(mdef.get("density", 0.0)),
                         cond=np.array(k[:6]), spec_heat=float(th.get("specHeat", 0.0)), therm_exp=np.array(x[:6]),
                         max_stress=np.array([float(v) for v in (ms.get("tensile", [0] * 3) + ms.get("compressive", [0] * 3)
                                                                 + ms.get("shear", [0] * 3))]),
                         max_strain=np.array([float(v) for v in (mn.get("tensile", [0] * 3) + mn.get("compressive", [0] * 3)
                                                                 + mn.get("shear", [0] * 3))]),
                         max_strain_energy=float(fl.get("maxStrainEnergy", 0.0))))
    info.materials = mats
    mid = {m["name"]: q for q, m in enumerate(mats)}
    el_mat = np.full(E, -1, np.int32)
    el_sec = np.zeros(E, np.int32)
    nL = 0
    secs = tree.get("sections") or []
    for s in secs:
        if s.get("type") == "shell":
            nL = max(nL, len((s.get("layup") or {}).get("layers") or []))
    layers = np.zeros((E, max(nL, 1), 6)) if nL else None
    z_off = np.zeros(E) if nL else None
    orient = np.tile([1.0, 0.0, 0.0, 0.0, 1.0, 0.0], (E, 1)) if nL else None
    for q, s in enumerate(secs, start=1):
        ids = info.element_sets.get(str(s.get("elementSet")))
        if ids is None:
            raise JobError(f"section element set {s.get('elementSet')} not found")
        if s.get("type") == "shell":
            lay = s.get("layup") or {}
            for e in ids:
                for li, L in enumerate(lay.get("layers") or []):
                    mt = mats[mid[str(L[0])]]
                    layers[e - 1, li] = [mt["eprops"][0], mt["eprops"][1], mt["eprops"][3], mt["eprops"][6],
                                         float(L[1]), math.radians(float(L[2]))]
                z_off[ids - 1] = float(lay.get("zOffset", 0.0))
                if lay.get("layers"):
                    el_mat[e - 1] = mid[str(lay["layers"][0][0])]
            if "orientation" in s:
                orient[ids - 1] = np.array([float(v) for v in s["orientation"]])
        else:
            if s.get("type") not in ("solid", None):
                raise JobError(f"section type {s.get('type')} is not on the device path")
            el_mat[ids - 1] = mid[str(s.get("material"))]
        el_sec[ids - 1] = q
    if (el_mat < 0).any():
        raise JobError("elements without a section: " + str(np.nonzero(el_mat < 0)[0][:5] + 1))
    info.el_section = el_sec
    info.el_material = el_mat
    eprops = np.stack([mats[q]["eprops"] for q in el_mat])
    m = Model(name=os.path.basename(path), nodes=nodes, elements=elements, etype=etype, eprops=eprops,
              density=np.array([mats[q]["density"] for q in el_mat]),
              cond=np.stack([mats[q]["cond"] for q in el_mat]),
              spec_heat=np.array([mats[q]["spec_heat"] for q in el_mat]))
    tex = np.stack([mats[q]["therm_exp"] for q in el_mat])
    if np.any(tex != 0.0):
        m.therm_exp = tex
    if nL:
        m.layers, m.z_offset, m.orient = layers, z_off, orient
    m.node_sets = dict(info.node_sets)
    m.element_sets = dict(info.element_sets)
    cons = tree.get("constraints") or {}
    for r in cons.get("matrix") or []:
        who = r[1]
        if isinstance(who, int):
            m.mpc.append(MPCTerm(eqn=int(r[0]), nodes=np.array([who], np.int32), dof=int(r[2]), coef=float(r[3]),
                                 is_set=False))
        else:
            m.mpc.append(MPCTerm(eqn=int(r[0]), nodes=info.node_sets[str(who)], dof=int(r[2]), coef=float(r[3]),
                                 is_set=True, set_name=str(who)))
    m.mpc_rhs = [float(v) for v in (cons.get("rhs") or [])]
    for blk in tree.get("loads") or []:
        t = blk.get("type")
        at = blk.get("activeTime", 0.0)
        win = (float(at[0]), float(at[1])) if isinstance(at, list) else (float(at), 1e100)
        rows = []
        for r in blk.get("setLoads") or []:
            who = r[0]
            vals = [float(v) for v in r[1:8]] + [0.0] * max(0, 7 - len(r[1:8]))
            ids = [int(who)] if isinstance(who, int) else list(info.node_sets[str(who)])
            rows.extend((nd, vals) for nd in ids)
        info.load_blocks.append(dict(type=t, window=win, rows=rows))
    st = tree.get("initialState") or {}
    for key in ("displacement", "velocity", "acceleration"):
        if st.get(key):
            info.initial[key] = np.array([[float(v) for v in r] for r in st[key]])
    return m, info


@dataclass
class DesignInfo:
    """The design variables as readDesignVarInput leaves them."""

    category: List[str]
    subcat: List[str]
    component: List[int]
    layer: List[int]
    act_time: np.ndarray
    el: Dict[int, list]      # element (0-based) -> [(variable 1-based, coef)]
    nd: Dict[int, list]      # node (0-based) -> [(variable 1-based, coef)]


def read_design_vars(path: str, info: ModelInfo, num_els: int, num_nodes: int) -> DesignInfo:
    with open(path) as f:
        tree = parse_blocks(f.read())
    dvs = tree.get("designVariables") or []
    cat, sub, comp, lay, act = [], [], [], [], []
    el: Dict[int, list] = {}
    nd: Dict[int, list] = {}
    for k, v in enumerate(dvs, start=1):
        cat.append(str(v.get("category")))
        sub.append(str(v.get("subCategory", "")))
        comp.append(int(v.get("component", 0)))
        lay.append(int(v.get("layer", 0)))
        at = v.get("activeTime", 0.0)
        act.append((float(at[0]), float(at[1])) if isinstance(at, list) else (float(at), 1e100) if "activeTime" in v
                   else (0.0, 0.0))
        coefs = v.get("coefficients", 1.0)
        if "elementSet" in v:
            ids = info.element_sets[str(v["elementSet"])]
            cs = np.broadcast_to(np.asarray(coefs, float), (len(ids),)) if not isinstance(coefs, list) \
                else np.asarray(coefs, float)
            for e, c in zip(ids, cs):
                lst = el.setdefault(int(e) - 1, [])
                if all(q != k for q, _ in lst):  # a variable listed twice keeps its first coefficient
                    lst.append((k, float(c)))
        if "nodeSet" in v:
            ids = info.node_sets[str(v["nodeSet"])]
            cs = np.broadcast_to(np.asarray(coefs, float), (len(ids),)) if not isinstance(coefs, list) \
                else np.asarray(coefs, float)
            for n_, c in zip(ids, cs):
                lst = nd.setdefault(int(n_) - 1, [])
                if all(q != k for q, _ in lst):
                    lst.append((k, float(c)))
    return DesignInfo(cat, sub, comp, lay, np.array(act).reshape(-1, 2), el, nd)


_ELASTIC_SUB = {"modulus": "modulus", "poissonRatio": "poissonRatio", "shearModulus": "shearModulus"}


def design_for_library(d: DesignInfo, num_els: int, num_nodes: int, time: float) -> astro.DesignVariables:
    """The library's form of the variables at `time` (node loads outside
    their activeTime window carry coefficient 0)."""
    cats = []
    for c, s in zip(d.category, d.subcat):
        if c == "elastic" and s in _ELASTIC_SUB:
            cats.append(_ELASTIC_SUB[s])
        elif c in ("density", "thermalCond", "nodeCoord"):
            cats.append(c)
        elif c == "nodeLoad" and s == "elastic":
            cats.append("nodeLoad")
        else:
            raise JobError(f"design category {c} / {s} is not on the device path")

    def ranges(lists, count, is_node):
        rng = [0]
        ids, cs = [], []
        for i in range(count):
            for k, c in lists.get(i, []):
                if is_node and cats[k - 1] == "nodeLoad":
                    t0, t1 = d.act_time[k - 1]
                    if not (t0 <= time <= t1):
                        c = 0.0
                ids.append(k)
                cs.append(c)
            rng.append(len(ids))
        return np.array(rng, np.int32), np.array(ids, np.int32), np.array(cs)

    er, eid, ec = ranges(d.el, num_els, False)
    nr, nid, nc = ranges(d.nd, num_nodes, True)
    comp = [max(1, c) for c in d.component]
    return astro.DesignVariables(category=cats, component=comp, elToDRange=er, elToD=eid, elToCoef=ec, ndToDRange=nr,
                                 ndToD=nid, ndToCoef=nc)


def apply_design_values(model: Model, info: ModelInfo, d: DesignInfo, values, time: float) -> Model:
    """A copy of the model with every property moved by coef * r_dVec
    (r_getMaterialStiffness, r_getMaterialDensity, r_getMaterialThermCond,
    r_getNodeCoord, r_getNodeDLoad)."""
    import copy

    m = copy.deepcopy(model)
    if values is None or not np.any(values):
        return m
    lib = design_for_library(d, m.num_els, m.num_nodes, time)
    idx = {"modulus": 0, "poissonRatio": 3, "shearModulus": 6}
    for e, lst in d.el.items():
        for k, c in lst:
            cat, cp = lib.category[k - 1], lib.component[k - 1] - 1
            if cat in idx:
                m.eprops[e, idx[cat] + cp] += c * values[k - 1]
            elif cat == "density":
                m.density[e] += c * values[k - 1]
            elif cat == "thermalCond":
                m.cond[e, cp] += c * values[k - 1]
    extra = {}
    for n_, lst in d.nd.items():
        for k, c in lst:
            cat, cp = lib.category[k - 1], lib.component[k - 1] - 1
            if cat == "nodeCoord":
                m.nodes[n_, cp] += c * values[k - 1]
            elif cat == "nodeLoad":
                t0, t1 = d.act_time[k - 1]
                if t0 <= time <= t1:
                    extra.setdefault(n_, np.zeros(7))[cp] += c * values[k - 1]
    if extra:
        ids = np.array([n + 1 for n in extra], np.int32)
        vals = np.stack([extra[n] for n in extra])
        m.load_nodes = np.concatenate([m.load_nodes, ids]).astype(np.int32)
        m.load_vals = np.vstack([m.load_vals.reshape(-1, 7), vals])
    return m


def model_at(model: Model, info: ModelInfo, design: Optional[DesignInfo], values, time: float) -> Model:
    """The model a step at `time` solves: the nodal load blocks whose
    activeTime window holds `time` (the reference adds a block when
    loadsActTime(1) <= time <= loadsActTime(2)), then the design values
    applied to every property they move."""
    import copy

    m = copy.deepcopy(model)
    rows = {}
    for b in info.load_blocks:
        if b["type"] != "nodal":
            raise JobError(f"load type {b['type']} is not on the device path")
        if b["window"][0] <= time <= b["window"][1]:
            for nd, vals in b["rows"]:
                rows.setdefault(nd, np.zeros(7))
                rows[nd] += vals
    ids = np.array(sorted(rows), np.int32)
    m.load_nodes = ids
    m.load_vals = np.stack([rows[i] for i in ids]) if rows else np.zeros((0, 7))
    if design is not None:
        m = apply_design_values(m, info, design, values, time)
    return m


@dataclass
class ObjTermInfo:
    category: str
    operator: str
    act_time: tuple
    component: int
    layer: int
    coef: float
    expo: float
    el_set: str
    nd_set: str
    tgt_tag: str
    tgt: list


def read_objective(path: str) -> List[ObjTermInfo]:
    """readObjectiveInput, with its defaults: coefficient and exponent 0,
    activeTime [0, 0], target 'none'."""
    with open(path) as f:
        tree = parse_blocks(f.read())
    out = []
    for t in tree.get("objectiveTerms") or []:
        at = t.get("activeTime", None)
        act = (float(at[0]), float(at[1])) if isinstance(at, list) else ((float(at), 1e100) if at is not None
                                                                        else (0.0, 0.0))
        tv = t.get("targetValue", None)
        if tv is None:
            tag, tg = "none", []
        elif isinstance(tv, str) and tv.startswith("max"):
            tag, tg = tv, []
        elif isinstance(tv, list):
            tag, tg = "vector", [float(v) for v in tv]
        else:
            tag, tg = "single", [float(tv)]
        out.append(ObjTermInfo(str(t.get("category")), str(t.get("operator", "")), act, int(t.get("component", 0)),
                               int(t.get("layer", 0)), float(t.get("coefficient", 0.0)), float(t.get("exponent", 0.0)),
                               str(t.get("elementSet", "")), str(t.get("nodeSet", "")), tag, tg))
    return out


def read_design_var_values(path: str, num_dv: int, log=None) -> np.ndarray:
    """readDesignVarValues: lines '- [label, value]'; labels above numDVar
    are reported and skipped."""
    v = np.zeros(num_dv)
    with open(path) as f:
        for ln in f:
            if "-" in ln and "[" in ln:
                inner = ln[ln.index("[") + 1:ln.index("]")]
                a, b = inner.split(",")[:2]
                k = int(a)
                if k <= num_dv:
                    v[k - 1] = float(b.replace("d", "e").replace("D", "e"))
                elif log:
                    log(f"Error: Design variable label {k} in value input file exceeds the size of the design "
                        "variable vector.")
    return v


# ---------------------------------------------------------------------------
class Job:
    """A job in a working directory: the state the reference keeps in its
    global modules, and the commands of its script."""

    def __init__(self, workdir: str = "."):
        self.wd = workdir
        self.model: Optional[Model] = None
        self.info: Optional[ModelInfo] = None
        self.design: Optional[DesignInfo] = None
        self.dvec: Optional[np.ndarray] = None
        self.terms: List[ObjTermInfo] = []
        self.obj_val: Optional[np.ndarray] = None
        self.dLdD: Optional[np.ndarray] = None
        self.state: Dict[str, np.ndarray] = {}
        self.history: List[Dict[str, np.ndarray]] = []
        self.num_steps = 0
        self.delT = 0.0
        self.load_time = 0.0
        self.ctx = None
        self.tctx = None
        self._log = open(os.path.join(workdir, "jobLogFile.txt"), "w")
        self.log(" Beginning Job")

    def log(self, msg: str):
        self._log.write(msg + "\n")
        self._log.flush()

    def close(self):
        for c in (self.ctx, self.tctx):
            if c is not None:
                c.close()
        self.ctx = self.tctx = None
        self._log.close()

    def path(self, name):
        return name if os.path.isabs(name) else os.path.join(self.wd, name)

    # -- script -------------------------------------------------------------
    def run_script(self, script: str):
        """Parse the command script and run its commands in order."""
        with open(self.path(script)) as f:
            text = f.read()
        self.log(f" Script name: {script}")
        cmds = []
        cur = None
        body: List[str] = []
        for ln in text.splitlines():
            s = ln.strip()
            if s.startswith("*"):
                if cur:
                    cmds.append((cur, body))
                cur, _, rest = s.partition(":")
                body = []
            elif cur is not None:
                body.append(ln)
        if cur:
            cmds.append((cur, body))
        for name, lines in cmds:
            args = parse_blocks("\n".join(lines)) if lines else {}
            fn = getattr(self, "cmd_" + name[1:], None)
            if fn is None:
                self.log(f" Warning: command {name} is not part of this job runner; skipped")
                continue
            fn(args if isinstance(args, dict) else {})

    # -- readers ----------------------------------------------------------------
    def cmd_readModelInput(self, a):
        self.log(f" reading model input, file: {a['fileName']}")
        self.model, self.info = read_model(self.path(a["fileName"]))

    def cmd_readDesignVarInput(self, a):
        self._need_model()
        self.log(f" reading design variable input, file: {a['fileName']}")
        self.design = read_design_vars(self.path(a["fileName"]), self.info, self.model.num_els, self.model.num_nodes)
        self.dvec = np.zeros(len(self.design.category))

    def cmd_readObjectiveInput(self, a):
        self.log(f" reading objective input, file: {a['fileName']}")
        self.terms = read_objective(self.path(a["fileName"]))

    def cmd_readDesignVarValues(self, a):
        if self.design is None:
            raise JobError("*readDesignVarValues before *readDesignVarInput")
        self.dvec = read_design_var_values(self.path(a["fileName"]), len(self.design.category), self.log)

    def _need_model(self):
        if self.model is None:
            raise JobError("command before *readModelInput")

    # -- solve ------------------------------------------------------------------
    def _model_at(self, time):
        return model_at(self.model, self.info, self.design, self.dvec, time)

    def _options(self, coupled=False):
        return astro.Options(literalMaxBW=False, thermLiteralD8=False, rtol=1e-12, coupledProfile=coupled)

    def cmd_solve(self, a):
        self._need_model()
        yes = lambda k, d: str(a.get(k, d)).strip().lower().startswith("y")  # noqa: E731
        elastic = yes("elastic", "yes")
        thermal = yes("thermal", "no")
        dynamic = yes("dynamic", "no")
        self.load_time = float(a.get("staticLoadTime", 0.0) or 0.0)
        self.delT = float(a.get("timeStep", 0.0) or 0.0)
        beta = float(a.get("newmarkBeta", 0.25) or 0.25)
        gamma = float(a.get("newmarkGamma", 0.5) or 0.5)
        period = float(a.get("simPeriod", 0.0) or 0.0)
        save = yes("saveSolnHist", "no")
        nonlinear = yes("nonLinearGeom", "no")
        shell = self.model is not None and bool(np.all(self.model.etype == SHELL4))
        self.history = []
        for c in (self.ctx, self.tctx):
            if c is not None:
                c.close()
        self.ctx = self.tctx = None
        if not dynamic:
            self.num_steps = 0
            m = self._model_at(self.load_time)
            # stepSolve's order (Working/AStrO_commandFunctions.f90:689-747): the
            # temperatures first, then the displacements, whose element stress
            # C (strain - T alpha) reads them (one context holds both)
            if thermal:
                self.tctx = astro.Context(m, self._options())
                self.tctx.assemble_thermal()
                self.tctx.thermal_solve()
                self._grab_thermal()
            if elastic:
                # shells factor over the coupled profile: the reference's
                # component-block factor leaves GMRES short of the tolerance on
                # thin shells (1000 cycles on a 1 x 0.1 x 0.002 strip)
                self.ctx = self.tctx if self.tctx is not None else astro.Context(m, self._options(shell))
                self._initial(self.ctx)
                if nonlinear:
                    # stepSolve's Newton loop with nonLinearGeom: at most 20 steps
                    # (Working/AStrO_commandFunctions.f90:706-747)
                    self.ctx.nonlinear_solve(max_it=20, tol=1e-12)
                else:
                    self.ctx.static_solve()
                self._grab_elastic()
            return
        if self.delT <= 0.0 or period <= 0.0:
            raise JobError("a dynamic *solve needs timeStep and simPeriod")
        n, t = 0, 0.0
        while t < period:  # the reference's step loop (Working/AStrO_commandFunctions.f90:791-794)
            t += self.delT
            n += 1
        self.num_steps = n
        m = self._model_at(0.0)
        if elastic:
            # one load vector scaled per step by the activeTime windows of its blocks
            wins = {b["window"] for b in self.info.load_blocks}
            if len(wins) > 1:
                raise JobError("dynamic loads with more than one activeTime window are not on the device path")
            win = wins.pop() if wins else (0.0, 1e100)
            scales = np.array([1.0 if win[0] <= (k + 1) * self.delT <= win[1] else 0.0 for k in range(n)])
            self.dyn_scales = scales
            mm = self._model_at(win[0])
            self.ctx = astro.Context(mm, self._options())
            self.ctx.dynamic_init(beta, gamma, self.delT)
            if save:
                for k in range(n):
                    self.ctx.dynamic_step(float(scales[k]))
                    self._grab_elastic()
                    self.history.append(dict(self.state))
            else:
                self.ctx.dynamic_run(scales)
                self._grab_elastic()
        if thermal:
            self.tctx = astro.Context(m, self._options())
            self.tctx.assemble_thermal()
            self.tctx.thermal_dynamic_init(gamma, self.delT)
            if save:
                for k in range(n):
                    self.tctx.thermal_dynamic_run(1)
                    self._grab_thermal()
                    if len(self.history) < k + 1:
                        self.history.append({})
                    self.history[k].update({q: self.state[q] for q in ("temperature", "tdot")})
            else:
                self.tctx.thermal_dynamic_run(n)
                self._grab_thermal()
        if save:
            self._write_history(n)

    def _write_history(self, n):
        """writeBinarySolution after every step and for the start (solve,
        Working/AStrO_commandFunctions.f90:786-812): the previous-state
        vectors of step k are the solution of step k (zero at the start)."""
        c = self.ctx or self.tctx
        N, dim = self.model.num_nodes, c.scalar("elMatDim")
        z, zn = np.zeros(dim), np.zeros(N)
        for k in range(n + 1):
            h = self.history[k - 1] if k else {}
            write_binary_solution(self.path(history_file_name(k)),
                                  [h.get("temperature", zn), h.get("tdot", zn), h.get("displacement", z),
                                   h.get("velocity", z), h.get("acceleration", z)])

    def _initial(self, ctx):
        disp = self.info.initial.get("displacement")
        if disp is None:
            return
        nd = ctx.geti("nDofIndex").reshape(-1, 6)
        u = np.zeros(ctx.scalar("elMatDim"))
        for r in disp:
            for c in range(6):
                d = nd[int(r[0]) - 1, c]
                if d:
                    u[d - 1] = r[1 + c]
        ctx.setd("nodeDisp", u)

    def _grab_elastic(self):
        c = self.ctx
        self.nDof = c.geti("nDofIndex").reshape(-1, 6)
        self.rank = c.geti("currentRank")
        self.state["displacement"] = c.getd("nodeDisp")
        if self.num_steps:
            self.state["velocity"] = c.getd("nodeVel")
            self.state["acceleration"] = c.getd("nodeAcc")
        c.elastic_soln_load()
        self.state["elasticLoad"] = c.getd("elasticLoad")

    def _grab_thermal(self):
        c = self.tctx
        self.rank = c.geti("currentRank")
        self.state["temperature"] = c.getd("nodeTemp")
        if self.num_steps:
            self.state["tdot"] = c.getd("nodeTdot")
        c.thermal_soln_load()
        self.state["thermalLoad"] = c.getd("thermalLoad")

    # -- objective ----------------------------------------------------------------
    def _lib_terms(self, time):
        """The library's terms of those active at time (time None: every
        term, for the Newmark gradient, which applies the windows per step)."""
        out = []
        for t in self.terms:
            if time is not None and not (t.act_time[0] <= time <= t.act_time[1]):
                continue
            if t.tgt_tag.startswith("max"):
                raise JobError(f"target {t.tgt_tag} is not on the device path")
            if t.category in ("displacement", "temperature"):
                items = self.info.node_sets[t.nd_set]
            else:
                items = self.info.element_sets[t.el_set]
            tg = None if t.tgt_tag == "none" else (t.tgt[0] if t.tgt_tag == "single" else np.array(t.tgt))
            out.append(astro.ObjectiveTerm(t.category, items, operator=t.operator or "powerNorm",
                                           component=max(1, t.component), coef=t.coef, expo=t.expo, target=tg))
        return out

    def _term_ctx(self):
        thermal_only = all(t.category == "temperature" for t in self.terms)
        c = self.tctx if thermal_only and self.tctx is not None else (self.ctx or self.tctx)
        if c is None:
            raise JobError("objective before *solve")
        return c

    def cmd_calcObjective(self, a):
        time = float(a.get("staticLoadTime", self.load_time) or self.load_time)
        c = self._term_ctx()
        vals = np.zeros(len(self.terms))
        active = [i for i, t in enumerate(self.terms) if t.act_time[0] <= time <= t.act_time[1]]
        lib = self._lib_terms(time)
        if lib:
            vals[active] = c.objective(lib)
        self.obj_val = vals

    def cmd_calcObjGradient(self, a):
        if self.design is None:
            raise JobError("*calcObjGradient without design variables")
        if self.num_steps:
            return self._dynamic_gradient()
        time = float(a.get("staticLoadTime", self.load_time) or self.load_time)
        c = self._term_ctx()
        c.set_design_vars(design_for_library(self.design, self.model.num_els, self.model.num_nodes, time))
        lib = self._lib_terms(time)
        self.dLdD = c.objective_gradient(lib) if lib else np.zeros(len(self.design.category))

    def _dynamic_gradient(self):
        """getTotaldLdD's dynamic branch (Working/AStrO_commandFunctions.f90:
        1416-1464) on the device from the solution history the solve kept
        (saveSolnHist; the reference reads it back from its binary files)."""
        if self.ctx is None or not self.history or "displacement" not in self.history[0]:
            raise JobError("*calcObjGradient of a dynamic solve needs the elastic solution history (saveSolnHist)")
        c = self.ctx
        # node-load variables follow the loads' first step of activity
        c.set_design_vars(design_for_library(self.design, self.model.num_els, self.model.num_nodes, self.delT))
        lib = self._lib_terms(None)
        if not lib:
            self.dLdD = np.zeros(len(self.design.category))
            return
        n = c.scalar("elMatDim")
        U = np.vstack([np.zeros(n)] + [h["displacement"] for h in self.history])
        A = np.vstack([np.zeros(n)] + [h["acceleration"] for h in self.history])
        windows = np.array([t.act_time for t in self.terms], np.float64)
        _, self.dLdD = c.dynamic_objective_gradient(lib, U, A, load_scale=self.dyn_scales, windows=windows)

    # -- writers ----------------------------------------------------------------
    def cmd_writeNodeResults(self, a):
        fields = a.get("fields") or []
        ns = a.get("nodeSet")
        nodes = self.info.node_sets.get(str(ns)) if ns else None
        if ns and nodes is None:
            self.log(f" Warning: no node set named {ns} was found")
            self.log(" Writing results for all nodes")
        if nodes is None:
            nodes = np.arange(1, self.model.num_nodes + 1)
        fn = self.path(str(a.get("fileName", "nodeResults.yaml")))
        ts = a.get("timeSteps")
        if ts is None:
            write_node_results(fn, fields, nodes, self.state, self.nDof if hasattr(self, "nDof") else None,
                               self.rank, self.num_steps * self.delT, self.num_steps, self.log)
            return
        if not self.num_steps:
            self.log(" Error: A dynamic analysis must be run prior to writing time step dependent results.")
            return
        steps = range(1, self.num_steps + 1) if ts == "all" else sorted(
            set(k for it in (ts if isinstance(ts, list) else [ts])
                for k in (range(it[0], it[1] + 1, it[2] if len(it) > 2 else 1) if isinstance(it, list) else [it])))
        root, ext = os.path.splitext(fn)
        for k in steps:
            if k > len(self.history):
                self.log(f" Error: time step {k} was never recorded in the solution history.")
                continue
            write_node_results(f"{root}_step{k}{ext}", fields, nodes, self.history[k - 1],
                               self.nDof if hasattr(self, "nDof") else None, self.rank,
                               self.num_steps * self.delT, self.num_steps, self.log)

    def cmd_writeElementResults(self, a):
        fields = a.get("fields") or []
        es = a.get("elementSet")
        els = self.info.element_sets.get(str(es)) if es else np.arange(1, self.model.num_els + 1)
        fn = self.path(str(a.get("fileName", "elementResults.yaml")))
        data = {}
        for f in fields:
            if f in ("stress", "strain", "strainEnergyDen") and "stress" not in data:
                st, sn = [], []
                for gp in range(8):
                    s3 = [(1 if (gp >> ax) & 1 else -1) / math.sqrt(3.0) for ax in range(3)]
                    a_, b_ = self.ctx.element_stress(s3)
                    st.append(a_)
                    sn.append(b_)
                data["stress"], data["strain"] = np.stack(st, 1), np.stack(sn, 1)
            if f in ("heatFlux", "tempGradient") and "heatFlux" not in data:
                fl, gr = [], []
                for gp in range(8):
                    s3 = [(1 if (gp >> ax) & 1 else -1) / math.sqrt(3.0) for ax in range(3)]
                    a_, b_ = self.tctx.element_flux(s3)
                    fl.append(a_)
                    gr.append(b_)
                data["heatFlux"], data["tempGradient"] = np.stack(fl, 1), np.stack(gr, 1)
        write_element_results(fn, fields, els, data, self.info, self.log)

    def cmd_writeObjective(self, a):
        inc = a.get("include") or []
        grad = str(a.get("writeGradient", "no")).lower().startswith("y")
        write_objective(self.path(str(a.get("fileName", "objective.yaml"))), self.terms, self.obj_val, inc,
                        self.dLdD if grad else None)

    def cmd_writeDesignVarValues(self, a):
        inc = a.get("include") or []
        write_dvar_values(self.path(str(a.get("fileName", "designVarValues.yaml"))), self.design, self.dvec, inc)


# ---------------------------------------------------------------------------
def write_node_results(fn, fields, nodes, state, nDof, rank, time, tstep, log=None):
    """writeNodeResults (Working/AStrO_output.f90:10-152)."""
    lines = [ld_line("time: ", float(time)), ld_line("timeStep: ", int(tstep))]
    six = {"displacement": ("displacement", 1.0), "velocity": ("velocity", 1.0), "acceleration": ("acceleration", 1.0),
           "reactionForce": ("elasticLoad", -1.0)}
    one = {"temperature": ("temperature", 1.0), "tdot": ("tdot", 1.0), "reactionHeatGen": ("thermalLoad", -1.0)}
    for f in fields:
        if f in one:
            key, sg = one[f]
            v = state[key]
            lines.append(ld_line(f + ":"))
            for nd in nodes:
                # fltStr is a 128-character record: the number, then blanks
                lines.append(ld_line("    - [", str(int(nd)).ljust(16), ", ", g18(sg * v[rank[nd - 1] - 1]).ljust(128),
                                     "]"))
        elif f in six:
            key, sg = six[f]
            v = state[key]
            lines.append(ld_line(f + ":"))
            for nd in nodes:
                parts = []
                for c in range(6):
                    d = nDof[nd - 1, c]
                    parts.append(g18(sg * v[d - 1] if d > 0 else 0.0))
                lines.append(ld_line("    - [", str(int(nd)).ljust(16), ", ", ",".join(parts), "]"))
        elif log:
            log(f" Warning: Unrecognized field name {f} listed in writeNodeResults")
    with open(fn, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_element_results(fn, fields, els, data, info, log=None):
    """writeElementResults, solid branch (Working/AStrO_output.f90:154-637):
    stress, strain, strainEnergyDen, heatFlux, tempGradient at the 8
    integration points, with the maximum-stress failure index ('NA' when
    the material has no strengths)."""
    lines = []
    heads = {"stress": "S11, S22, S33, S12, S13, S23, max index ",
             "strain": "E11, E22, E33, E12, E13, E23, max index ",
             "strainEnergyDen": "strain energy density, max Index ",
             "heatFlux": "q1, q2, q3 ", "tempGradient": "Tg1, Tg2, Tg3 "}
    for f in fields:
        if f not in heads:
            if log:
                log(f" Warning: Unrecognized field name {f} listed in writeNodeResults")
            continue
        lines.append(ld_line(f + ":"))
        tri = f in ("heatFlux", "tempGradient")
        lines.append(ld_line("## element, section, integration point, " + ("" if tri else "layer, ") + heads[f]))
        for e in els:
            sec = int(info.el_section[e - 1])
            mat = info.materials[int(info.el_material[e - 1])]
            for gp in range(8):
                # the reference joins 16-character fields into a 32-character
                # intStr, so the record keeps the element and part of the section
                if tri:
                    lab = (str(e).ljust(16) + ", " + str(sec).ljust(16) + ", " + str(gp + 1).ljust(16))[:32]
                    vals = ",".join(g18(v) for v in data[f][e - 1, gp])
                    lines.append(ld_line("    - [", lab, ", ", vals, "]"))
                    continue
                lab = (str(e).ljust(16) + ", " + str(sec).ljust(16) + ", " + str(gp + 1).ljust(16) + ", "
                       + " NA".ljust(16))[:32]
                st, sn = data["stress"][e - 1, gp], data["strain"][e - 1, gp]
                if f == "strainEnergyDen":
                    sed = float(data["strainEnergyDen"][e - 1, gp]) if "strainEnergyDen" in data else \
                        float(np.sum(0.5 * st * sn))
                    ms = mat["max_strain_energy"]
                    txt = g18(sed) + "," + (g18(sed / ms) if ms != 0 else " NA               ")
                    lines.append(ld_line("    - [", lab, ", ", txt, "]"))
                    continue
                v = st if f == "stress" else sn
                idx = _failure_index(v, mat["max_stress"] if f == "stress" else mat["max_strain"])
                txt = ",".join(g18(x) for x in v) + ("," + g18(idx) if idx != 0.0 else ", NA")
                lines.append(ld_line("    - [", lab, ", ", txt, "]"))
    with open(fn, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _failure_index(s, strength):
    """Largest of s_i / tensile_i, s_i / compressive_i, |s_shear| / shear over
    the nonzero strengths (the reference divides by the compressive strength
    with its sign as given)."""
    best = 0.0
    for i in range(3):
        for j, val in ((i, s[i]), (i + 3, s[i]), (i + 6, abs(s[i + 3]))):
            if strength[j] != 0.0:
                best = max(best, val / strength[j])
    return best


def write_objective(fn, terms, values, include, grad=None):
    """writeObjective (Working/AStrO_output.f90:681-758)."""
    vals = np.zeros(len(terms)) if values is None else values
    lines = [ld_line("objective:"), ld_line(("    totalValue: ".ljust(18) + g18(float(np.sum(vals)))).ljust(128))]
    lines.append(ld_line("    termValues:"))
    for t, v in zip(terms, vals):
        lines.append(ld_line("        - value: " + g18(v)))
        for f in include:
            if f == "category":
                lines.append(ld_line("          category: ", t.category.ljust(16)))
            elif f == "operator":
                lines.append(ld_line("          operator: ", t.operator.ljust(16)))
            elif f == "component":
                lines.append(ld_line("          component: ", int(t.component)))
            elif f == "layer":
                lines.append(ld_line("          layer: ", int(t.layer)))
            elif f == "elementSet":
                lines.append(ld_line("          elementSet: ", t.el_set.ljust(64)))
            elif f == "nodeSet":
                lines.append(ld_line("          nodeSet: ", t.nd_set.ljust(64)))
            elif f == "coefficient":
                # the number overwrites the blank after the colon (writeLine(23:40))
                lines.append(ld_line("          coefficient:" + g18(t.coef)))
            elif f == "exponent":
                lines.append(ld_line("          exponent: " + g18(t.expo)))
            elif f == "activeTime":
                lines.append(ld_line("      activeTime: [" + g18(t.act_time[0]) + ", " + g18(t.act_time[1]) + "]"))
    if grad is not None:
        lines.append(ld_line("objectiveGradient:"))
        for k, g in enumerate(grad, start=1):
            # intStr and fltStr are 32-character records
            lines.append(ld_line(("    - [" + str(k).ljust(32) + ", " + g18(g).ljust(32) + "]").ljust(128)))
    with open(fn, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def history_file_name(step: int) -> str:
    """The name writeBinarySolution opens (Working/AStrO_output.f90:766-769):
    'SolutionHistory' // stepString // '.out' with stepString a 16-character
    variable holding the step left-justified, into a 32-character name, so the
    step's trailing blanks stay inside and '.out' is cut to '.' (readBinarySolution
    builds the same name, D33)."""
    return ("SolutionHistory" + f"{int(step):<16d}" + ".out")[:32].rstrip()


def write_binary_solution(path: str, arrays) -> None:
    """One record per array in Fortran unformatted sequential form (4-byte
    little-endian length before and after the doubles), as writeBinarySolution
    writes prevTemp, prevTdot, prevDisp, [prevIntDisp,] prevVel, prevAcc."""
    with open(path, "wb") as f:
        for a in arrays:
            b = np.ascontiguousarray(a, "<f8").tobytes()
            n = np.array([len(b)], "<i4").tobytes()
            f.write(n + b + n)


def read_binary_solution(path: str):
    """The records of a binary solution history file (readBinarySolution,
    Working/AStrO_input.f90:1740-1768) as float64 arrays."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    i = 0
    while i < len(data):
        n = int(np.frombuffer(data[i:i + 4], "<i4")[0])
        out.append(np.frombuffer(data[i + 4:i + 4 + n], "<f8").copy())
        if int(np.frombuffer(data[i + 4 + n:i + 8 + n], "<i4")[0]) != n:
            raise JobError(f"{path}: record markers disagree")
        i += 8 + n
    return out


def write_dvar_values(fn, design, values, include):
    """writeDVarValues (Working/AStrO_output.f90:639-679); its subCategory
    line prints the category, as the reference's does."""
    lines = [ld_line("designVariableValues:")]
    for k in range(len(design.category)):
        lines.append(ld_line("    - value: " + g18(values[k])))
        for f in include:
            if f == "category":
                lines.append(ld_line("      category: ", design.category[k].ljust(16)))
            elif f == "subCategory":
                lines.append(ld_line("      subCategory: ", design.category[k].ljust(16)))
            elif f == "component":
                lines.append(ld_line("      component: ", int(design.component[k])))
            elif f == "layer":
                lines.append(ld_line("      layer: ", int(design.layer[k])))
            elif f == "activeTime":
                t0, t1 = design.act_time[k]
                lines.append(ld_line("      activeTime: [" + g18(t0) + ", " + g18(t1) + "]"))
    with open(fn, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main(argv=None):
    """python -m misc_amd.job <script> [workdir]"""
    import sys

    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m misc_amd.job <script> [workdir]")
        return 2
    wd = argv[1] if len(argv) > 1 else os.path.dirname(os.path.abspath(argv[0]))
    job = Job(wd)
    try:
        job.run_script(os.path.abspath(argv[0]))
    finally:
        job.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
