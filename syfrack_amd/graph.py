"""Node graphs: the reference's form graph, lowered to the C ABI's syf_graph.

Mirrors crates/syfrack-core/src/scene/graph/mod.rs for this path: a Graph holds nodes (id, kind
with its parameters, name, enabled), links from an output port to an input port, the global
safety factor and the bounding-radius hint (mod.rs:427-451).  Node kinds, trap measures and
trap modes keep the order of NodeKind, Measure and TrapMode (mod.rs:81-257), which is the
numbering of SYF_NODE_*, SYF_MEASURE_* and SYF_TRAP_*.  Parameters are addressed as
``nodes[ID].name`` with ``.x/.y/.z`` for vectors, within the bounds of mod.rs:44-73 (Params for
Graph, mod.rs:527-662).

``Graph.bundled(name)`` loads the forms shipped with the reference (graphs/*.ron, embedded by
scene/graph/bundled.rs), kept here as JSON; ``Graph.from_ron`` reads a graph file of the
reference.  Validation, the step scale, the iteration product and the trap scales come from the
library (syf_graph_info), which also compiles the graph for the device when a job is created
(syf_job_create_graph).
"""
from __future__ import annotations

import copy
import ctypes
import json
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from . import abi
from .ron import from_ron

KINDS = ("Input", "Output", "Translate", "Rotate", "Scale", "ScaleAxes", "Mirror", "BoxFold", "BoxFoldAxes",
         "SphereInversion", "SphereFold", "Abs", "MengerFold", "Kaleido", "Power", "AddOrigin", "AddSeed",
         "AddConstant", "Twist", "Bend", "Tile", "RepeatFinite", "Iterate", "Trap", "Sphere", "Torus", "Cylinder",
         "Box", "RoundBox", "Rock", "Plane", "NormFence", "BulbFence", "MaterialTag", "Bounded", "Instances",
         "Union", "Intersection", "Subtraction", "SmoothUnion", "SmoothIntersection")
MEASURES = ("Norm", "AxisX", "AxisY", "AxisZ", "PlaneY", "AngleXY", "AngleXZ", "AngleYZ", "LogNorm")
TRAP_MODES = ("Minimum", "Maximum", "Last")

# Bounds (mod.rs:44-73; DIRECTION params.rs:102).
SAFETY = (0.1, 1.0)
OFFSET = (-10.0, 10.0)
ANGLE_DEG = (-180.0, 180.0)
SCALE = (-4.0, 4.0)
AXIS_SCALE = (0.001, 100.0)
MATERIAL_INDEX = (0.0, 8.0)
BOUND_MARGIN = (0.001, 10.0)
MIRROR_OFFSET = (-10.0, 10.0)
FOLD_LIMIT = (0.5, 2.0)
AXIS_FOLD_LIMIT = (0.01, 4.0)
INVERSION_CORE = (0.0001, 2.0)
INVERSION_RADIUS = (0.01, 10.0)
MIN_RADIUS = (0.05, 1.0)
FIXED_RADIUS = (0.5, 2.0)
MENGER_SCALE = (2.0, 4.0)
MENGER_OFFSET = (0.5, 2.0)
PLANES = (2.0, 12.0)
PHASE = (-180.0, 180.0)
POWER = (2.0, 16.0)
RATE = (-2.0, 2.0)
PERIOD = (0.1, 100.0)
ITERATIONS = (1.0, 64.0)
BAILOUT = (2.0, 1.0e6)
SLOT = (0.0, 3.0)
TRAP_SCALE = (0.01, 100.0)
RADIUS = (0.01, 100.0)
HALF = (0.01, 500.0)
PLANE_OFFSET = (-100.0, 100.0)
INSET = (0.0, 2.0)
SMOOTH_K = (0.001, 10.0)
DIRECTION = (-1.0, 1.0)
EROSION = (0.0, 0.6)
ROCK_SEED = (0.0, 65535.0)
REPEAT_COUNT = (1.0, 32.0)

# Per kind, in the visiting order of Params for Graph: (field, form, slot, bounds).  Forms: "f"
# a float in v[slot], "v3" a vector in v[slot..slot+3], "u" an integer in u[slot], "u3" three
# integers in u[0..3], "b" a flag bit, "measure"/"mode" an enum index in u[slot], "body" the
# Iterate body.
SPEC: Dict[str, list] = {
    "Input": [], "Output": [], "AddOrigin": [], "BulbFence": [], "Union": [], "Intersection": [],
    "Subtraction": [],
    "Translate": [("offset", "v3", 0, OFFSET)],
    "Rotate": [("angles", "v3", 0, ANGLE_DEG)],
    "Scale": [("factor", "f", 0, SCALE)],
    "ScaleAxes": [("factors", "v3", 0, AXIS_SCALE)],
    "Mirror": [("normal", "v3", 0, DIRECTION), ("offset", "f", 3, MIRROR_OFFSET)],
    "BoxFold": [("limit", "f", 0, FOLD_LIMIT)],
    "BoxFoldAxes": [("limit", "v3", 0, AXIS_FOLD_LIMIT), ("x", "b", 0, None), ("y", "b", 1, None),
                    ("z", "b", 2, None)],
    "SphereInversion": [("radius", "f", 0, INVERSION_RADIUS), ("min_radius", "f", 1, INVERSION_CORE)],
    "SphereFold": [("min_radius", "f", 0, MIN_RADIUS), ("fixed_radius", "f", 1, FIXED_RADIUS)],
    "Abs": [("x", "b", 0, None), ("y", "b", 1, None), ("z", "b", 2, None)],
    "MengerFold": [("scale", "f", 0, MENGER_SCALE), ("offset", "v3", 1, MENGER_OFFSET)],
    "Kaleido": [("planes", "u", 0, PLANES), ("phase", "f", 0, PHASE)],
    "Power": [("power", "f", 0, POWER)],
    "AddSeed": [("multiplier", "v3", 0, SCALE)],
    "AddConstant": [("multiplier", "v3", 0, SCALE), ("julia", "v3", 3, OFFSET)],
    "Twist": [("rate", "f", 0, RATE)],
    "Bend": [("rate", "f", 0, RATE)],
    "Tile": [("period", "v3", 0, PERIOD)],
    "RepeatFinite": [("period", "v3", 0, PERIOD), ("count", "u3", 0, REPEAT_COUNT)],
    "Iterate": [("iterations", "u", 0, ITERATIONS), ("bailout", "f", 0, BAILOUT), ("body", "body", 0, None)],
    "Trap": [("slot", "u", 0, SLOT), ("measure", "measure", 1, None), ("scale", "f", 0, TRAP_SCALE),
             ("mode", "mode", 2, None)],
    "Sphere": [("radius", "f", 0, RADIUS)],
    "Torus": [("radius", "f", 0, RADIUS), ("tube", "f", 1, RADIUS)],
    "Cylinder": [("radius", "f", 0, RADIUS), ("half_height", "f", 1, HALF)],
    "Box": [("half", "v3", 0, HALF)],
    "RoundBox": [("half", "v3", 0, HALF), ("radius", "f", 3, INSET)],
    "Rock": [("half", "v3", 0, HALF), ("erosion", "f", 3, EROSION), ("seed", "u", 0, ROCK_SEED)],
    "Plane": [("normal", "v3", 0, DIRECTION), ("offset", "f", 3, PLANE_OFFSET)],
    "NormFence": [("inset", "f", 0, INSET)],
    "MaterialTag": [("index", "u", 0, MATERIAL_INDEX)],
    "Bounded": [("half", "v3", 0, HALF), ("margin", "f", 3, BOUND_MARGIN)],
    "Instances": [("half", "v3", 0, HALF), ("margin", "f", 3, BOUND_MARGIN)],
    "SmoothUnion": [("k", "f", 0, SMOOTH_K)],
    "SmoothIntersection": [("k", "f", 0, SMOOTH_K)],
}
DEFAULTS = {("Trap", "mode"): "Minimum"}

GRAPH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "graphs")


@dataclass
class Node:
    id: int
    kind: str
    params: dict = field(default_factory=dict)
    name: str = ""
    enabled: bool = True


@dataclass
class Graph:
    """A form graph (Graph, mod.rs:442-451).  links: ((from_id, from_port), (to_id, to_port))."""
    nodes: List[Node]
    links: List[Tuple[Tuple[int, int], Tuple[int, int]]]
    safety: float = 1.0
    bounding_radius_hint: float = 2.0

    # ------------------------------------------------------------ construction

    @staticmethod
    def new_minimal() -> "Graph":
        """Graph::new_minimal (mod.rs:454-486): Input -> Sphere(1) -> Output."""
        return Graph([Node(0, "Input", name="Entrée"), Node(1, "Sphere", {"radius": 1.0}, name="Sphère"),
                      Node(2, "Output", name="Sortie")], [((0, 0), (1, 0)), ((1, 0), (2, 0))], 1.0, 2.0)

    @staticmethod
    def bundled_names() -> List[str]:
        return sorted(f[:-5] for f in os.listdir(GRAPH_DIR) if f.endswith(".json"))

    @staticmethod
    def bundled(name: str) -> "Graph":
        """Graph::bundled(name) (bundled.rs): the forms shipped with the reference."""
        path = os.path.join(GRAPH_DIR, name + ".json")
        if not os.path.exists(path):
            raise KeyError(f"unknown bundled graph {name!r}")
        with open(path, encoding="utf-8") as fh:
            return Graph.from_dict(json.load(fh))

    @staticmethod
    def from_ron(text: str) -> "Graph":
        """A graph file of the reference (RON, as graphs/*.ron)."""
        return Graph.from_dict(from_ron(text))

    @staticmethod
    def from_dict(d: dict) -> "Graph":
        """From the parsed RON of a graph or the JSON this module writes."""
        nodes = []
        for nd in d["nodes"]:
            kind = nd["kind"]
            if isinstance(kind, dict):
                name = kind["__struct__"]
                params = {k: v for k, v in kind.items() if k != "__struct__"}
            elif isinstance(kind, str):
                name, params = kind, dict(nd.get("params", {}))
            else:
                raise ValueError(f"node {nd.get('id')}: unreadable kind {kind!r}")
            if name not in KINDS:
                raise ValueError(f"node {nd.get('id')}: unknown kind {name!r}")
            nodes.append(Node(int(nd["id"]), name, _norm_params(params), nd.get("name", ""), bool(nd.get("enabled", True))))
        links = []
        for l in d["links"]:
            if isinstance(l, dict):
                a, b = l["from"], l["to"]
            else:
                a, b = l
            links.append(((int(a[0]), int(a[1])), (int(b[0]), int(b[1]))))
        return Graph(nodes, links, float(d.get("safety", 1.0)), float(d.get("bounding_radius_hint", 2.0)))

    def to_dict(self) -> dict:
        return {"nodes": [{"id": n.id, "kind": n.kind, "params": _plain(n.params), "name": n.name,
                           "enabled": n.enabled} for n in self.nodes],
                "links": [[list(a), list(b)] for a, b in self.links],
                "safety": self.safety, "bounding_radius_hint": self.bounding_radius_hint}

    # ------------------------------------------------------------ access

    def node(self, nid: int) -> Optional[Node]:
        for n in self.nodes:
            if n.id == nid:
                return n
        return None

    def input_of(self, to: int, port: int) -> Optional[Tuple[int, int]]:
        for a, b in self.links:
            if b == (to, port):
                return a
        return None

    def next_id(self) -> int:
        return max((n.id + 1 for n in self.nodes), default=0)

    def copy(self) -> "Graph":
        return copy.deepcopy(self)

    _PATH = re.compile(r"^(?:fractal\.)?nodes\[(\d+)\]\.([a-z_]+)(?:\.([xyz]))?$")

    def set_param(self, path: str, value: float) -> None:
        """Scene::set_param for a graph parameter (``fractal.safety``, ``fractal.bounding_radius_hint``,
        ``fractal.nodes[ID].name[.x]``), with the graph bounds; raises ValueError naming the path."""
        p = path[len("fractal."):] if path.startswith("fractal.") else path
        if p == "safety":
            _check(path, value, SAFETY)
            self.safety = float(value)
            return
        if p == "bounding_radius_hint":
            self.bounding_radius_hint = float(value)
            return
        m = self._PATH.match(path)
        if not m:
            raise ValueError(f"{path}: not a graph parameter")
        nid, key, axis = int(m.group(1)), m.group(2), m.group(3)
        n = self.node(nid)
        if n is None:
            raise ValueError(f"{path}: no node {nid}")
        spec = {s[0]: s for s in SPEC[n.kind]}
        if key not in spec or spec[key][1] in ("b", "measure", "mode", "body"):
            raise ValueError(f"{path}: no such parameter on {n.kind}")
        _, form, _, bounds = spec[key]
        _check(path, value, bounds)
        if form in ("v3", "u3"):
            if axis is None:
                raise ValueError(f"{path}: name a component (.x, .y or .z)")
            v = list(n.params[key])
            v["xyz".index(axis)] = int(round(value)) if form == "u3" else float(value)
            n.params[key] = tuple(v) if form == "v3" else v
        elif form == "u":
            n.params[key] = int(round(value))
        else:
            n.params[key] = float(value)

    # ------------------------------------------------------------ C ABI

    def c_struct(self) -> "CGraph":
        """The syf_graph of this graph, with the arrays it points to."""
        return CGraph(self)

    def info(self) -> Tuple[float, int, Tuple[float, float, float, float]]:
        """(step scale, iteration product, trap scales) from the library; raises ValueError with
        the library's message when the graph is invalid (Graph::validate)."""
        lib = abi.load()
        cg = self.c_struct()
        step = ctypes.c_float()
        prod = ctypes.c_uint32()
        scales = (ctypes.c_float * 4)()
        fam = ctypes.c_uint32()
        rc = lib.syf_graph_info(ctypes.byref(cg.s), ctypes.byref(step), ctypes.byref(prod), scales, ctypes.byref(fam))
        if rc != abi.SYF_OK:
            raise ValueError(lib.syf_last_error().decode())
        self._kernel_family = fam.value
        return step.value, prod.value, tuple(scales)

    def kernel_family(self) -> int:
        """The classic family (abi.FAMILY_*) a specialised job of this graph runs on, or
        abi.FAMILY_GRAPH (syf_graph_info)."""
        self.info()
        return self._kernel_family

    def source(self) -> str:
        """The device source of this graph's specialised topology (syf_graph_source), the
        counterpart of the WGSL of graph_wgsl::compile."""
        lib = abi.load()
        cg = self.c_struct()
        n = ctypes.c_size_t()
        if lib.syf_graph_source(ctypes.byref(cg.s), None, 0, ctypes.byref(n)) != abi.SYF_OK:
            raise ValueError(lib.syf_last_error().decode())
        buf = ctypes.create_string_buffer(n.value)
        abi.check(lib.syf_graph_source(ctypes.byref(cg.s), buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def validate(self) -> None:
        self.info()

    def step_scale(self) -> float:
        return self.info()[0]

    def iteration_product(self) -> int:
        return self.info()[1]

    def trap_scales(self) -> Tuple[float, float, float, float]:
        return self.info()[2]


class CGraph:
    """syf_graph plus the node, link and body arrays it points to (kept alive with it)."""

    def __init__(self, g: Graph):
        n = len(g.nodes)
        self.nodes = (abi.SyfGraphNode * max(n, 1))()
        bodies: List[int] = []
        for i, nd in enumerate(g.nodes):
            c = self.nodes[i]
            c.id = nd.id
            c.kind = KINDS.index(nd.kind)
            c.enabled = 1 if nd.enabled else 0
            for key, form, slot, _ in SPEC[nd.kind]:
                val = nd.params.get(key, DEFAULTS.get((nd.kind, key)))
                if val is None:
                    raise ValueError(f"node {nd.id} ({nd.kind}): missing parameter {key!r}")
                if form == "f":
                    c.v[slot] = float(val)
                elif form == "v3":
                    for k in range(3):
                        c.v[slot + k] = float(val[k])
                elif form == "u":
                    c.u[slot] = int(val)
                elif form == "u3":
                    for k in range(3):
                        c.u[k] = int(val[k])
                elif form == "b":
                    if val:
                        c.flags |= 1 << slot
                elif form == "measure":
                    c.u[slot] = MEASURES.index(val)
                elif form == "mode":
                    c.u[slot] = TRAP_MODES.index(val)
                elif form == "body":
                    c.body_first = len(bodies)
                    c.body_count = len(val)
                    bodies.extend(int(b) for b in val)
        self.links = (abi.SyfGraphLink * max(len(g.links), 1))()
        for i, ((a, ap), (b, bp)) in enumerate(g.links):
            self.links[i].from_node, self.links[i].from_port = a, ap
            self.links[i].to_node, self.links[i].to_port = b, bp
        self.bodies = (ctypes.c_uint32 * max(len(bodies), 1))(*bodies)
        s = abi.SyfGraph()
        s.nodes = ctypes.cast(self.nodes, ctypes.c_void_p)
        s.node_count = n
        s.links = ctypes.cast(self.links, ctypes.c_void_p)
        s.link_count = len(g.links)
        s.bodies = ctypes.cast(self.bodies, ctypes.c_void_p)
        s.body_len = len(bodies)
        s.safety = g.safety
        s.bounding_radius_hint = g.bounding_radius_hint
        self.s = s

    def ref(self):
        return ctypes.byref(self.s)


def _check(path: str, value: float, bounds) -> None:
    if bounds is None:
        return
    lo, hi = bounds
    v = float(value)
    if not (math.isfinite(v) and lo <= v <= hi):
        raise ValueError(f"{path} = {value} is outside [{lo}, {hi}]")


def _norm_params(params: dict) -> dict:
    out = {}
    for k, v in params.items():
        if isinstance(v, (tuple, list)) and k not in ("body", "count", "placements"):
            out[k] = tuple(float(x) for x in v)
        elif isinstance(v, (tuple, list)):
            out[k] = list(v)
        else:
            out[k] = v
    return out


def _plain(params: dict) -> dict:
    return {k: (list(v) if isinstance(v, tuple) else v) for k, v in params.items()}
