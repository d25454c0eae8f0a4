"""SWC morphology reader.

Follows ``braincell.io.swc`` (``io/swc/reader.py:67-725``, ``rules.py``,
``soma.py``, ``types.py``): rows are parsed and normalised, the sample tree
is compressed into unbranched branches, the soma is represented the way
NEURON imports it, and the result is a :class:`~braincell_amd.morph.
Morphology` whose branch order is the order in which the branches are
discovered.

Normalisation (``rules.py:22-506``, all applied as warnings):

- type codes outside {0: custom, 1: soma, 2: axon, 3: basal_dendrite,
  4: apical_dendrite} become ``custom``;
- a radius that is not a positive number becomes 0.5 um;
- parent index 0, a parent equal to the sample or one that does not exist
  becomes a root (-1);
- a sample equal to its parent in x, y, z and radius is merged into it;
- rows are sorted by (parent, index) and renumbered 1..N.

Errors (a malformed row, duplicate indices, not exactly one root, a cycle)
raise ``ValueError`` with every problem listed.

Branch extraction (``reader.py:137-330``): a branch runs from its first
sample while there is exactly one child of the same type.  A root that
branches at once keeps its first same-type child (the others attach at
x = 0); so does the first sample of a branch under the soma.  Soma forms
(``reader.py:464-678``, ``soma.py:36-76``): one sample -> a synthetic
cylinder (centre +- r along x, radius r) with children at x = 0.5; the
three-sample NeuroMorpho soma (two samples at distance r from the centre,
equal radii) -> the same cylinder around the centre; an unbranched
multi-sample soma -> one soma branch with children at 0 / 1 / 0.5 (first /
last / inner sample); a branched soma -> one soma branch per unbranched
soma section.  A child starts at its parent's attachment sample
(``reader.py:366-433``), duplicated only when it differs from the child's
first sample in position or radius; under a soma the child keeps its own
first radius, and in ``mode="neuron"`` a multi-sample child at a soma's
midpoint does not copy the soma's centre.  ``parent_x`` of a child
attached at a sample is that sample's path-length fraction along the
parent (``reader.py:435-462``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .morph import Branch, Morphology

__all__ = ["SwcIssue", "SwcReport", "SwcReader", "read_swc", "SWC_TYPE_MAP"]

SWC_TYPE_MAP = {0: "custom", 1: "soma", 2: "axon", 3: "basal_dendrite", 4: "apical_dendrite"}
SWC_MODES = ("neuron", "neuromorpho")
MIN_SYNTHETIC_LENGTH_UM = 1e-6
_REL_TOL = 1e-6
_THREE_POINT_TOL_SCALE = 0.02


@dataclass(frozen=True)
class SwcIssue:
    level: str  # "error" | "warning"
    code: str
    message: str
    line_number: int = None
    node_id: int = None

    def __str__(self) -> str:
        where = []
        if self.line_number is not None:
            where.append(f"line {self.line_number}")
        if self.node_id is not None:
            where.append(f"node {self.node_id}")
        loc = f" ({', '.join(where)})" if where else ""
        return f"[{self.level.upper()}] {self.code}{loc}: {self.message}"


@dataclass
class SwcReport:
    issues: list = field(default_factory=list)

    def add(self, level, code, message, line_number=None, node_id=None):
        self.issues.append(SwcIssue(level, code, message, line_number, node_id))

    @property
    def errors(self):
        return [i for i in self.issues if i.level == "error"]

    @property
    def warnings(self):
        return [i for i in self.issues if i.level == "warning"]

    @property
    def has_errors(self) -> bool:
        return bool(self.errors)

    def codes(self):
        return [i.code for i in self.issues]

    def format(self, errors_only: bool = False) -> str:
        return "\n".join(str(i) for i in (self.errors if errors_only else self.issues))


@dataclass
class _Row:
    node_id: int
    type_code: int
    x: float
    y: float
    z: float
    radius: float
    parent_id: int
    line_number: int

    @property
    def point(self) -> np.ndarray:
        return np.array([self.x, self.y, self.z], dtype=np.float64)

    @property
    def type(self) -> str:
        return SWC_TYPE_MAP.get(self.type_code, "custom")


@dataclass
class _Attach:
    node_id: int = None
    point: tuple = None
    radius: float = None
    parent_x: float = None


@dataclass
class _Sec:
    point_ids: tuple
    type: str
    parent: int = None
    attach: _Attach = None
    points: tuple = None  # synthetic geometry (soma cylinders)
    radii: tuple = None


def _int_token(tok: str):
    try:
        return int(tok), False
    except ValueError:
        try:
            v = float(tok)
        except ValueError:
            return None, False
        if math.isfinite(v) and v == int(v):
            return int(v), True
        return None, False


def _float_token(tok: str):
    try:
        v = float(tok)
    except ValueError:
        return None
    return v if math.isfinite(v) else None


@dataclass(frozen=True)
class SwcReader:
    mode: str = "neuron"

    def __post_init__(self):
        if self.mode not in SWC_MODES:
            raise ValueError(f"mode must be one of {SWC_MODES!r}, got {self.mode!r}.")

    # ---------------------------------------------------------------- API
    def read(self, path, *, return_report: bool = False):
        rows, report = self._normalise(Path(path))
        if report.has_errors:
            raise ValueError(f"SWC validation failed for {path}:\n{report.format(errors_only=True)}")
        nodes = {r.node_id: r for r in rows}
        children = {r.node_id: [] for r in rows}
        roots = []
        for r in rows:
            if r.parent_id == -1:
                roots.append(r.node_id)
            else:
                children[r.parent_id].append(r.node_id)
        for c in children.values():
            c.sort()
        secs = self._extract(rows, nodes, children, roots[0])
        morph = self._build(secs, nodes)
        return (morph, report) if return_report else morph

    def check(self, path) -> SwcReport:
        return self._normalise(Path(path))[1]

    # ------------------------------------------------------- normalisation
    def _normalise(self, path: Path):
        rep = SwcReport()
        rows = []
        seen = set()
        with path.open(encoding="utf-8", errors="replace") as fh:
            lines = list(fh)
        for ln, raw in enumerate(lines, start=1):
            s = raw.strip()
            if not s or s.startswith("#"):
                continue
            f = s.split()
            if len(f) < 7:
                rep.add("error", "format.column_count", f"expected 7 columns, got {len(f)}", ln)
                continue
            nid, _ = _int_token(f[0])
            if nid is None or nid < 0:
                rep.add("error", "identity.invalid_id", f"invalid sample index {f[0]!r}", ln)
                continue
            if nid in seen:
                rep.add("error", "identity.duplicate_id", f"duplicate sample index {nid}", ln, nid)
                continue
            seen.add(nid)
            tcode, _ = _int_token(f[1])
            if tcode is None:
                rep.add("warning", "format.type_int", f"type {f[1]!r} is not an integer; set to 0", ln, nid)
                tcode = 0
            xyz = [_float_token(t) for t in f[2:5]]
            for k, v in enumerate(xyz):
                if v is None:
                    rep.add("warning", "geometry.xyz_double", f"coordinate {f[2 + k]!r} is not a number; set to 0",
                            ln, nid)
                    xyz[k] = 0.0
            r = _float_token(f[5])
            if r is None or r <= 0.0:
                rep.add("warning", "geometry.radius_positive_double", f"radius {f[5]!r} is not positive; set to 0.5",
                        ln, nid)
                r = 0.5
            pid, _ = _int_token(f[6])
            if pid is None:
                rep.add("error", "identity.invalid_parent_id", f"invalid parent index {f[6]!r}", ln, nid)
                continue
            if tcode not in SWC_TYPE_MAP:
                rep.add("warning", "semantics.unknown_type", f"type {tcode} is not mapped; read as custom", ln, nid)
                tcode = 0
            rows.append(_Row(nid, tcode, xyz[0], xyz[1], xyz[2], r, pid, ln))
        if not rows:
            rep.add("error", "format.empty_file", "SWC file has no samples.")
            return rows, rep
        ids = {r.node_id for r in rows}
        for r in rows:
            if r.parent_id == 0 or (r.parent_id != -1 and (r.parent_id == r.node_id or r.parent_id not in ids)):
                rep.add("warning", "topology.invalid_parent", f"parent {r.parent_id} is invalid; set to -1",
                        r.line_number, r.node_id)
                r.parent_id = -1
            elif r.parent_id < -1:
                rep.add("warning", "topology.invalid_parent", f"parent {r.parent_id} is invalid; set to -1",
                        r.line_number, r.node_id)
                r.parent_id = -1
        # merge samples equal to their parent in x, y, z and radius
        while True:
            by_id = {r.node_id: r for r in rows}
            dup = None
            for r in rows:
                p = by_id.get(r.parent_id)
                if p is not None and (r.x, r.y, r.z, r.radius) == (p.x, p.y, p.z, p.radius):
                    dup = (r, p)
                    break
            if dup is None:
                break
            r, p = dup
            rep.add("warning", "geometry.duplicate_xyzr_node", f"sample {r.node_id} equals its parent {p.node_id}; merged",
                    r.line_number, r.node_id)
            for c in rows:
                if c.parent_id == r.node_id:
                    c.parent_id = p.node_id
            rows = [c for c in rows if c is not r]
        if not any(r.type_code == 1 for r in rows):
            rep.add("warning", "semantics.no_soma_samples", "SWC has no soma samples.")
        # parents first (by parent index, then index), renumbered 1..N
        order = sorted(rows, key=lambda r: (r.parent_id != -1, r.parent_id if r.parent_id != -1 else -1, r.node_id))
        if [r.line_number for r in order] != [r.line_number for r in rows]:
            rep.add("warning", "topology.sorted_order", "rows reordered so parents come before children")
        rows = order
        new = {r.node_id: i for i, r in enumerate(rows, start=1)}
        if any(r.node_id != i for i, r in enumerate(rows, start=1)):
            rep.add("warning", "identity.sequential_index", "sample indices renumbered 1..N")
        for r in rows:
            r.node_id = new[r.node_id]
            if r.parent_id != -1:
                r.parent_id = new[r.parent_id]
        roots = [r for r in rows if r.parent_id == -1]
        if len(roots) != 1:
            rep.add("error", "topology.root_count", f"expected exactly one root sample, found {len(roots)}")
            return rows, rep
        # every sample reaches the root (no cycles)
        parent = {r.node_id: r.parent_id for r in rows}
        state = {}
        for r in rows:
            path_, cur = [], r.node_id
            while cur != -1 and cur not in state:
                if cur in path_:
                    rep.add("error", "topology.cycle", f"sample {cur} lies on a cycle", node_id=cur)
                    return rows, rep
                path_.append(cur)
                cur = parent[cur]
            for n in path_:
                state[n] = True
        return rows, rep

    # ----------------------------------------------------------- extraction
    def _extract(self, rows, nodes, children, root):
        if nodes[root].type != "soma":
            return self._extract_generic(nodes, children, root)
        soma = self._soma_component(nodes, children, root)
        srows = [r for r in rows if r.node_id in soma]
        if len(srows) == 1:
            secs, tasks = self._single_point_soma(srows[0], nodes, children)
        else:
            three = self._three_point_soma(srows, children)
            if three is not None:
                secs, tasks = self._single_point_soma(three, nodes, children, point_ids=tuple(r.node_id for r in srows))
            elif all(len(self._soma_kids(n, nodes, children)) <= 1 for n in soma):
                secs, tasks = self._regular_soma(tuple(r.node_id for r in srows), nodes, children)
            else:
                secs, tasks = self._branched_soma(nodes, children, root)
        return self._append_subtrees(secs, tasks, nodes, children)

    def _walk_same_type(self, start, nodes, children, ids):
        t = nodes[start].type
        cur = start
        while True:
            kids = children[cur]
            if len(kids) != 1 or nodes[kids[0]].type != t:
                return cur
            ids.append(kids[0])
            cur = kids[0]

    def _extract_generic(self, nodes, children, root):
        secs = []
        rtype = nodes[root].type
        ids = [root]
        side = ()
        kids = children[root]
        same = [c for c in kids if nodes[c].type == rtype]
        cur = root
        if len(kids) != 1 and same:
            ids.append(same[0])
            cur = same[0]
            side = tuple(c for c in kids if c != same[0])
        end = self._walk_same_type(cur, nodes, children, ids)
        secs.append(_Sec(tuple(ids), rtype))

        def walk(parent, attach, start):
            pts = [start]
            e = self._walk_same_type(start, nodes, children, pts)
            k = len(secs)
            secs.append(_Sec(tuple(pts), nodes[start].type, parent, attach))
            for c in children[e]:
                walk(k, _Attach(node_id=e), c)

        for c in side:
            walk(0, _Attach(node_id=root), c)
        for c in children[end]:
            walk(0, _Attach(node_id=end), c)
        return secs

    def _soma_component(self, nodes, children, root):
        comp, stack = set(), [root]
        while stack:
            n = stack.pop()
            if n in comp or nodes[n].type != "soma":
                continue
            comp.add(n)
            p = nodes[n].parent_id
            if p in nodes and nodes[p].type == "soma":
                stack.append(p)
            stack.extend(c for c in children[n] if nodes[c].type == "soma")
        return comp

    def _soma_kids(self, n, nodes, children):
        return [c for c in children[n] if nodes[c].type == "soma"]

    def _synthetic_soma(self, center, radius, point_ids):
        off = np.array([radius, 0.0, 0.0])
        return _Sec(point_ids, "soma", points=(tuple(center - off), tuple(center), tuple(center + off)),
                    radii=(radius, radius, radius))

    def _single_point_soma(self, row, nodes, children, point_ids=None):
        c = row.point
        r = max(row.radius, MIN_SYNTHETIC_LENGTH_UM)
        sec = self._synthetic_soma(c, r, point_ids or (row.node_id,))
        tasks = [(_Attach(node_id=row.node_id, point=tuple(c), radius=r, parent_x=0.5), k, 0)
                 for k in children[row.node_id] if nodes[k].type != "soma"]
        if point_ids is None:
            tasks = [(_Attach(point=a.point, radius=a.radius, parent_x=0.5), k, p) for a, k, p in tasks]
        return [sec], tasks

    def _three_point_soma(self, srows, children):
        """NeuroMorpho's three-sample soma: a centre and two childless
        samples that hang from it at distance r, all radii equal."""
        if len(srows) != 3:
            return None
        for cand in srows:
            others = [r for r in srows if r is not cand]
            if all(o.parent_id == cand.node_id for o in others):
                break
        else:
            return None
        if any(children.get(o.node_id) for o in others):
            return None
        tol = max(_THREE_POINT_TOL_SCALE * cand.radius, _REL_TOL)
        d = [float(np.linalg.norm(o.point - cand.point)) for o in others]
        if min(d) <= 0.0:
            return None
        radii = np.array([r.radius for r in srows])
        if not np.allclose(radii, radii[0], rtol=_REL_TOL, atol=tol):
            return None
        if not all(math.isclose(x, cand.radius, rel_tol=_REL_TOL, abs_tol=tol) for x in d):
            return None
        return srows[0]

    @staticmethod
    def _section_x(ids, node):
        if node == ids[0]:
            return 0.0
        if node == ids[-1]:
            return 1.0
        return 0.5

    def _regular_soma(self, soma_ids, nodes, children):
        tasks = []
        for n in soma_ids:
            for k in children[n]:
                if nodes[k].type != "soma":
                    tasks.append((_Attach(node_id=n, parent_x=self._section_x(soma_ids, n)), k, 0))
        return [_Sec(soma_ids, "soma")], tasks

    def _branched_soma(self, nodes, children, root):
        secs, tasks = [], []

        def section(parent, start, attach):
            ids = [start]
            cur = start
            side = ()
            if parent is None:
                sk = self._soma_kids(cur, nodes, children)
                if len(sk) > 1:
                    ids.append(sk[0])
                    cur = sk[0]
                    side = tuple(sk[1:])
            while True:
                sk = self._soma_kids(cur, nodes, children)
                if len(sk) != 1:
                    break
                ids.append(sk[0])
                cur = sk[0]
            k = len(secs)
            secs.append(_Sec(tuple(ids), "soma", parent, attach))
            eff = tuple(ids)
            if attach is not None and attach.node_id is not None and not (attach.parent_x == 0.5 and len(ids) > 1) \
                    and attach.node_id != ids[0]:
                eff = (attach.node_id,) + eff
            for n in ids:
                px = self._section_x(eff, n)
                for c in children[n]:
                    if nodes[c].type != "soma":
                        tasks.append((_Attach(node_id=n, parent_x=px), c, k))
            for c in side:
                section(k, c, _Attach(node_id=start, parent_x=0.0))
            for c in self._soma_kids(cur, nodes, children):
                section(k, c, _Attach(node_id=cur, parent_x=1.0))

        section(None, root, None)
        return secs, tasks

    def _append_subtrees(self, secs, tasks, nodes, children):
        def walk(parent, attach, start):
            btype = nodes[start].type
            pts = [start]
            cur = start
            side = ()
            if secs[parent].type == "soma":
                kids = children[cur]
                same = [c for c in kids if nodes[c].type == btype]
                if len(kids) != 1 and same:
                    pts.append(same[0])
                    cur = same[0]
                    side = tuple(c for c in kids if c != same[0])
            e = self._walk_same_type(cur, nodes, children, pts)
            k = len(secs)
            secs.append(_Sec(tuple(pts), btype, parent, attach))
            for c in side:
                walk(k, _Attach(node_id=start), c)
            for c in children[e]:
                walk(k, _Attach(node_id=e), c)

        for attach, child, parent in tasks:
            walk(parent, attach, child)
        return secs

    # -------------------------------------------------------------- building
    def _attach_x(self, parent_sec, attach, nodes):
        if attach is None:
            return 1.0
        if attach.parent_x is not None:
            return float(attach.parent_x)
        pts = [nodes[n].point for n in parent_sec.point_ids]
        if len(pts) == 1:
            return 1.0
        i = parent_sec.point_ids.index(attach.node_id)
        seg = [float(np.linalg.norm(pts[j + 1] - pts[j])) for j in range(len(pts) - 1)]
        tot = sum(seg)
        return 1.0 if tot <= 0.0 else float(sum(seg[:i]) / tot)

    def _branch(self, sec, nodes, parent_type=None):
        if sec.points is not None:
            return _branch_from_points(sec.points, sec.radii, sec.type)
        pts = [nodes[n].point for n in sec.point_ids]
        rad = [nodes[n].radius for n in sec.point_ids]
        if sec.attach is not None:
            a = sec.attach
            if a.point is not None and a.radius is not None:
                ap, ar = np.asarray(a.point, dtype=np.float64), float(a.radius)
            else:
                ap, ar = nodes[a.node_id].point, nodes[a.node_id].radius
            copy = True
            if self.mode == "neuron" and parent_type == "soma":
                copy = len(sec.point_ids) == 1 or a.parent_x != 0.5
            if parent_type == "soma":
                ar = rad[0]
            if copy and (not np.allclose(pts[0], ap) or not np.isclose(rad[0], ar)):
                pts.insert(0, ap)
                rad.insert(0, ar)
        return _branch_from_points(pts, rad, sec.type)

    def _build(self, secs, nodes):
        if not secs:
            raise ValueError("SWC extraction produced no branches.")
        morph = Morphology.from_root(self._branch(secs[0], nodes))
        for s in secs[1:]:
            p = s.parent
            morph.attach(p, self._branch(s, nodes, parent_type=secs[p].type),
                         parent_x=self._attach_x(secs[p], s.attach, nodes), child_x=0.0)
        return morph


def _branch_from_points(points, radii, btype) -> Branch:
    """Frusta between consecutive samples (``Branch.from_points``)."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(radii, dtype=np.float64)
    if p.shape[0] < 2:
        raise ValueError("a branch needs at least two samples")
    lengths = np.linalg.norm(np.diff(p, axis=0), axis=1)
    return Branch(lengths, r[:-1], r[1:], btype)


def read_swc(path, *, mode: str = "neuron", return_report: bool = False):
    """Read an SWC file into a :class:`Morphology` (``Morphology.from_swc``)."""
    return SwcReader(mode=mode).read(path, return_report=return_report)
