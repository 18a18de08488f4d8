"""Structured hex meshes of FCBGA stacks from a layer/region spec, in the mesh
contract of the reference mesher.

The arrays follow `buildStructuredMesh` (src/core/06-mesher.js:14-72): nodes
exist only where an element touches them and are numbered z fastest, then y,
then x; `cell_elems` holds up to two superposed elements per cell; `col_start`
gives each (i, j) column's contiguous node range.  3-2-1 support follows
`constraints321` (src/core/06-mesher.js:112-116) and the per-group stiffness
matrices follow `isoD` / `columnD` (src/core/03-materials.js:222-266).

A spec is a plain dict (JSON-serialisable):

    {"name": str, "xs": [...], "ys": [...], "zs": [...],
     "materials": [{"name", "E" (MPa), "nu", "cte" (ppm/K), "part",
                    "stage" (default 0), "column" (default false)}, ...],
     "cells": [{"k0", "k1", "mat", "in": [x0, x1, y0, y1] | null,
                "out": [x0, x1, y0, y1] | null, "scale": 1.0,
                "also": mat | null, "also_scale": 1.0,
                "bumps": {"x0", "y0", "pitch", "n", "d"} | null}, ...],
     "Ta", "Tb", "stage", "constraint_span", "feature_x", "feature_y",
     "bga": {"pitch", "inset", "sx", "sy"}, "tol"}

A cell (i, j, k) takes the first rule with k0 <= k < k1 whose region test
passes at the cell centre (inside "in", outside "out", inside one of the
n x n square bump footprints of side d centred in the pitch cells of a bump
array starting at (x0, y0), inside the sphere {"c": [x, y, z], "r"}); no rule
= void.  "constrain_faces": true holds every DOF of the six faces' nodes
instead of the 3-2-1 support.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

NU_MAX_GLOBAL = 0.45           # src/core/00-constants.js:19
COLUMN_INPLANE_FRACTION = 1e-3  # src/core/00-constants.js:28
UNBORN_SCALE = 1e-6            # src/core/00-constants.js:18
T_STRAIN_ORIGIN = 25.0         # src/core/00-constants.js:14
CONSTRAINT_SPAN = 0.4          # src/core/00-constants.js:52

# part codes (src/core/06-mesher.js:3-4)
PART = dict(SUB=0, SOLDER=1, UF_BUMP=2, UF_FILLET=3, DIE=4, TIM=5, LID=6, LIDADH=7,
            STIFF=8, STIFFADH=9)
# parts that split a column into separate shell aggregates (src/core/10-solver.js:248)
COMPLIANT_PARTS = frozenset([1, 2, 3, 5, 7, 9])


def linspace_ref(a: float, b: float, n: int) -> list:
    """n + 1 points a + (b - a) i / n (the verification suite's linspace, src/core/15-verify.js:3)."""
    return [a + (b - a) * i / n for i in range(n + 1)]


def iso_d(E: float, nu: float, nu_max: float = NU_MAX_GLOBAL) -> np.ndarray:
    """Isotropic 6x6 stiffness (Voigt xx, yy, zz, xy, yz, zx; engineering shear).

    Same construction as `isoD` -> `orthoD` (src/core/03-materials.js:222-254):
    clamp nu to nu_max, invert the 3x3 normal compliance, G on the shear diagonal.
    """
    nu = min(nu, nu_max)
    S = np.array([[1 / E, -nu / E, -nu / E], [-nu / E, 1 / E, -nu / E], [-nu / E, -nu / E, 1 / E]])
    C = np.linalg.inv(S)
    D = np.zeros((6, 6))
    D[:3, :3] = C
    G = E / (2 * (1 + nu))
    D[3, 3] = D[4, 4] = D[5, 5] = G
    return D


def column_d(E: float, nu: float) -> np.ndarray:
    """Solder column phase stiffness (src/core/03-materials.js:259-266): axial E,
    transverse shear G in xz / yz, an in-plane numerical floor only."""
    D = np.zeros((6, 6))
    G = E / (2 * (1 + nu))
    fl = COLUMN_INPLANE_FRACTION
    D[0, 0] = D[1, 1] = fl * E
    D[3, 3] = fl * G
    D[2, 2] = E
    D[4, 4] = G
    D[5, 5] = G
    return D


@dataclass
class Mesh:
    xs: np.ndarray
    ys: np.ndarray
    zs: np.ndarray
    node_id: np.ndarray      # int32 [nx*ny*nz], -1 absent
    cell_elems: np.ndarray   # int32 [2*ncells], -1 none
    e_group: np.ndarray      # int32 [ne]
    e_part: np.ndarray       # int8 [ne]
    e_scale: np.ndarray      # float64 [ne]
    e_ijk: np.ndarray        # int32 [ne, 3]
    conn: np.ndarray         # int32 [ne, 8]
    coords: np.ndarray       # float64 [nn, 3]
    col_start: np.ndarray    # int32 [nx*ny+1]
    constraint_dofs: np.ndarray  # int32
    constraint_nodes: list
    feature_x: np.ndarray
    feature_y: np.ndarray
    groups: list = field(default_factory=list)
    e_section: np.ndarray = None  # int32 [ne]: layered section of an element, -1 homogeneous (None: none)
    nsections: int = 0
    # stacked-module stages: the FEModel's moduleConstraints (the package's 3-2-1 plus each
    # module's support, src/core/09-assembly.js:227); None: no module stages
    module_constraint_dofs: np.ndarray = None

    @property
    def nx(self): return len(self.xs)
    @property
    def ny(self): return len(self.ys)
    @property
    def nz(self): return len(self.zs)
    @property
    def nn(self): return len(self.coords)
    @property
    def ne(self): return len(self.e_group)
    @property
    def ndof(self): return 3 * self.nn


# corner order of a hex (src/core/06-mesher.js:46)
HEX_LOC = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)


def _inside(r, x, y):
    return r[0] <= x <= r[1] and r[2] <= y <= r[3]


def in_bumps(b: dict, X, Y):
    """Points (X, Y) inside a bump footprint of the array b: n x n pitch cells
    from (x0, y0), each with a d x d square centred in it."""
    def axis(V, v0):
        t = (np.asarray(V) - v0) / b["pitch"]
        m = np.floor(t)
        off = (t - m) * b["pitch"]
        lo, hi = (b["pitch"] - b["d"]) / 2, (b["pitch"] + b["d"]) / 2
        return (m >= 0) & (m < b["n"]) & (off >= lo) & (off <= hi)
    return axis(X, b["x0"]) & axis(Y, b["y0"])


def bump_index(b: dict, x: float, y: float) -> int:
    """Bump of the array b (row-major in x then y) whose footprint holds (x, y), or -1."""
    mx = math.floor((x - b["x0"]) / b["pitch"])
    my = math.floor((y - b["y0"]) / b["pitch"])
    if not (0 <= mx < b["n"] and 0 <= my < b["n"]):
        return -1
    if not bool(in_bumps(b, np.array([x]), np.array([y]))[0]):
        return -1
    return int(mx * b["n"] + my)


def cell_assignment(spec: dict) -> np.ndarray:
    """Per cell (i, j, k): up to two (material, scale) entries; returns int arrays
    mat[ncells, 2] (-1 none) and scale[ncells, 2]."""
    xs, ys, zs = spec["xs"], spec["ys"], spec["zs"]
    nx, ny, nz = len(xs), len(ys), len(zs)
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1
    mat = -np.ones((ncx, ncy, ncz, 2), dtype=np.int32)
    scl = np.ones((ncx, ncy, ncz, 2), dtype=np.float64)
    xc = [(xs[i] + xs[i + 1]) / 2 for i in range(ncx)]
    yc = [(ys[j] + ys[j + 1]) / 2 for j in range(ncy)]
    rules = spec["cells"]
    # rules are region tests at the cell centre: evaluate with numpy per rule
    X, Y = np.meshgrid(np.array(xc), np.array(yc), indexing="ij")
    for k in range(ncz):
        done = np.zeros((ncx, ncy), dtype=bool)
        for r in rules:
            if not (r["k0"] <= k < r["k1"]):
                continue
            ok = ~done
            if r.get("in") is not None:
                a = r["in"]
                ok &= (X >= a[0]) & (X <= a[1]) & (Y >= a[2]) & (Y <= a[3])
            if r.get("out") is not None:
                a = r["out"]
                ok &= ~((X >= a[0]) & (X <= a[1]) & (Y >= a[2]) & (Y <= a[3]))
            if r.get("bumps") is not None:
                ok &= in_bumps(r["bumps"], X, Y)
            if r.get("sphere") is not None:
                sc = r["sphere"]
                zc = (zs[k] + zs[k + 1]) / 2
                ok &= (X - sc["c"][0]) ** 2 + (Y - sc["c"][1]) ** 2 + (zc - sc["c"][2]) ** 2 <= sc["r"] ** 2
            mat[ok, k, 0] = r["mat"]
            scl[ok, k, 0] = r.get("scale", 1.0)
            if r.get("also") is not None:
                mat[ok, k, 1] = r["also"]
                scl[ok, k, 1] = r.get("also_scale", 1.0)
            done |= ok
    return mat.reshape(-1, 2), scl.reshape(-1, 2)


def build_mesh(spec: dict) -> Mesh:
    """Mesh contract of `buildStructuredMesh` (src/core/06-mesher.js:14-72) for a spec."""
    xs = np.asarray(spec["xs"], dtype=np.float64)
    ys = np.asarray(spec["ys"], dtype=np.float64)
    zs = np.asarray(spec["zs"], dtype=np.float64)
    nx, ny, nz = len(xs), len(ys), len(zs)
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1
    mats = spec["materials"]
    cmat, cscl = cell_assignment(spec)
    # element numbering: cell order (i, j, k), phases in order
    has0 = cmat[:, 0] >= 0
    has1 = cmat[:, 1] >= 0
    cnt = has0.astype(np.int64) + has1.astype(np.int64)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    ne = int(cnt.sum())
    cell_elems = -np.ones(2 * len(cmat), dtype=np.int32)
    cell_elems[0::2] = np.where(has0, first, -1)
    cell_elems[1::2] = np.where(has1, first + has0, -1)
    cells = np.arange(len(cmat))
    e_cell = np.empty(ne, dtype=np.int64)
    e_slot = np.empty(ne, dtype=np.int64)
    e_cell[first[has0]] = cells[has0]
    e_slot[first[has0]] = 0
    e_cell[(first + has0)[has1]] = cells[has1]
    e_slot[(first + has0)[has1]] = 1
    e_group = cmat[e_cell, e_slot].astype(np.int32)
    e_scale = cscl[e_cell, e_slot].astype(np.float64)
    e_part = np.array([mats[g]["part"] for g in range(len(mats))], dtype=np.int8)[e_group]
    ek = e_cell % ncz
    ej = (e_cell // ncz) % ncy
    ei = e_cell // (ncz * ncy)
    e_ijk = np.stack([ei, ej, ek], axis=1).astype(np.int32)
    # node existence
    exist = np.zeros((nx, ny, nz), dtype=bool)
    occ = (cnt > 0).reshape(ncx, ncy, ncz)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                exist[di:di + ncx, dj:dj + ncy, dk:dk + ncz] |= occ
    flat = exist.reshape(-1)
    node_id = -np.ones(flat.size, dtype=np.int32)
    node_id[flat] = np.arange(int(flat.sum()), dtype=np.int32)
    I, J, K = np.nonzero(exist)
    coords = np.stack([xs[I], ys[J], zs[K]], axis=1)
    conn = np.empty((ne, 8), dtype=np.int32)
    for a in range(8):
        li, lj, lk = HEX_LOC[a]
        conn[:, a] = node_id[((ei + li) * ny + ej + lj) * nz + ek + lk]
    col_counts = exist.reshape(nx * ny, nz).sum(axis=1)
    col_start = np.concatenate([[0], np.cumsum(col_counts)]).astype(np.int32)
    m = Mesh(xs=xs, ys=ys, zs=zs, node_id=node_id, cell_elems=cell_elems, e_group=e_group,
             e_part=e_part, e_scale=e_scale, e_ijk=e_ijk, conn=conn, coords=coords,
             col_start=col_start, constraint_dofs=np.zeros(0, np.int32), constraint_nodes=[],
             feature_x=np.asarray(spec.get("feature_x", []), np.int32),
             feature_y=np.asarray(spec.get("feature_y", []), np.int32), groups=list(mats))
    if spec.get("constrain_faces"):
        # every DOF of the nodes on the six faces (a submodel's prescribed boundary)
        onf = np.zeros((nx, ny, nz), bool)
        onf[0], onf[-1], onf[:, 0], onf[:, -1], onf[:, :, 0], onf[:, :, -1] = True, True, True, True, True, True
        nid = node_id.reshape(nx, ny, nz)
        nodes = np.sort(nid[onf & (nid >= 0)])
        m.constraint_nodes = list(nodes)
        m.constraint_dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)
        return m
    span = spec.get("constraint_span")
    if span is None:
        span = CONSTRAINT_SPAN * min((xs[-1] - xs[0]) / 2, (ys[-1] - ys[0]) / 2)
    nodes, dofs = constraints_321(m, span)
    m.constraint_nodes = nodes
    m.constraint_dofs = dofs
    return m


def nearest_node_on_plane(m: Mesh, x: float, y: float, k: int) -> int:
    """First node (in i, j order) of plane k closest to (x, y) (src/core/06-mesher.js:75-84)."""
    best, bd = -1, math.inf
    for i in range(m.nx):
        for j in range(m.ny):
            nid = int(m.node_id[(i * m.ny + j) * m.nz + k])
            if nid < 0:
                continue
            d = math.hypot(m.xs[i] - x, m.ys[j] - y)
            if d < bd - 1e-12:
                bd, best = d, nid
    return best


def constraints_321(m: Mesh, a: float):
    """3-2-1 rigid-body support on the bottom plane (src/core/06-mesher.js:112-116)."""
    n0 = nearest_node_on_plane(m, 0.0, 0.0, 0)
    n1 = nearest_node_on_plane(m, a, 0.0, 0)
    n2 = nearest_node_on_plane(m, 0.0, a, 0)
    dofs = np.array([3 * n0, 3 * n0 + 1, 3 * n0 + 2, 3 * n1 + 1, 3 * n1 + 2, 3 * n2 + 2], dtype=np.int32)
    return [n0, n1, n2], dofs


def group_step_data(spec: dict, Ta: float, Tb: float, stage: int):
    """Per-group step data of a T-independent material set (the `prepareStep`
    output of src/core/09-assembly.js:346-365 for constant tables): mean
    stiffness D (36), thermal stress increment s0 = D (F(Tb) - F(Ta)) (6) with
    F = alpha (T - 25), and the born flag (stage of the group <= step stage)."""
    mats = spec["materials"]
    ng = len(mats)
    D = np.zeros((ng, 36))
    s0 = np.zeros((ng, 6))
    born = np.zeros(ng, dtype=np.uint8)
    for g, mt in enumerate(mats):
        Dg = column_d(mt["E"], mt["nu"]) if mt.get("column") else iso_d(mt["E"], mt["nu"])
        D[g] = Dg.reshape(-1)
        b = mt.get("stage", 0) <= stage
        born[g] = 1 if b else 0
        if b:
            a = mt["cte"] * 1e-6
            dF = a * (Tb - T_STRAIN_ORIGIN) - a * (Ta - T_STRAIN_ORIGIN)
            s0[g] = Dg[:, 0] * dF + Dg[:, 1] * dF + Dg[:, 2] * dF
    return D, s0, born


def bga_sites(pitch: float, inset: float, sx: float, sy: float):
    """BGA ball sites centred on the substrate (src/core/05-geometry.js:23-29, 172-185)."""
    def axis(L):
        span0 = L - 2 * inset
        n = int(math.floor(span0 / pitch + 1e-9)) + 1
        span = (n - 1) * pitch
        return n, -span / 2, span / 2
    nxb, fx, lx = axis(sx)
    nyb, fy, ly = axis(sy)
    X = np.array([fx + i * pitch for i in range(nxb) for j in range(nyb)])
    Y = np.array([fy + j * pitch for i in range(nxb) for j in range(nyb)])
    return X, Y, (fx, lx, fy, ly)


def save_spec(spec: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(spec, f, indent=1)


def load_spec(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def mesh_from_arrays(xs, ys, zs, node_id, cell_elems, e_group, e_part, e_scale, constraint_dofs, feature_x=(),
                     feature_y=(), ngroups=None, e_section=None, nsections=0, module_constraint_dofs=None) -> Mesh:
    """A Mesh from the arrays of the reference's mesh contract (buildStructuredMesh /
    buildPackageMesh, src/core/06-mesher.js:66-71, 500-545): the derived arrays (element
    cells, connectivity, coordinates, column starts) are rebuilt from node_id and cell_elems."""
    xs, ys, zs = (np.asarray(a, np.float64) for a in (xs, ys, zs))
    nx, ny, nz = len(xs), len(ys), len(zs)
    node_id = np.asarray(node_id, np.int32)
    cell_elems = np.asarray(cell_elems, np.int32)
    ne = len(e_group)
    e_cell = np.empty(ne, np.int64)
    slots = np.nonzero(cell_elems >= 0)[0]
    e_cell[cell_elems[slots]] = slots // 2
    ncz, ncy = nz - 1, ny - 1
    ek, ej, ei = e_cell % ncz, (e_cell // ncz) % ncy, e_cell // (ncz * ncy)
    e_ijk = np.stack([ei, ej, ek], axis=1).astype(np.int32)
    conn = np.empty((ne, 8), np.int32)
    for a in range(8):
        li, lj, lk = HEX_LOC[a]
        conn[:, a] = node_id[((ei + li) * ny + ej + lj) * nz + ek + lk]
    exist = node_id.reshape(nx, ny, nz) >= 0
    I, J, K = np.nonzero(exist)
    coords = np.stack([xs[I], ys[J], zs[K]], axis=1)
    col_counts = exist.reshape(nx * ny, nz).sum(axis=1)
    col_start = np.concatenate([[0], np.cumsum(col_counts)]).astype(np.int32)
    ng = int(ngroups if ngroups is not None else np.max(e_group) + 1)
    return Mesh(xs=xs, ys=ys, zs=zs, node_id=node_id, cell_elems=cell_elems,
                e_group=np.asarray(e_group, np.int32), e_part=np.asarray(e_part, np.int8),
                e_scale=np.asarray(e_scale, np.float64), e_ijk=e_ijk, conn=conn, coords=coords,
                col_start=col_start, constraint_dofs=np.asarray(constraint_dofs, np.int32), constraint_nodes=[],
                feature_x=np.asarray(feature_x, np.int32), feature_y=np.asarray(feature_y, np.int32),
                groups=[None] * ng,
                e_section=None if e_section is None else np.asarray(e_section, np.int32), nsections=int(nsections),
                module_constraint_dofs=None if module_constraint_dofs is None
                else np.asarray(module_constraint_dofs, np.int32))
