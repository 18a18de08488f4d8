"""Post-processing of step displacements on the host: JEITA warpage of the
substrate bottom surface (warpageJEITA, src/core/12-post.js:60-99) with the
least-squares plane of fitPlane (src/core/01-util.js:176-190)."""
from __future__ import annotations

import numpy as np


def _interval(a, x):
    """Index i with a[i] <= x < a[i+1], clamped to [0, n-2] (src/core/01-util.js:12-22)."""
    i = np.searchsorted(a, x, side="right") - 1
    return np.clip(i, 0, len(a) - 2)


def sample_plane_w(mesh, u, x, y, k=0):
    """Bilinear w on plane k at points (x, y); NaN where a corner node is absent
    (samplePlane, src/core/12-post.js:30-39)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    i = _interval(mesh.xs, x)
    j = _interval(mesh.ys, y)
    tx = (x - mesh.xs[i]) / (mesh.xs[i + 1] - mesh.xs[i])
    ty = (y - mesh.ys[j]) / (mesh.ys[j + 1] - mesh.ys[j])
    nid = lambda a, b: mesh.node_id[((i + a) * mesh.ny + j + b) * mesh.nz + k]
    n00, n10, n01, n11 = nid(0, 0), nid(1, 0), nid(0, 1), nid(1, 1)
    ok = (n00 >= 0) & (n10 >= 0) & (n01 >= 0) & (n11 >= 0)
    w = lambda n: np.where(ok, u[3 * np.maximum(n, 0) + 2], np.nan)
    return (1 - tx) * (1 - ty) * w(n00) + tx * (1 - ty) * w(n10) + (1 - tx) * ty * w(n01) + tx * ty * w(n11)


def fit_plane(x, y, w):
    mx, my, mw = x.mean(), y.mean(), w.mean()
    dx, dy, dw = x - mx, y - my, w - mw
    sxx, sxy, syy, sxw, syw = (dx * dx).sum(), (dx * dy).sum(), (dy * dy).sum(), (dx * dw).sum(), (dy * dw).sum()
    det = sxx * syy - sxy * sxy
    c1 = c2 = 0.0
    if abs(det) > 1e-300:
        c1 = (sxw * syy - syw * sxy) / det
        c2 = (syw * sxx - sxw * sxy) / det
    return mw - c1 * mx - c2 * my, c1, c2


def warpage_jeita(mesh, u, bga):
    """Signed JEITA warpage: residual range over the BGA sites after removing
    the least-squares plane, sign from the two diagonals of the zone."""
    bx, by, (x0, x1, y0, y1) = bga
    w = sample_plane_w(mesh, u, bx, by)
    c0, c1, c2 = fit_plane(bx, by, w)
    res = w - (c0 + c1 * bx + c2 * by)
    mag = float(res.max() - res.min())
    t = np.arange(201) / 200.0
    total = 0.0
    for xa, ya, xb, yb in ((x0, y0, x1, y1), (x0, y1, x1, y0)):
        wa, wb = sample_plane_w(mesh, u, [xa], [ya])[0], sample_plane_w(mesh, u, [xb], [yb])[0]
        wv = sample_plane_w(mesh, u, xa + t * (xb - xa), ya + t * (yb - ya))
        rel = wv - (wa + t * (wb - wa))
        total += max(rel.max(), 0.0) + min(rel.min(), 0.0)
    sign = 1 if total > 0 else (-1 if total < 0 else 0)
    return {"signed": sign * mag, "mag": mag, "sign": sign, "plane": (c0, c1, c2),
            "rmin": float(res.min()), "rmax": float(res.max())}


def warpage_jeita_grad(mesh, u, bga):
    """Signed JEITA warpage and its gradient with respect to u (3 nn).  Where
    the sites of the largest and smallest residual and the sign stay the
    same the warpage is linear in u: J = sign (res[imax] - res[imin]) with
    res = (I - H) w, H the least-squares plane fit over the sites (symmetric)
    and w the bilinear samples of u_z on plane 0, so dJ/dw = sign (I - H)
    (e_imax - e_imin), taken back to the nodes through the sampling weights."""
    bx, by, _ = bga
    bx, by = np.asarray(bx, np.float64), np.asarray(by, np.float64)
    r = warpage_jeita(mesh, u, bga)
    w = sample_plane_w(mesh, u, bx, by)
    c0, c1, c2 = fit_plane(bx, by, w)
    res = w - (c0 + c1 * bx + c2 * by)
    imax, imin = int(np.argmax(res)), int(np.argmin(res))
    e = np.zeros(len(bx))
    e[imax] += 1.0
    e[imin] -= 1.0
    X = np.stack([np.ones_like(bx), bx, by], 1)
    dw = r["sign"] * (e - X @ np.linalg.solve(X.T @ X, X.T @ e))
    g = np.zeros(3 * mesh.nn)
    i, j = _interval(mesh.xs, bx), _interval(mesh.ys, by)
    tx = (bx - mesh.xs[i]) / (mesh.xs[i + 1] - mesh.xs[i])
    ty = (by - mesh.ys[j]) / (mesh.ys[j + 1] - mesh.ys[j])
    for a, b, wt in ((0, 0, (1 - tx) * (1 - ty)), (1, 0, tx * (1 - ty)), (0, 1, (1 - tx) * ty), (1, 1, tx * ty)):
        n = mesh.node_id[((i + a) * mesh.ny + j + b) * mesh.nz]
        np.add.at(g, 3 * n + 2, wt * dw)
    return r["signed"], g


# ---------------------------------------------------------------------------
# Stress post-processing on the device stress histories
# ---------------------------------------------------------------------------

HEX_XI = np.array([-1, 1, 1, -1, -1, 1, 1, -1], np.float64)
HEX_ETA = np.array([-1, -1, 1, 1, -1, -1, 1, 1], np.float64)
HEX_ZETA = np.array([-1, -1, -1, -1, 1, 1, 1, 1], np.float64)
GP2 = 1.0 / np.sqrt(3.0)
DIE_EDGE_BAND = 0.5     # mm excluded from the die interior peak (src/core/00-constants.js:86)
CORNER_PATCH = 0.1      # mm, averaging patch (src/core/00-constants.js:87)
PATCH_SAMPLES = 5       # samples per patch side (src/core/00-constants.js:88)


def _gp_to_node():
    """Extrapolation from the 8 Gauss points to the 8 nodes: row n holds the
    trilinear functions of the Gauss points at sqrt(3) x the node's natural
    coordinates (GP_TO_NODE, src/core/07-element.js:222-234)."""
    s = np.sqrt(3.0)
    M = np.zeros((8, 8))
    for n in range(8):
        x, y, z = HEX_XI[n] * s, HEX_ETA[n] * s, HEX_ZETA[n] * s
        M[n] = 0.125 * (1 + HEX_XI * x) * (1 + HEX_ETA * y) * (1 + HEX_ZETA * z)
    return M


GP_TO_NODE = _gp_to_node()


def max_principal(s):
    """Largest eigenvalue of symmetric 3x3 tensors in Voigt order (..., 6), by
    the closed form of Smith (1961) (eigSym3, src/core/01-util.js:148-173)."""
    s = np.asarray(s, np.float64)
    a, b, c, d, e, f = (s[..., q] for q in range(6))
    p1 = d * d + e * e + f * f
    diag = p1 <= 1e-30 * (a * a + b * b + c * c + 1e-300)
    q = (a + b + c) / 3
    p2 = (a - q) ** 2 + (b - q) ** 2 + (c - q) ** 2 + 2 * p1
    p = np.sqrt(p2 / 6)
    with np.errstate(invalid="ignore", divide="ignore"):
        B0, B1, B2, B3, B4, B5 = (a - q) / p, (b - q) / p, (c - q) / p, d / p, e / p, f / p
        detB = B0 * (B1 * B2 - B4 * B4) - B3 * (B3 * B2 - B4 * B5) + B5 * (B3 * B4 - B1 * B5)
        r = np.clip(detB / 2, -1, 1)
        phi = np.arccos(r) / 3
        e1 = q + 2 * p * np.cos(phi)
    return np.where(diag, np.maximum(np.maximum(a, b), c), e1)


def plane_index(mesh, z):
    """Plane nearest to z, first on ties (planeIndex, src/core/12-post.js:50-54)."""
    return int(np.argmin(np.abs(np.asarray(mesh.zs) - z)))


def element_in_cell(mesh, i, j, k, part=None):
    """Element of a part in cell (i, j, k), or -1 (elementInCell, src/core/12-post.js:19-27)."""
    c = (i * (mesh.ny - 1) + j) * (mesh.nz - 1) + k
    for q in range(2):
        e = int(mesh.cell_elems[2 * c + q])
        if e >= 0 and (part is None or int(mesh.e_part[e]) == part):
            return e
    return -1


class StressView:
    """The solution at one temperature for stress post-processing (StateView,
    src/core/11-analysis.js:336-365): each element's stress-history state is
    accumulated on the device from the steps of the path that lead to T,
    given as (Step, du) pairs in path order, and cached."""

    def __init__(self, model, steps, T):
        self.model, self.steps, self.T = model, list(steps), T
        self.mesh = model.mesh
        self._rows = {}
        self._states = np.zeros((0, 81))
        self._elems = np.zeros(0, np.int32)

    def states(self, elems):
        elems = np.asarray(elems, np.int32)
        new = [int(e) for e in dict.fromkeys(elems.tolist()) if int(e) not in self._rows]
        if new:
            st = None
            for step, du in self.steps:
                st = self.model.state_accumulate(step, du, new, st)
            base = len(self._elems)
            for q, e in enumerate(new):
                self._rows[e] = base + q
            self._elems = np.concatenate([self._elems, np.asarray(new, np.int32)])
            self._states = np.vstack([self._states, st])
        return np.array([self._rows[int(e)] for e in elems], np.int32)

    def points(self, elems, xez):
        """Stress and strain (n x 6 each) at natural points xez (n x 3) of
        elements elems (n) (StateView.point, src/core/11-analysis.js:356-364)."""
        rows = self.states(elems)
        return self.model.state_stress(self._elems, self._states, rows, xez)


def _rect_contains(r, x, y):
    return (x >= r["x0"]) & (x <= r["x1"]) & (y >= r["y0"]) & (y <= r["y1"])


def die_surface_stress(view, geom, die_idx, which):
    """Nodal-averaged stress on the top or bottom surface of a silicon part from
    the Gauss-point stresses of its adjacent elements, extrapolated to nodes and
    averaged among them; interior peak outside the edge band and 100 um corner
    and mid-edge patch averages (dieSurfaceStress, src/core/12-post.js:180-257).
    geom["siParts"][die_idx]: {name, rect: {x0, x1, y0, y1}, ztop, zb, gkey, part}."""
    mesh = view.mesh
    d = geom["siParts"][die_idx]
    r = d["rect"]
    top = which == "top"
    k_plane = plane_index(mesh, d["ztop"] if top else d["zb"])
    k_elem = k_plane - 1 if top else k_plane
    zeta_face = 1.0 if top else -1.0
    xs, ys = np.asarray(mesh.xs), np.asarray(mesh.ys)
    i0 = int(_interval(xs, r["x0"] + 1e-9)); i1 = int(_interval(xs, r["x1"] - 1e-9)) + 1
    j0 = int(_interval(ys, r["y0"] + 1e-9)); j1 = int(_interval(ys, r["y1"] - 1e-9)) + 1
    nxd, nyd = i1 - i0 + 1, j1 - j0 + 1
    group = [g["name"] for g in mesh.groups].index(d["gkey"])
    band = DIE_EDGE_BAND
    inner = {"x0": r["x0"] + band, "x1": r["x1"] - band, "y0": r["y0"] + band, "y1": r["y1"] - band}
    cells = []
    for i in range(i0, i1):
        for j in range(j0, j1):
            e = element_in_cell(mesh, i, j, k_elem, d["part"])
            if e >= 0 and int(mesh.e_group[e]) == group:
                cells.append((i, j, e))
    ssum = np.zeros((nxd * nyd, 6))
    cnt = np.zeros(nxd * nyd, np.int64)
    gp_peak = gp_peak_all = -np.inf
    if cells:
        ce = np.array([c[2] for c in cells], np.int32)
        gpx = np.stack([HEX_XI * GP2, HEX_ETA * GP2, HEX_ZETA * GP2], 1)
        sig, _ = view.points(np.repeat(ce, 8), np.tile(gpx, (len(ce), 1)))
        sig = sig.reshape(len(ce), 8, 6)
        sp = max_principal(sig)
        ci = np.array([c[0] for c in cells]); cj = np.array([c[1] for c in cells])
        face = HEX_ZETA == zeta_face
        gx = 0.5 * (xs[ci] + xs[ci + 1])[:, None] + (HEX_XI * GP2)[None, :] * 0.5 * (xs[ci + 1] - xs[ci])[:, None]
        gy = 0.5 * (ys[cj] + ys[cj + 1])[:, None] + (HEX_ETA * GP2)[None, :] * 0.5 * (ys[cj + 1] - ys[cj])[:, None]
        fsp = sp[:, face]
        gp_peak_all = float(fsp.max())
        ins = _rect_contains(inner, gx[:, face], gy[:, face])
        if ins.any():
            gp_peak = float(fsp[ins].max())
        for a in np.nonzero(face)[0]:
            v = np.einsum("g,ngc->nc", GP_TO_NODE[a], sig)
            gi = ci - i0 + (1 if HEX_XI[a] > 0 else 0)
            gj = cj - j0 + (1 if HEX_ETA[a] > 0 else 0)
            o = gi * nyd + gj
            np.add.at(ssum, o, v)
            np.add.at(cnt, o, 1)
    smax = np.full(nxd * nyd, np.nan); sxx = smax.copy(); syy = smax.copy()
    has = cnt > 0
    tens = ssum[has] / cnt[has][:, None]
    smax[has] = max_principal(tens); sxx[has] = tens[:, 0]; syy[has] = tens[:, 1]
    gxs, gys = xs[i0:i1 + 1].copy(), ys[j0:j1 + 1].copy()
    peak, px, py = -np.inf, np.nan, np.nan
    for a in range(nxd):
        for b in range(nyd):
            v = smax[a * nyd + b]
            if not np.isfinite(v) or not _rect_contains(inner, gxs[a], gys[b]):
                continue
            if v > peak:
                peak, px, py = float(v), float(gxs[a]), float(gys[b])

    def sample_grid(x, y):
        a, b = int(_interval(gxs, x)), int(_interval(gys, y))
        tx = (x - gxs[a]) / (gxs[a + 1] - gxs[a]); ty = (y - gys[b]) / (gys[b + 1] - gys[b])
        v00, v10 = smax[a * nyd + b], smax[(a + 1) * nyd + b]
        v01, v11 = smax[a * nyd + b + 1], smax[(a + 1) * nyd + b + 1]
        return (1 - tx) * (1 - ty) * v00 + tx * (1 - ty) * v10 + (1 - tx) * ty * v01 + tx * ty * v11

    def patch(xa, ya):
        P, m = CORNER_PATCH, PATCH_SAMPLES
        tot = 0.0
        for a in range(m):
            for b in range(m):
                tot += sample_grid(xa + P * (a + 0.5) / m, ya + P * (b + 0.5) / m)
        return tot / (m * m)

    P = CORNER_PATCH
    # corners 1..4 clockwise from (+x, -y) (dieCorners, src/core/12-post.js:316-323)
    corners = [(r["x1"], r["y0"], 1, -1), (r["x0"], r["y0"], -1, -1), (r["x0"], r["y1"], -1, 1), (r["x1"], r["y1"], 1, 1)]
    corner_vals = [patch(x - P if sx > 0 else x, y - P if sy > 0 else y) for x, y, sx, sy in corners]
    cx, cy = 0.5 * (r["x0"] + r["x1"]), 0.5 * (r["y0"] + r["y1"])
    mid_vals = [patch(cx - P / 2, r["y0"]), patch(r["x0"], cy - P / 2), patch(cx - P / 2, r["y1"] - P),
                patch(r["x1"] - P, cy - P / 2)]
    return {"die": d.get("name"), "which": which, "xs": gxs, "ys": gys, "nx": nxd, "ny": nyd, "smax": smax,
            "sxx": sxx, "syy": syy, "peakInterior": peak, "peakX": px, "peakY": py, "gpPeakInterior": gp_peak,
            "gpPeakAll": gp_peak_all, "corners": corner_vals, "midEdges": mid_vals, "T": view.T}


def die_bump_sites(d):
    """Bump sites of a field {nbx, nby, bx0, by0, px, py, peripheral, rows}
    (dieBumpSites, src/core/05-geometry.js:188-198)."""
    xs, ys = [], []
    for i in range(d["nbx"]):
        for j in range(d["nby"]):
            rr = min(i, j, d["nbx"] - 1 - i, d["nby"] - 1 - j)
            if d.get("peripheral") and rr >= d["rows"]:
                continue
            xs.append(d["bx0"] + i * d["px"]); ys.append(d["by0"] + j * d["py"])
    return np.array(xs), np.array(ys)


def joint_loads(view, d, zmid, part, A):
    """Axial and shear forces N = szz A, V = |(tyz, tzx)| A at the joint sites
    of a bump field at height zmid in elements of a part (jointLoads,
    src/core/12-post.js:279-309)."""
    mesh = view.mesh
    sx, sy = die_bump_sites(d)
    n = len(sx)
    N = np.full(n, np.nan); V = np.full(n, np.nan); szz = np.full(n, np.nan)
    kb = int(_interval(np.asarray(mesh.zs), zmid))
    xs, ys, zs = np.asarray(mesh.xs), np.asarray(mesh.ys), np.asarray(mesh.zs)
    qs, es, pts = [], [], []
    for q in range(n):
        i, j = int(_interval(xs, sx[q])), int(_interval(ys, sy[q]))
        e = element_in_cell(mesh, i, j, kb, part)
        if e < 0:
            continue
        # natural coordinates of the site in its element (naturalCoords, src/core/12-post.js:4-11)
        ei, ej, ek = (int(v) for v in mesh.e_ijk[e])
        pts.append([np.clip(2 * (sx[q] - xs[ei]) / (xs[ei + 1] - xs[ei]) - 1, -1, 1),
                    np.clip(2 * (sy[q] - ys[ej]) / (ys[ej + 1] - ys[ej]) - 1, -1, 1),
                    np.clip(2 * (zmid - zs[ek]) / (zs[ek + 1] - zs[ek]) - 1, -1, 1)])
        qs.append(q); es.append(e)
    if qs:
        sig, _ = view.points(np.asarray(es, np.int32), np.asarray(pts))
        qa = np.asarray(qs)
        szz[qa] = sig[:, 2]
        N[qa] = sig[:, 2] * A
        V[qa] = np.hypot(sig[:, 4], sig[:, 5]) * A
    iN = iV = -1
    wN = wV = -np.inf
    for q in range(n):
        if np.isfinite(N[q]) and abs(N[q]) > wN:
            wN, iN = abs(N[q]), q
        if np.isfinite(V[q]) and V[q] > wV:
            wV, iV = V[q], q
    order = sorted((q for q in range(n) if np.isfinite(V[q])), key=lambda q: -np.hypot(N[q], V[q]))[:10]
    return {"x": sx, "y": sy, "N": N, "V": V, "szz": szz,
            "worstN": N[iN] if iN >= 0 else np.nan, "worstV": V[iV] if iV >= 0 else np.nan,
            "worstNidx": iN, "worstVidx": iV,
            "worst": [{"q": q, "x": sx[q], "y": sy[q], "N": N[q], "V": V[q]} for q in order], "T": view.T, "abump": A}
