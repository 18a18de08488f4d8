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
