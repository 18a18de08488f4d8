"""Workload specs: the bare-die FCBGA (48 x 48 x 14 hexes), the lidded
full-fidelity FCBGA (160 x 160 x 52 hexes, graded) and small test boxes.

Every value the workload statement leaves open is pinned here, once:
substrate split (build-up 0.2 mm per side, core 0.8 mm), Poisson ratios of
the build-up, bump and TIM layers (0.3), the die and bump layers present only
under the die, the BGA field (pitch 1.0 mm, inset 1.0 mm), the grading law of
the lidded mesh, the lid geometry and the six temperatures.
"""
from __future__ import annotations

import math

import numpy as np

from .meshspec import PART, linspace_ref

# ---- materials (E in MPa, CTE in ppm/K) -----------------------------------------
CORE = dict(name="core", E=26000.0, nu=0.2, cte=15.0, part=PART["SUB"])
BUILDUP = dict(name="buildup", E=5000.0, nu=0.3, cte=40.0, part=PART["SUB"])
BUMP_UF = dict(name="bump_uf", E=8000.0, nu=0.3, cte=30.0, part=PART["UF_BUMP"])
SI = dict(name="si", E=131000.0, nu=0.28, cte=2.6, part=PART["DIE"])
TIM = dict(name="tim", E=10.0, nu=0.3, cte=30.0, part=PART["TIM"])
LID = dict(name="lid_cu", E=117000.0, nu=0.34, cte=17.0, part=PART["LID"])
LIDADH = dict(name="lid_adh", E=2000.0, nu=0.3, cte=60.0, part=PART["LIDADH"])


def _stack(layers):
    """zs from (thickness, n_elements) pairs; cumulative sums in float64."""
    zs = [0.0]
    for t, n in layers:
        z0 = zs[-1]
        for q in range(1, n + 1):
            zs.append(z0 + t * q / n)
    return zs


def cfg1_spec() -> dict:
    """Bare-die FCBGA, uniform 48 x 48 in-plane x 14 layers.

    Substrate 40 x 40 x 1.2 mm (build-up 0.2 / core 0.8 / build-up 0.2 mm, 2/4/2
    layers), bump/underfill 0.08 mm (2 layers) and die 20 x 20 x 0.775 mm
    (4 layers) only under the die; 150 -> 25 C in one step; PCG to 1e-10.
    """
    xs = linspace_ref(-20.0, 20.0, 48)
    zs = _stack([(0.2, 2), (0.8, 4), (0.2, 2), (0.08, 2), (0.775, 4)])
    die = [-10.0, 10.0, -10.0, 10.0]
    return {
        "name": "cfg1_bare_die_48x48x14",
        "xs": xs, "ys": list(xs), "zs": zs,
        "materials": [BUILDUP, CORE, BUMP_UF, SI],
        "cells": [
            {"k0": 0, "k1": 2, "mat": 0},
            {"k0": 2, "k1": 6, "mat": 1},
            {"k0": 6, "k1": 8, "mat": 0},
            {"k0": 8, "k1": 10, "mat": 2, "in": die},
            {"k0": 10, "k1": 14, "mat": 3, "in": die},
        ],
        "Ta": 150.0, "Tb": 25.0, "stage": 0,
        "constraint_span": 8.0,
        "feature_x": [12, 36], "feature_y": [12, 36],
        "bga": {"pitch": 1.0, "inset": 1.0, "sx": 40.0, "sy": 40.0},
        "top_plane": 14,
        "tol": 1e-10,
        # stress recovery checks: die top/bottom surface maps; joint loads of a
        # 39 x 39 field at 0.5 mm pitch (0.1 mm joints) in the smeared bump layer
        "stress_die": {"rect": die, "ztop": zs[14], "zb": zs[10], "gkey": "si"},
        "stress_joints": {"nbx": 39, "nby": 39, "bx0": -9.5, "by0": -9.5, "px": 0.5, "py": 0.5,
                          "zmid": 0.5 * (zs[8] + zs[10]), "part": PART["UF_BUMP"], "A": math.pi * 0.05 ** 2},
    }


# SAC305 solder of the bump-array workload: the Anand constants of the workload
# statement; elastic E(T) and nu and the CTE of the reference's sac305 record
# (materials-db.json: 57.4 / 46.5 / 38.2 / 30.0 GPa at -40 / 25 / 75 / 125 C,
# linear, nu 0.40, 22.5 ppm/K)
SAC305_ANAND = dict(A=5.87e6, QR=7460.0, xi=2.0, m=0.0942, shat=58.3, n=0.015, h0=9350.0, a=1.5, s0=45.9)
SAC305 = dict(name="sac305", E=46500.0, nu=0.40, cte=22.5, part=PART["SOLDER"],
              E_table={"T": [-40.0, 25.0, 75.0, 125.0], "v": [57400.0, 46500.0, 38200.0, 30000.0]},
              anand=SAC305_ANAND)


def graded_sizes(length: float, h_small: float, h_max: float, ratio: float = 1.25) -> list:
    """Cell sizes covering `length`, smallest first: h_small growing by `ratio`
    until h_max, then cells of about h_max; all scaled to fill the length."""
    sizes, h, tot = [], h_small, 0.0
    while h < h_max and tot + h <= length:
        sizes.append(h)
        tot += h
        h *= ratio
    n_uni = max(0, int(round((length - tot) / h_max)))
    sizes += [h_max] * n_uni
    f = length / sum(sizes)
    return [v * f for v in sizes]


def bump_array_lines(a0: float, pitch: float, n: int, d: float, per_bump: int) -> list:
    """Grid lines over a bump array from a0: in every pitch cell the bump's
    per_bump equal cells (width d, centred), the gaps one cell each."""
    g = (pitch - d) / 2
    lines = [a0]
    for m in range(n):
        b0 = a0 + m * pitch + g
        lines += [b0 + d * q / per_bump for q in range(per_bump + 1)]
    lines.append(a0 + n * pitch)
    return lines


def cfg3_lines(arr0: float, arr1: float, arr_lines: list, half: float = 20.0, die_half: float = 10.0,
               n_out: int = 12, h: float = 40.0 / 48) -> list:
    """Grid lines on [-half, half]: n_out uniform cells from -half to the die
    edge -die_half, graded (ratio <= 1.25, cells up to h) from the bump array's
    gap size toward the die edge on the left and toward the substrate edge on
    the right of the array [arr0, arr1]."""
    left = linspace_ref(-half, -die_half, n_out)
    gap = arr_lines[1] - arr_lines[0]
    s_left = graded_sizes(arr0 - (-die_half), 1.5 * gap, h)
    mid = [arr0]
    for v in s_left:
        mid.append(mid[-1] - v)
    mid = list(reversed(mid))
    mid[0] = -die_half
    s_right = graded_sizes(half - arr1, 1.5 * gap, h)
    right = [arr1]
    for v in s_right:
        right.append(right[-1] + v)
    right[-1] = half
    return left[:-1] + mid[:-1] + arr_lines[:-1] + right


def cfg3_spec() -> dict:
    """Bare-die FCBGA of cfg1_spec with a 20 x 20 array of discrete SAC305
    bumps at 150 um pitch, each 3 x 3 x 4 hexes, under the die corner.

    The array covers [7, 10] x [7, 10] mm (its outer bumps at the die corner
    x, y = 10, where the distance to the neutral point is largest); a bump is a
    100 um square column (3 cells of 33.3 um), the gaps 50 um (one cell), 25 um
    at the array's edges.  Outside the array the bump layer is the smeared
    bump/underfill layer of cfg1 (between the bumps, underfill of the same
    properties).  The bump layer is 0.08 mm in 4 layers; the in-plane lines are
    cfg1's 0.833 mm spacing graded (ratio <= 1.25) down to the array.
    Loading: stress-free at 25 C, 2 temperature cycles -40 <-> 125 C with 15 min
    ramps and 15 min dwells, 15 one-minute steps per ramp or dwell (the first
    ramp 25 -> 125 C), 120 steps.
    """
    bumps = {"x0": 7.0, "y0": 7.0, "pitch": 0.15, "n": 20, "d": 0.1}
    arr = bump_array_lines(7.0, 0.15, 20, 0.1, 3)
    xs = cfg3_lines(7.0, 10.0, arr)
    zs = _stack([(0.2, 2), (0.8, 4), (0.2, 2), (0.08, 4), (0.775, 4)])
    die = [-10.0, 10.0, -10.0, 10.0]
    return {
        "name": "cfg3_bare_die_bump_array_20x20",
        "xs": xs, "ys": list(xs), "zs": zs,
        "materials": [BUILDUP, CORE, BUMP_UF, SI, SAC305],
        "cells": [
            {"k0": 0, "k1": 2, "mat": 0},
            {"k0": 2, "k1": 6, "mat": 1},
            {"k0": 6, "k1": 8, "mat": 0},
            {"k0": 8, "k1": 12, "mat": 4, "in": die, "bumps": bumps},
            {"k0": 8, "k1": 12, "mat": 2, "in": die},
            {"k0": 12, "k1": 16, "mat": 3, "in": die},
        ],
        "bumps": bumps, "bump_layer": [8, 12], "solder_group": 4,
        "Tref": 25.0,
        "cycle": {"Tmin": -40.0, "Tmax": 125.0, "cycles": 2, "steps_per_segment": 15, "segment_s": 900.0},
        "Ta": 25.0, "Tb": 125.0, "stage": 0,
        "constraint_span": 8.0,
        "feature_x": [xs.index(-10.0), xs.index(10.0)], "feature_y": [xs.index(-10.0), xs.index(10.0)],
        "bga": {"pitch": 1.0, "inset": 1.0, "sx": 40.0, "sy": 40.0},
        "top_plane": 16,
        "tol": 1e-10,
    }


def cfg3_small_spec(n: int = 2, solder_anand: dict = None, half: float = 1.0) -> dict:
    """A small stand-in of the bump-array workload for the oracle: a substrate
    of 2 half x 2 half mm (build-up 0.1 / core 0.4 / build-up 0.1 mm, 1/2/1
    layers), the 0.08 mm bump layer (4 layers) and a 1 x 1 x 0.3 mm die (2
    layers), with an n x n array of the same bumps (150 um pitch, 3 x 3 x 4
    hexes) at the die corner [0.5 - 0.15 n, 0.5]^2; same materials and loading
    as cfg3_spec."""
    a0 = 0.5 - 0.15 * n
    bumps = {"x0": a0, "y0": a0, "pitch": 0.15, "n": n, "d": 0.1}
    arr = bump_array_lines(a0, 0.15, n, 0.1, 3)
    xs = cfg3_lines(a0, 0.5, arr, half=half, die_half=0.5, n_out=2 if half > 0.75 else 1, h=0.25)
    zs = _stack([(0.1, 1), (0.4, 2), (0.1, 1), (0.08, 4), (0.3, 2)])
    die = [-0.5, 0.5, -0.5, 0.5]
    sol = dict(SAC305)
    if solder_anand is not None:
        sol["anand"] = dict(SAC305_ANAND, **solder_anand)
    return {
        "name": f"cfg3_small_{n}x{n}",
        "xs": xs, "ys": list(xs), "zs": zs,
        "materials": [BUILDUP, CORE, BUMP_UF, SI, sol],
        "cells": [
            {"k0": 0, "k1": 1, "mat": 0},
            {"k0": 1, "k1": 3, "mat": 1},
            {"k0": 3, "k1": 4, "mat": 0},
            {"k0": 4, "k1": 8, "mat": 4, "in": die, "bumps": bumps},
            {"k0": 4, "k1": 8, "mat": 2, "in": die},
            {"k0": 8, "k1": 10, "mat": 3, "in": die},
        ],
        "bumps": bumps, "bump_layer": [4, 8], "solder_group": 4,
        "Tref": 25.0,
        "cycle": {"Tmin": -40.0, "Tmax": 125.0, "cycles": 2, "steps_per_segment": 15, "segment_s": 900.0},
        "Ta": 25.0, "Tb": 125.0, "stage": 0,
        "constraint_span": 0.4,
        "feature_x": [xs.index(-0.5), xs.index(0.5)], "feature_y": [xs.index(-0.5), xs.index(0.5)],
        "bga": {"pitch": 0.5, "inset": 0.25, "sx": 2 * half, "sy": 2 * half},
        "top_plane": 10,
        "tol": 1e-10,
    }


class MT19937_64:
    """The 64-bit Mersenne Twister (std::mt19937_64: Matsumoto and Nishimura's
    MT19937-64 with the standard's default seeding)."""
    NN, MM = 312, 156
    MATRIX_A, UM, LM = 0xB5026F5AA96619E9, 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.mt = [seed & self.MASK]
        for i in range(1, self.NN):
            prev = self.mt[-1]
            self.mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & self.MASK)
        self.mti = self.NN

    def next(self) -> int:
        mt = self.mt
        if self.mti >= self.NN:
            for i in range(self.NN):
                x = (mt[i] & self.UM) | (mt[(i + 1) % self.NN] & self.LM)
                xa = x >> 1
                if x & 1:
                    xa ^= self.MATRIX_A
                mt[i] = mt[(i + self.MM) % self.NN] ^ xa
            self.mti = 0
        x = mt[self.mti]
        self.mti += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self.MASK

    def uniform(self) -> float:
        """A double in [0, 1) from the top 53 bits."""
        return (self.next() >> 11) * 2.0 ** -53


CFG5_SEED = 20261203
UNDERFILL5 = dict(name="underfill", E=8000.0, nu=0.3, cte=30.0, part=PART["UF_BUMP"])


def cfg5_draws(count: int = 256, seed: int = CFG5_SEED) -> list:
    """The seeded parameters of the batched bump submodels: per submodel, from
    one std::mt19937_64 stream in this order, the voxel count n = 24 + (r mod 33)
    (uniform on 24..56), the sphere diameter as a fraction 0.6 + 0.3 u of the
    cube edge, the submodel's position (x, y) = 4 + 6 (u1, u2) mm from the die
    centre (corner and edge region of a 20 mm die), and 3 x 9 amplitudes (within
    2e-8 mm/K) and phases of the boundary field's smooth part."""
    g = MT19937_64(seed)
    out = []
    for q in range(count):
        n = 24 + g.next() % 33
        frac = 0.6 + 0.3 * g.uniform()
        x0, y0 = 4.0 + 6.0 * g.uniform(), 4.0 + 6.0 * g.uniform()
        amp = [[2e-8 * (2 * g.uniform() - 1) for _ in range(9)] for _ in range(3)]
        ph = [[2 * math.pi * g.uniform() for _ in range(9)] for _ in range(3)]
        out.append({"id": q, "n": int(n), "frac": frac, "x0": x0, "y0": y0, "amp": amp, "phase": ph})
    return out


def cfg5_spec(d: dict, edge: float = 0.15) -> dict:
    """One batched bump submodel: a cube of edge 0.15 mm in n^3 voxels, the
    bottom n/6 voxel layers the substrate (core of cfg1), the top n/6 silicon,
    between them underfill (E 8 GPa, CTE 30 ppm/K) around a SAC305 sphere of
    diameter frac x edge centred in the joint layer, cut by the two slabs into
    a sphere cap; the cube's six faces follow the boundary field; one cycle of
    the bump-array workload's history (60 steps)."""
    n = d["n"]
    lines = linspace_ref(0.0, edge, n)
    slab = max(3, int(round(n / 6)))
    zc = (lines[slab] + lines[n - slab]) / 2
    sol = dict(SAC305)
    return {
        "name": f"cfg5_submodel_{d['id']}_n{n}",
        "xs": lines, "ys": list(lines), "zs": list(lines),
        "materials": [CORE, sol, UNDERFILL5, SI],
        "cells": [
            {"k0": 0, "k1": slab, "mat": 0},
            {"k0": n - slab, "k1": n, "mat": 3},
            {"k0": slab, "k1": n - slab, "mat": 1, "sphere": {"c": [edge / 2, edge / 2, zc], "r": d["frac"] * edge / 2}},
            {"k0": slab, "k1": n - slab, "mat": 2},
        ],
        "constrain_faces": True,
        "solder_group": 1, "Tref": 25.0, "bump_layer": [slab, n - slab], "draw": d,
        "cycle": {"Tmin": -40.0, "Tmax": 125.0, "cycles": 1, "steps_per_segment": 15, "segment_s": 900.0},
        "Ta": 25.0, "Tb": 125.0, "stage": 0, "feature_x": [], "feature_y": [],
        "bga": {"pitch": 0.05, "inset": 0.025, "sx": edge, "sy": edge}, "top_plane": n, "tol": 1e-10,
    }


CFG5_SHEAR_SHARE = 0.15  # share of the die/substrate CTE mismatch the joint layer takes (package compliance)


def cfg5_boundary(spec: dict, coords, T: float):
    """Prescribed displacement (mm) at points coords (k, 3) of a submodel at T:
    (T - Tref) g(x).  In-plane: the substrate's expansion (15 ppm/K) at the
    submodel's position (x0, y0) from the die centre, plus, across the joint
    layer (linearly in z from the substrate slab's top to the silicon slab's
    bottom, constant in the slabs), CFG5_SHEAR_SHARE of the silicon/substrate
    mismatch (2.6 - 15 ppm/K): the bump's shear.  Out of plane: the expansion
    integrated through the layers (substrate 15, joint 30, silicon 2.6 ppm/K).
    Plus the seeded smooth part, sum over k in {0,1,2}^2 of
    a_k cos(pi (k1 x + k2 z) / edge + phase_k) per component."""
    d = spec["draw"]
    edge = spec["xs"][-1] - spec["xs"][0]
    k0, k1 = spec["bump_layer"]
    zb, zt = spec["zs"][k0], spec["zs"][k1]
    a_sub, a_si, a_joint = 15e-6, 2.6e-6, 30e-6
    c = np.asarray(coords, np.float64)
    z = c[:, 2]
    zeta = np.clip((z - zb) / (zt - zb), 0.0, 1.0)
    a_in = a_sub + CFG5_SHEAR_SHARE * (a_si - a_sub) * zeta
    uz = a_sub * np.minimum(z, zb) + a_joint * (np.clip(z, zb, zt) - zb) + a_si * np.maximum(z - zt, 0.0)
    g = np.stack([a_in * (d["x0"] + c[:, 0]), a_in * (d["y0"] + c[:, 1]), uz], axis=1)
    for comp in range(3):
        for q in range(9):
            q1, q2 = q // 3, q % 3
            g[:, comp] += d["amp"][comp][q] * np.cos(np.pi * (q1 * c[:, 0] + q2 * c[:, 2]) / edge + d["phase"][comp][q])
    return (T - spec["Tref"]) * g


def small_box_spec() -> dict:
    """A 6 x 5 x 4 box exercising every assembly branch: three materials, a void
    corner, an element scale, two superposed phases in one layer and a group born
    in a later stage (stiffness x 1e-6, no load)."""
    xs = linspace_ref(0.0, 6.0, 6)
    ys = linspace_ref(0.0, 4.0, 5)
    zs = [0.0, 0.3, 0.5, 0.9, 1.2]
    mats = [
        dict(name="A", E=130000.0, nu=0.28, cte=2.6, part=PART["DIE"]),
        dict(name="B", E=20000.0, nu=0.34, cte=16.4, part=PART["SUB"]),
        dict(name="C", E=5000.0, nu=0.3, cte=40.0, part=PART["UF_BUMP"]),
        dict(name="D", E=9000.0, nu=0.35, cte=22.0, part=PART["SOLDER"], column=True),
        dict(name="E", E=50000.0, nu=0.3, cte=10.0, part=PART["LID"], stage=1),
    ]
    return {
        "name": "small_box",
        "xs": xs, "ys": ys, "zs": zs,
        "materials": mats,
        "cells": [
            {"k0": 0, "k1": 1, "mat": 1},
            {"k0": 1, "k1": 2, "mat": 2, "scale": 0.7, "also": 3, "also_scale": 0.3},
            {"k0": 2, "k1": 3, "mat": 0, "out": [4.5, 6.0, 3.0, 4.0]},
            {"k0": 3, "k1": 4, "mat": 4, "in": [0.0, 3.0, 0.0, 4.0]},
        ],
        "Ta": 125.0, "Tb": 25.0, "stage": 0,
        "constraint_span": 0.4 * min(3.0, 2.0),
        "feature_x": [3], "feature_y": [],
        "bga": {"pitch": 1.0, "inset": 0.5, "sx": 6.0, "sy": 4.0},
        "top_plane": 3,
        "tol": 1e-13,
        # stress recovery checks: die surface of part A (with the void corner),
        # solder-phase joint loads in the superposed layer
        "stress_die": {"rect": [0.0, 6.0, 0.0, 4.0], "ztop": zs[3], "zb": zs[2], "gkey": "A"},
        "stress_joints": {"nbx": 6, "nby": 4, "bx0": 0.5, "by0": 0.5, "px": 1.0, "py": 1.0,
                          "zmid": 0.5 * (zs[1] + zs[2]), "part": PART["SOLDER"], "A": 0.05},
    }


def plate_spec(n: int = 16, nz_sub: int = 2) -> dict:
    """Bimaterial plate (the Timoshenko check of src/core/15-verify.js:164-181):
    40 x 40 mm, two 0.5 mm layers, cooled by 100 C."""
    xs = linspace_ref(-20.0, 20.0, n)
    zs = _stack([(0.5, nz_sub), (0.5, nz_sub)])
    return {
        "name": f"plate_{n}",
        "xs": xs, "ys": list(xs), "zs": zs,
        "materials": [dict(name="B", E=110000.0, nu=0.34, cte=16.4, part=PART["SUB"]),
                      dict(name="A", E=130000.0, nu=0.28, cte=2.6, part=PART["DIE"])],
        "cells": [{"k0": 0, "k1": nz_sub, "mat": 0}, {"k0": nz_sub, "k1": 2 * nz_sub, "mat": 1}],
        "Ta": 25.0, "Tb": -75.0, "stage": 0,
        "constraint_span": 8.0, "feature_x": [], "feature_y": [],
        "bga": {"pitch": 1.0, "inset": 1.0, "sx": 40.0, "sy": 40.0},
        "top_plane": 2 * nz_sub, "tol": 1e-10,
    }


# staircase grading of the lidded mesh, in units of 0.005 mm: (size, count) runs
# from the centre to the die edge, die edge to lid-foot edge, and across the foot
_RUN_CENTRE_TO_DIE = [(100, 8), (80, 3), (64, 6), (50, 4), (40, 3), (32, 3), (25, 4), (20, 3)]
_RUN_DIE_TO_FOOT = [(20, 5), (25, 6), (32, 7), (40, 6), (50, 7), (64, 4), (80, 2), (64, 5)]
_RUN_FOOT = [(50, 4)]


def graded_lines() -> list:
    """161 grid lines on [-20, 20] mm graded toward the die edges (+-10 mm).

    Element sizes take eight values 0.1 .. 0.5 mm, neighbouring sizes differ by
    at most 1.28x (the reference mesher's growth limit is 1.3), the smallest sit
    at the die edges and 4 x 0.25 mm span the 1 mm lid foot.  Lines are integer
    multiples of 0.005 mm, so the die edges (+-10) and the foot edges (+-19)
    fall exactly on grid lines.
    """
    units = [0]
    for size, count in _RUN_CENTRE_TO_DIE + _RUN_DIE_TO_FOOT + _RUN_FOOT:
        for _ in range(count):
            units.append(units[-1] + size)
    assert units[-1] == 4000 and len(units) == 81
    right = [u / 200.0 for u in units]
    return [-v for v in reversed(right[1:])] + right


def cfg2_spec() -> dict:
    """Lidded FCBGA at full fidelity: 160 x 160 in-plane (graded toward the die
    edges) x 52 layers.

    Stack (bottom to top): build-up 0.2 mm (4 layers), core 0.8 mm (12),
    build-up 0.2 mm (4); bump/underfill 0.08 mm under the die with the lid
    adhesive (E 2 GPa, CTE 60 ppm/K) at the same height under the lid foot (4);
    die 0.775 mm (14); TIM 0.05 mm over the die (2); Cu lid plate 0.5 mm over
    the whole package (12).  The Cu lid foot is a 1.0 mm ring at the substrate
    edge spanning the die and TIM layers; between foot and die is the lid
    cavity (void).  Six temperatures 25..260 C from 150 C stress-free, each
    solve cold.
    """
    xs = graded_lines()
    zs = _stack([(0.2, 4), (0.8, 12), (0.2, 4), (0.08, 4), (0.775, 14), (0.05, 2), (0.5, 12)])
    assert len(zs) - 1 == 52
    die = [-10.0, 10.0, -10.0, 10.0]
    foot_in = [-19.0, 19.0, -19.0, 19.0]
    return {
        "name": "cfg2_lidded_160x160x52",
        "xs": xs, "ys": list(xs), "zs": zs,
        "materials": [BUILDUP, CORE, BUMP_UF, SI, TIM, LID, LIDADH],
        "cells": [
            {"k0": 0, "k1": 4, "mat": 0},
            {"k0": 4, "k1": 16, "mat": 1},
            {"k0": 16, "k1": 20, "mat": 0},
            {"k0": 20, "k1": 24, "mat": 2, "in": die},
            {"k0": 20, "k1": 24, "mat": 6, "out": foot_in},
            {"k0": 24, "k1": 38, "mat": 3, "in": die},
            {"k0": 24, "k1": 40, "mat": 5, "out": foot_in},
            {"k0": 38, "k1": 40, "mat": 4, "in": die},
            {"k0": 40, "k1": 52, "mat": 5},
        ],
        "Ta": 150.0, "Tb": 25.0, "stage": 0,
        "constraint_span": 8.0,
        "feature_x": [xs.index(v) for v in (-19.0, -10.0, 10.0, 19.0)],
        "feature_y": [xs.index(v) for v in (-19.0, -10.0, 10.0, 19.0)],
        "bga": {"pitch": 1.0, "inset": 1.0, "sx": 40.0, "sy": 40.0},
        "top_plane": 52,
        "temperatures": [25.0, 72.0, 119.0, 166.0, 213.0, 260.0],
        "tol": 1e-10,
        # stress recovery checks: die top/bottom surface maps; joint loads of a
        # 39 x 39 field at 0.5 mm pitch (0.1 mm joints) in the smeared bump layer
        "stress_die": {"rect": die, "ztop": zs[38], "zb": zs[24], "gkey": "si"},
        "stress_joints": {"nbx": 39, "nby": 39, "bx0": -9.5, "by0": -9.5, "px": 0.5, "py": 0.5,
                          "zmid": 0.5 * (zs[20] + zs[24]), "part": PART["UF_BUMP"], "A": math.pi * 0.05 ** 2},
    }


def cfg4_spec() -> dict:
    """Adjoint sensitivity workload: the lidded mesh of cfg2_spec with its
    layers split into 18 material groups (same materials, so the same
    solution).  Objectives: the warpage proxy on the lid top plane (k = 52) and
    the von Mises p-norm (p = 8, over Gauss points, unweighted) of the die
    elements (groups die_lower and die_upper)."""
    sp = cfg2_spec()
    die = [-10.0, 10.0, -10.0, 10.0]
    foot_in = [-19.0, 19.0, -19.0, 19.0]

    def m(base, name):
        return dict(base, name=name)

    mats = [m(BUILDUP, "buildup_bot_outer"), m(BUILDUP, "buildup_bot_inner"), m(CORE, "core_bot"),
            m(CORE, "core_mid"), m(CORE, "core_top"), m(BUILDUP, "buildup_top_inner"),
            m(BUILDUP, "buildup_top_outer"), m(BUMP_UF, "bump_uf_lower"), m(BUMP_UF, "bump_uf_upper"),
            m(LIDADH, "lid_adh"), m(SI, "die_lower"), m(SI, "die_upper"), m(TIM, "tim"),
            m(LID, "lid_foot_lower"), m(LID, "lid_foot_upper"), m(LID, "lid_plate_lower"),
            m(LID, "lid_plate_mid"), m(LID, "lid_plate_upper")]
    sp.update({
        "name": "cfg4_lidded_18_groups",
        "materials": mats,
        "cells": [
            {"k0": 0, "k1": 2, "mat": 0}, {"k0": 2, "k1": 4, "mat": 1},
            {"k0": 4, "k1": 8, "mat": 2}, {"k0": 8, "k1": 12, "mat": 3}, {"k0": 12, "k1": 16, "mat": 4},
            {"k0": 16, "k1": 18, "mat": 5}, {"k0": 18, "k1": 20, "mat": 6},
            {"k0": 20, "k1": 22, "mat": 7, "in": die}, {"k0": 22, "k1": 24, "mat": 8, "in": die},
            {"k0": 20, "k1": 24, "mat": 9, "out": foot_in},
            {"k0": 24, "k1": 31, "mat": 10, "in": die}, {"k0": 31, "k1": 38, "mat": 11, "in": die},
            {"k0": 24, "k1": 32, "mat": 13, "out": foot_in}, {"k0": 32, "k1": 40, "mat": 14, "out": foot_in},
            {"k0": 38, "k1": 40, "mat": 12, "in": die},
            {"k0": 40, "k1": 44, "mat": 15}, {"k0": 44, "k1": 48, "mat": 16}, {"k0": 48, "k1": 52, "mat": 17},
        ],
        "temperatures": [25.0],
        "objective": {"k_top": 52, "vm_groups": [10, 11], "p": 8.0},
    })
    return sp
