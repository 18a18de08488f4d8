"""The reference's tornado fixture (tests/golden/tornado_golden.json.gz and
tornado_blocks.f64.gz, written by gen_golden.js 'tornado'): the lidded default
preset at the Draft mesh, the base case and the low and high case of a set of
its sensitivity inputs, each with its thermal-path plan, step data, mesh and
the reference's outputs, and the reference's tornado table."""
import gzip
import json
import os

import numpy as np

from conftest import GOLDEN
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec
from fcbga_amd import screening as scr

_cache = {}


def load():
    if "g" not in _cache:
        with gzip.open(os.path.join(GOLDEN, "tornado_golden.json.gz")) as f:
            _cache["g"] = json.load(f)
        with open(os.path.join(GOLDEN, "tornado_blocks.f64.gz"), "rb") as f:
            _cache["blocks"] = np.frombuffer(gzip.decompress(f.read()), dtype="<f8").copy()
    return _cache["g"], _cache["blocks"]


def mesh(g, key):
    """The Mesh of a case (meshes are keyed by the reference's meshFingerprint)."""
    m = g["meshes"][key]
    out = meshspec.mesh_from_arrays(m["xs"], m["ys"], m["zs"], m["nodeId"], m["cellElems"], m["eGroup"], m["ePart"],
                                    m["eScale"], m["constraints"], m["featureX"], m["featureY"], m["ngroups"],
                                    m["eSection"], m["nsections"])
    out.groups = [{"name": k} for k in m["groupKeys"]]
    return out


def block(g, blocks, i):
    o = g["blockOffsets"]
    return blocks[o[i]:o[i + 1]]


def step_of(g, blocks):
    """hip_solver.Step of a plan step of a case (its group and section blocks)."""
    def make(s):
        gr = np.array([block(g, blocks, i) for i in s["groups"]])
        D, s0, born = gr[:, :36], gr[:, 36:42], gr[:, 42].astype(np.uint8)
        if s["sections"]:
            sec = np.array([block(g, blocks, i) for i in s["sections"]])
            return hs.Step(s["Ta"], s["Tb"], s["stage"], D, s0, born, sec[:, :108], sec[:, 108:126],
                           sec[:, 126].astype(np.uint8))
        return hs.Step(s["Ta"], s["Tb"], s["stage"], D, s0, born)
    return make


def plan(case):
    return {"chain": case["chain"], "up": case["up"], "down": case["down"], "hasCPI": False, "hasDDI": False,
            "chainJoin": [], "cpi": [], "ddiJoin": [], "ddi": []}


def bga(g, key):
    b = g["meshes"][key]["bga"]
    z = b["zone"]
    return np.array(b["x"]), np.array(b["y"]), (z["x0"], z["x1"], z["y0"], z["y1"])


def geom(g, key):
    return {"siParts": g["meshes"][key]["siParts"]}


def screen(g, case):
    """The keyword arguments of sensitivity.worst_bump_dw for a case: its cycling
    temperatures, bump-layer mid-plane, dies' bump fields, cycling profile and the
    solder's Anand constants and property tables."""
    s = g["solder"]
    P = hs.FcbgaAnandParams()
    for k in ("A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu"):
        setattr(P, k, s["params"][k])
    sc = case["screen"]
    return {"temps": case["temps"]["cycling"], "zmid": sc["zmid"], "dies": sc["dies"], "solder_params": P,
            "solder": scr.MaterialTables.from_grids(s["E"], s["F"], s["params"]["nu"]), "cycling": sc["cycling"]}
