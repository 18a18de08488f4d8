"""The bump submodel golden case (tests/golden/gen_golden.js 'submodel'): the
reference's own set-up data of one corner bump of the bare-die preset, and a
host-side Submodel built from it."""
import gzip
import json
import os

import numpy as np

from conftest import GOLDEN
from fcbga_amd import submodel as sm


def f64(name):
    with gzip.open(os.path.join(GOLDEN, name)) as f:
        return np.frombuffer(f.read(), np.float64).copy()


def golden():
    with open(os.path.join(GOLDEN, "submodel_golden.json")) as f:
        return json.load(f)


def build(variant):
    g = golden()
    s = g["setup"]
    nx, ny, nz = len(s["xs"]), len(s["ys"]), len(s["zs"])
    ijk = np.array(s["eIJK"]).reshape(-1, 3)
    cells = (ijk[:, 0] * (ny - 1) + ijk[:, 1]) * (nz - 1) + ijk[:, 2]
    cell_mat = np.full((nx - 1) * (ny - 1) * (nz - 1), -1, np.int8)
    cell_mat[cells] = np.array(s["eGroup"], np.int8)
    solder = sm.Tables(f64("submodel_solderE.f64.gz"), f64("submodel_solderF.f64.gz"), s["solder"]["nu"])
    uf = sm.Tables(f64("submodel_ufE.f64.gz"), f64("submodel_ufF.f64.gz"), s["ufNu"])
    model = sm.Submodel(s, cell_mat, solder, uf, f64("submodel_subD.f64.gz"), f64("submodel_subF.f64.gz"),
                        f64("submodel_siD.f64.gz"), f64("submodel_siF.f64.gz"), s["boundaryTemps"],
                        f64("submodel_boundary.f64.gz"), s["tempsAll"], variant=variant)
    return g, model, cells
