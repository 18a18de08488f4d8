"""Bump screening through the temperature cycle (screenJoints, cycleHistory,
screeningSites: src/core/13-fatigue.js:275-455).

Golden data (tests/golden/gen_golden.js 'screen'): the reference's own
screenJoints (AnandPoint.stepColumn and, for the pillar case, stepSeriesColumn)
run on a synthetic bump-layer strain table written out with the results (the
reference's siteStrainTable needs a full package analysis; it is replaced by
the table, so the loop over sites and cycle steps is what is pinned), the
cycle histories and the site subsampling above the screening cap."""
import gzip
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from fcbga_amd import screening as scr


def golden():
    with open(os.path.join(GOLDEN, "screen_golden.json")) as f:
        return json.load(f)


def table_of(name, n, nt):
    with gzip.open(os.path.join(GOLDEN, f"screen_{name}_table.f64.gz")) as f:
        return np.frombuffer(f.read(), np.float64).reshape(n, nt, 6)


def arr(v):
    return np.array([np.nan if x is None else x for x in v], np.float64)


def case_inputs(g, c):
    s = g["solder"]
    P = [s["params"][k] for k in ("A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu")]
    solder = scr.MaterialTables(s["E_table"]["T"], [v / 1000 for v in s["E_table"]["v"]], s["cte"], s["params"]["nu"])
    pillar = None
    if c["pillar"]:
        pm = g["pillar_mat"]
        pillar = scr.MaterialTables(pm["E_table"]["T"], [v / 1000 for v in pm["E_table"]["v"]], pm["cte"], pm["nu"])
    hist = scr.cycle_history(c["cycling"])
    steps = scr.step_data(hist, c["temps"], solder, pillar, c["pillar"]["fs"] if c["pillar"] else 1.0)
    sites, idx = scr.screening_sites(c["d"], scr.SCREEN_MAX_DDI if c["pillar"] else None)
    assert idx is None and len(sites["x"]) == c["nSites"]
    table = table_of(c["name"], c["nSites"], len(c["temps"]))
    i25 = c["temps"].index(25) if 25 in c["temps"] else -1
    return P, steps, table, i25


def check_results(c, dW, dE, stab, rtol):
    for got, key in ((dW, "dW"), (dE, "dE"), (stab, "stab")):
        ref = arr(c[key])
        assert np.array_equal(np.isnan(got), np.isnan(ref)), key
        ok = ~np.isnan(ref)
        if not ok.any():
            continue
        scale = np.abs(ref[ok]).max() if key != "stab" else 1.0
        err = np.abs(got[ok] - ref[ok]).max()
        assert err <= rtol * scale, (key, err, scale)
    assert list(scr.rank(dW)) == c["top"]


def test_cycle_history_matches_reference():
    for h in golden()["histories"]:
        r = scr.cycle_history(h["cycling"])
        assert r["phase"] == h["phase"] and list(r["cycle"]) == h["cycle"]
        assert np.array_equal(r["T"], np.array(h["T"])) and np.array_equal(r["t"], np.array(h["t"]))
    # the reference default (10 C/min, 15 min dwells, -40..125 C): 73 steps for the first cycle, 86 per later one
    h3 = scr.cycle_history({"Tmin": -40, "Tmax": 125, "ramp": 10, "dwell": 15, "cycles": 3})
    assert len(h3["T"]) == 1 + 73 + 86 + 86


def test_screening_sites_subsampling_matches_reference():
    for s in golden()["subsample"]:
        sites, idx = scr.screening_sites(s["d"], s["cap"])
        assert (idx is not None) == s["subsampled"] and len(sites["x"]) == s["n"]
        if s["idx"] is not None:
            assert list(idx) == s["idx"]


@pytest.mark.parametrize("case", ["bump3", "bump2", "ddi2", "ring1"])
def test_oracle_screening_matches_reference(oracle, case):
    g = golden()
    c = next(c for c in g["cases"] if c["name"] == case)
    P, steps, table, i25 = case_inputs(g, c)
    dW, dE, stab = oracle.screen_joints(P, steps, table, len(c["temps"]), i25)
    check_results(c, dW, dE, stab, 1e-9)


def anand_params(g):
    from fcbga_amd import hip_solver as hs
    p = g["solder"]["params"]
    P = hs.FcbgaAnandParams()
    for k in ("A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu"):
        setattr(P, k, p[k])
    return P


VARIANTS = ["hostemu", pytest.param("hip", marks=pytest.mark.gpu)]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("case", ["bump3", "bump2", "ddi2", "ring1"])
def test_screen_joints_matches_reference(variant, case):
    """The C ABI (one device thread per joint) against the reference's
    screenJoints: dW, dE and stab per site (1e-9; the device evaluates the same
    FP64 operations without contraction, so the difference is the libm's last
    bits) and the same top-20 ranking."""
    from fcbga_amd import hip_solver as hs
    g = golden()
    c = next(c for c in g["cases"] if c["name"] == case)
    _, steps, table, i25 = case_inputs(g, c)
    dW, dE, stab = hs.screen_joints(anand_params(g), steps, table, i25, variant=variant)
    check_results(c, dW, dE, stab, 1e-9)


def test_screen_joints_rejects_bad_steps():
    from fcbga_amd import hip_solver as hs
    g = golden()
    c = g["cases"][0]
    _, steps, table, i25 = case_inputs(g, c)
    bad = dict(steps, wi=np.full(steps["nsteps"], len(c["temps"]) - 1, np.int32))
    with pytest.raises(hs.SolverError):
        hs.screen_joints(anand_params(g), bad, table, i25, variant="hostemu")
    with pytest.raises(hs.SolverError):
        hs.screen_joints(anand_params(g), dict(steps, fs=0.0), table, i25, variant="hostemu")
    with pytest.raises(hs.SolverError):
        hs.screen_joints(anand_params(g), steps, table, len(c["temps"]), variant="hostemu")
    dW, dE, stab = hs.screen_joints(anand_params(g), steps, table[:0], i25, variant="hostemu")
    assert len(dW) == 0
