"""Results files (.fcbgares) and the save / restore of a device-resident thermal
path (encodeResultsFile, decodeResultsFile, meshFingerprint, exportAnalysisState,
restoreAnalysisState, resultsChunkSteps: src/core/16-results-file.js:1-149;
Analysis.restorePath / finishPath: src/core/11-analysis.js:220-261).

Golden data (tests/golden/gen_golden.js 'results'): the bytes the reference's
encodeResultsFile writes for a value tree holding every kind of value the format
carries (results_tree.fcbgares.gz) and for exportAnalysisState over the
small_box mesh with the preview plan of a stiffener package and synthetic FP32
increments (results_state.fcbgares.gz); meshFingerprint of the small_box and
cfg1 meshes; the fields finishPath sums from the decoded increments;
resultsChunkSteps at five sizes."""
import gzip
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_spec
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec
from fcbga_amd import path as P
from fcbga_amd import results_file as R


def golden():
    with open(os.path.join(GOLDEN, "results_golden.json")) as f:
        return json.load(f)


def fixture_bytes(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return gzip.decompress(f.read())


def value_tree():
    """The tree gen_golden.js encodes, in the Python stand-ins."""
    shared = np.array([1.5, -2.25, 3e-8], np.float32)
    cfg = {"name": "résumé ✓", "packageType": "lid",
           "nested": [1, 0.1, -0.0, 1e21, 1.5e-7, 123456789012, 2 ** 53, -3.25e-300, 5e-324, 1e-6, 2.5e-7, 1e20],
           "flags": [True, False, None], "b": 1, "2": "two", "a": "x", "1": "one", "quote": 'a"b\\c\n\t\u0001'}
    return {"cfg": cfg, "nums": [math.nan, math.inf, -math.inf, 0],
            "fields": R.JsMap([(25, np.array([0.1, 0.2])), (-40, np.array([1.0, 2.0, 3.0]))]),
            "temps": R.JsSet([25, 150, "x"]),
            "arrays": [np.array([math.pi]), shared, shared, np.array([-1, 2], np.int32),
                       np.array([4294967295], np.uint32), np.array([-32768], np.int16),
                       np.array([65535], np.uint16), np.array([-128, 127], np.int8),
                       np.array([0, 255, 7], np.uint8)],
            "empty": np.zeros(0, np.float32)}


def test_js_number_matches_ecmascript():
    # Number::toString (ECMA-262 6.1.6.1.20) at each of its branches
    cases = {1e21: "1e+21", 1e20: "100000000000000000000", 1.5e-7: "1.5e-7", 1e-6: "0.000001", 1e-7: "1e-7",
             123.0: "123", -0.0: "0", 0.5: "0.5", 5e-324: "5e-324", 1.7976931348623157e308: "1.7976931348623157e+308",
             -2.5: "-2.5", 0.1 + 0.2: "0.30000000000000004", 123456.789: "123456.789", 2 ** 53: "9007199254740992"}
    for x, s in cases.items():
        assert R.js_number(x) == s, x


def test_encode_matches_reference_bytes():
    assert R.encode_results_file(value_tree()) == fixture_bytes("results_tree.fcbgares.gz")


def test_decode_reference_tree():
    v = R.decode_results_file(fixture_bytes("results_tree.fcbgares.gz"))
    ref = value_tree()
    assert list(v["cfg"].keys()) == ["1", "2", "name", "packageType", "nested", "flags", "b", "a", "quote"]
    assert v["cfg"]["quote"] == ref["cfg"]["quote"] and v["cfg"]["name"] == ref["cfg"]["name"]
    assert v["cfg"]["nested"] == [0 if x == 0 else x for x in ref["cfg"]["nested"]]
    assert math.isnan(v["nums"][0]) and v["nums"][1:] == [math.inf, -math.inf, 0]
    assert isinstance(v["fields"], R.JsMap) and list(v["fields"].keys()) == [25, -40]
    assert np.array_equal(v["fields"][-40], ref["fields"][-40]) and v["fields"][25].dtype == np.float64
    assert isinstance(v["temps"], R.JsSet) and v["temps"] == [25, 150, "x"]
    for a, b in zip(v["arrays"], ref["arrays"]):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert v["empty"].dtype == np.float32 and v["empty"].size == 0
    # the round trip writes the same bytes again
    assert R.encode_results_file(v) == fixture_bytes("results_tree.fcbgares.gz")


def test_decode_rejects_bad_files():
    good = fixture_bytes("results_tree.fcbgares.gz")
    with pytest.raises(R.ResultsFileError, match="too short"):
        R.decode_results_file(good[:10])
    with pytest.raises(R.ResultsFileError, match="Not an FCBGA"):
        R.decode_results_file(b"FCBGAREZ" + good[8:])
    with pytest.raises(R.ResultsFileError, match="newer version"):
        R.decode_results_file(good[:8] + (3).to_bytes(4, "little") + good[12:])
    with pytest.raises(R.ResultsFileError, match="damaged or incomplete"):
        R.decode_results_file(good[:-2])


@pytest.mark.parametrize("name", ["small_box", "cfg1"])
def test_mesh_fingerprint_matches_reference(name):
    g = golden()["fingerprints"][name]
    mesh = meshspec.build_mesh(golden_spec(name))
    assert (mesh.nn, mesh.ne) == (g["nn"], g["ne"])
    assert R.mesh_fingerprint(mesh, variant="hostemu") == g["fingerprint"]
    # any change of a hashed array changes the hash
    mesh.e_scale = mesh.e_scale.copy()
    mesh.e_scale[-1] *= 1 + 2 ** -52
    assert R.mesh_fingerprint(mesh, variant="hostemu") != g["fingerprint"]


def test_results_chunk_steps():
    for n, k in golden()["chunkSteps"]:
        assert R.results_chunk_steps(n) == k


def _state_inputs():
    g = golden()["state"]
    mesh = meshspec.build_mesh(golden_spec("small_box"))
    steps, path, temps = P.build_path(g["cfg"], [140], {0: [140], 1: [140], 2: [140], 3: [140]}, [217], "preview")
    ndof = g["ndof"]
    i = np.arange(ndof)
    du = [((k + 1) * 0.001 * ((i % 97) - 48)).astype(np.float32) for k in range(len(steps))]
    return g, mesh, steps, path, temps, du


def test_export_state_matches_reference_bytes():
    g, mesh, steps, path, temps, du = _state_inputs()
    assert len(steps) == g["nsteps"]
    fp = R.mesh_fingerprint(mesh, variant="hostemu")
    assert fp == g["fingerprint"]
    st = R.export_analysis_state(g["cfg"], g["opts"], fp, g["ndof"], steps, du)
    assert R.encode_results_file(st) == fixture_bytes("results_state.fcbgares.gz")


def test_restore_on_device_path_sums_like_finish_path():
    """The decoded FP32 increments loaded into the device path (host build) and
    summed there give the reference's finishPath fields bit for bit."""
    g, mesh, steps, path, temps, _ = _state_inputs()
    st = R.decode_results_file(fixture_bytes("results_state.fcbgares.gz"))
    model = hs.Model(mesh, variant="hostemu")
    assert model.ndof == g["ndof"]
    fields, u_cpi, _ = R.restore_device_path(model, st, mesh, steps, path, temps["all"], variant="hostemu")
    ref = golden()["restored"]
    assert sorted(fields) == sorted(float(T) for T in ref["fields"])
    for T, v in ref["fields"].items():
        assert np.array_equal(fields[float(T)], np.array(v)), T
    assert (u_cpi is None) == (ref["uCPI"] is None)


def test_restore_refuses_other_meshes_and_plans():
    g, mesh, steps, path, temps, _ = _state_inputs()
    st = R.decode_results_file(fixture_bytes("results_state.fcbgares.gz"))
    fp, n = g["fingerprint"], g["ndof"]
    R.check_restorable(st, n, fp, steps)
    with pytest.raises(R.ResultsFileError, match="meshes the saved configuration differently"):
        R.check_restorable(st, n, "00000000000000", steps)
    with pytest.raises(R.ResultsFileError, match="meshes the saved configuration differently"):
        R.check_restorable(st, n + 3, fp, steps)
    other = [dict(s) for s in steps]
    other[0]["Tb"] += 1
    with pytest.raises(R.ResultsFileError, match="plans the thermal path"):
        R.check_restorable(st, n, fp, other)
    cut = dict(st, du=st["du"][:-1])
    with pytest.raises(R.ResultsFileError, match="incomplete"):
        R.check_restorable(cut, n, fp, steps)


@pytest.mark.gpu
def test_device_path_export_and_restore():
    """A path solved on the device, exported (FP32 increments), written, read and
    restored on a fresh model sums to the FP64 sums of the same FP32 values; the
    single-step payload of a partial message is the FP64 increment."""
    spec = golden_spec("plate_16")
    mesh = meshspec.build_mesh(spec)
    cfg = {"packageType": "bare", "process": {"Tsub": 217, "Tjoin": 217, "Tuf": 165, "Tattach": 150},
           "reflow": {"peak": 260, "reportTemps": [25, 150, 220, 260]}, "cycling": {"Tmin": -40, "Tmax": 125}}
    steps, path, temps = P.build_path(cfg, [], {0: [], 1: [], 2: [], 3: []}, [217], "preview")
    model = hs.Model(mesh, variant="hip")
    P.run_path(model, steps, path, temps["all"], lambda s: hs.step_for_spec(spec, s["Ta"], s["Tb"], 0), tol=1e-10)
    for k in (0, len(steps) - 1):
        assert np.array_equal(model.path_du(k), model.path_sum([k]).astype(np.float32))
    opts = {"preset": "draft", "evaluations": "preview", "precond": "mg"}
    st = R.export_device_path(model, cfg, opts, mesh, steps)
    back = R.decode_results_file(R.encode_results_file(st))
    fresh = hs.Model(mesh, variant="hip")
    fields, _, _ = R.restore_device_path(fresh, back, mesh, steps, path, temps["all"])
    du = [d.astype(np.float64) for d in back["du"]]
    ref, _, _ = P.finish_path(steps, path, du, temps["all"])
    assert sorted(fields) == sorted(ref)
    for T in ref:
        assert np.array_equal(fields[T], ref[T]), T


@pytest.mark.gpu
def test_node_addon_device_path():
    """The JS host's view of the same path: node -> Node-API addon -> C ABI."""
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        pytest.skip("node not installed")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "js", "napi_path.js")
    out = subprocess.run([node, script, "plate_16"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["converged"] and rec["rel_err_sum"] <= 1e-9, rec
    assert rec["f32ok"] and rec["restore_exact"] and rec["badLoad"], rec
