"""Results files (.fcbgares) and the save / restore of a thermal path whose
increments live on the device (encodeResultsFile, decodeResultsFile,
meshFingerprint, exportAnalysisState, restoreAnalysisState, resultsChunkSteps:
src/core/16-results-file.js:1-149; constants src/core/00-constants.js:132-136;
Analysis.restorePath / finishPath: src/core/11-analysis.js:220-261).

Layout (little endian): bytes 0-7 the magic "FCBGARES", 8-11 the format
version (uint32), 12-15 the manifest length (uint32), the manifest (UTF-8 JSON
{tree, blobs}), then the typed arrays, each at an 8-byte aligned offset from the
start of the data block.  The tree is the saved value with every typed array
replaced by {"__ta": index} (an array referenced twice is stored once), every
Map by {"__map": [[key, value], ...]}, every Set by {"__set": [...]} and every
non-finite number by {"__num": "NaN" | "Infinity" | "-Infinity"}.

Python stand-ins for the JavaScript values: a typed array is a 1-D numpy array
of one of the eight element types below, a Map is a JsMap, a Set a JsSet, a
plain object a dict (string keys), an array a list; None is null (JavaScript's
undefined has no stand-in: a key that would hold it is left out).  The manifest
is written exactly as JSON.stringify writes it (integer-like keys of an object
first in ascending order, numbers in the shortest round-trip form of
Number.prototype.toString, -0 as 0), so a file written here is byte-identical to
the reference's for the same value.
"""
from __future__ import annotations

import json
import math
import re
import struct

import numpy as np

RESULTS_FILE_MAGIC = "FCBGARES"       # src/core/00-constants.js:132
RESULTS_FILE_VERSION = 2              # :133
RESULTS_FILE_ALIGN = 8                # :134
RESULTS_FILE_EXT = ".fcbgares"        # :135
RESULTS_FILE_CHUNK_BYTES = 64 * 1048576  # :136

TYPED = {"Float64Array": np.float64, "Float32Array": np.float32, "Int32Array": np.int32, "Uint32Array": np.uint32,
         "Int16Array": np.int16, "Uint16Array": np.uint16, "Int8Array": np.int8, "Uint8Array": np.uint8}
_TYPE_NAME = {np.dtype(v): k for k, v in TYPED.items()}


class JsMap(dict):
    """A JavaScript Map (insertion-ordered keys of any JSON-able kind)."""


class JsSet(list):
    """A JavaScript Set (insertion-ordered members)."""


class ResultsFileError(ValueError):
    pass


# ---------------------------------------------------------------------------
# JSON exactly as JSON.stringify writes it
# ---------------------------------------------------------------------------

def js_number(x) -> str:
    """Number.prototype.toString of a finite double (ECMAScript Number::toString, radix 10)."""
    x = float(x)
    if x == 0:
        return "0"
    if x < 0:
        return "-" + js_number(-x)
    mant, _, ex = repr(x).partition("e")
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    q = (int(ex) if ex else 0) - len(fp)       # value = int(ip + fp) * 10^q
    stripped = digits.rstrip("0")
    q += len(digits) - len(stripped)
    s, k = stripped, len(stripped)
    n = q + k                                  # value = 0.s * 10^n
    if k <= n <= 21:
        return s + "0" * (n - k)
    if 0 < n <= 21:
        return s[:n] + "." + s[n:]
    if -6 < n <= 0:
        return "0." + "0" * (-n) + s
    e = n - 1
    sign = "+" if e >= 0 else "-"
    return (s if k == 1 else s[0] + "." + s[1:]) + "e" + sign + str(abs(e))


_SURROGATE = re.compile("[\ud800-\udfff]")


def _js_string(s: str) -> str:
    out = json.dumps(s, ensure_ascii=False)
    # well-formed JSON.stringify writes a lone surrogate as an escape
    return _SURROGATE.sub(lambda m: "\\u%04x" % ord(m.group(0)), out)


_INDEX_KEY = re.compile(r"^(0|[1-9][0-9]*)$")


def _ordered_keys(d: dict):
    """Property order of a JavaScript object: array-index keys ascending, then the rest as inserted."""
    keys = [str(k) for k in d.keys()]
    idx = sorted((k for k in keys if _INDEX_KEY.match(k) and int(k) < 4294967295), key=int)
    rest = [k for k in keys if k not in set(idx)]
    return idx + rest


def _stringify(v) -> str:
    if v is None:
        return "null"
    if v is True:
        return "true"
    if v is False:
        return "false"
    if isinstance(v, (int, float, np.integer, np.floating)):
        return js_number(v) if math.isfinite(float(v)) else "null"
    if isinstance(v, str):
        return _js_string(v)
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(_stringify(x) for x in v) + "]"
    if isinstance(v, dict):
        strkeyed = {str(k): x for k, x in v.items()}
        return "{" + ",".join(_js_string(k) + ":" + _stringify(strkeyed[k]) for k in _ordered_keys(v)) + "}"
    raise TypeError(f"cannot write a {type(v).__name__} as JSON")


# ---------------------------------------------------------------------------
# value trees (packTree / unpackTree, src/core/16-results-file.js:15-50)
# ---------------------------------------------------------------------------

def _is_typed(v) -> bool:
    return isinstance(v, np.ndarray)


def pack_tree(v, blobs: list, seen: dict):
    if isinstance(v, (bool, np.bool_)) or v is None or isinstance(v, str):
        return bool(v) if isinstance(v, np.bool_) else v
    if isinstance(v, (int, float, np.integer, np.floating)):
        f = float(v)
        if math.isfinite(f):
            return v
        return {"__num": "NaN" if math.isnan(f) else ("Infinity" if f > 0 else "-Infinity")}
    if _is_typed(v):
        if v.ndim != 1 or v.dtype not in _TYPE_NAME:
            raise ResultsFileError(f"Cannot save a {v.dtype} array of shape {v.shape}.")
        if id(v) not in seen:
            blobs.append(v)
            seen[id(v)] = len(blobs) - 1
        return {"__ta": seen[id(v)]}
    if isinstance(v, JsMap):
        return {"__map": [[pack_tree(k, blobs, seen), pack_tree(x, blobs, seen)] for k, x in v.items()]}
    if isinstance(v, JsSet):
        return {"__set": [pack_tree(x, blobs, seen) for x in v]}
    if isinstance(v, (list, tuple)):
        return [pack_tree(x, blobs, seen) for x in v]
    if isinstance(v, dict):
        return {str(k): pack_tree(x, blobs, seen) for k, x in v.items()}
    raise ResultsFileError(f"Cannot save a {type(v).__name__}.")


def unpack_tree(v, arrays):
    if v is None or not isinstance(v, (dict, list)):
        return v
    if isinstance(v, list):
        return [unpack_tree(x, arrays) for x in v]
    if "__ta" in v:
        return arrays[v["__ta"]]
    if "__num" in v:
        return float(v["__num"])
    if "__map" in v:
        return JsMap((unpack_tree(k, arrays), unpack_tree(x, arrays)) for k, x in v["__map"])
    if "__set" in v:
        return JsSet(unpack_tree(x, arrays) for x in v["__set"])
    return {k: unpack_tree(x, arrays) for k, x in v.items()}


def _align(n: int, a: int = RESULTS_FILE_ALIGN) -> int:
    return -(-n // a) * a


def encode_results_file(value) -> bytes:
    """A value tree as a results file (encodeResultsFile + joinParts, src/core/16-results-file.js:55-81)."""
    blobs: list = []
    tree = pack_tree(value, blobs, {})
    table, off = [], 0
    for b in blobs:
        off = _align(off)
        table.append({"type": _TYPE_NAME[b.dtype], "offset": off, "length": int(b.shape[0])})
        off += b.nbytes
    manifest = _stringify({"tree": tree, "blobs": table}).encode("utf-8")
    data_start = _align(16 + len(manifest))
    size = data_start + (table[-1]["offset"] + blobs[-1].nbytes if blobs else 0)
    out = bytearray(size)
    out[0:8] = RESULTS_FILE_MAGIC.encode("ascii")
    out[8:16] = struct.pack("<II", RESULTS_FILE_VERSION, len(manifest))
    out[16:16 + len(manifest)] = manifest
    for b, e in zip(blobs, table):
        raw = np.ascontiguousarray(b).astype(b.dtype.newbyteorder("<"), copy=False).tobytes()
        out[data_start + e["offset"]:data_start + e["offset"] + len(raw)] = raw
    return bytes(out)


def decode_results_file(buf: bytes):
    """The value tree of a results file; typed arrays are copied out of the buffer
    (decodeResultsFile, src/core/16-results-file.js:83-100)."""
    buf = bytes(buf)
    if len(buf) < 16:
        raise ResultsFileError("Not a results file (too short).")
    if buf[:8] != RESULTS_FILE_MAGIC.encode("ascii"):
        raise ResultsFileError("Not an FCBGA screening tool results file.")
    version, n = struct.unpack("<II", buf[8:16])
    if version > RESULTS_FILE_VERSION:
        raise ResultsFileError("This results file was written by a newer version of the tool "
                               f"(format {version}).")
    man = json.loads(buf[16:16 + n].decode("utf-8"))
    data_start = _align(16 + n)
    arrays = []
    for b in man["blobs"]:
        t = TYPED.get(b["type"])
        if t is None:
            raise ResultsFileError("The results file is damaged or incomplete.")
        dt = np.dtype(t).newbyteorder("<")
        lo = data_start + b["offset"]
        hi = lo + b["length"] * dt.itemsize
        if hi > len(buf):
            raise ResultsFileError("The results file is damaged or incomplete.")
        arrays.append(np.frombuffer(buf[lo:hi], dtype=dt).astype(t))
    return unpack_tree(man["tree"], arrays)


# ---------------------------------------------------------------------------
# analysis state
# ---------------------------------------------------------------------------

def mesh_fingerprint(mesh, variant: str = "hip") -> str:
    """The 14-hex-digit mesh hash a results file carries (meshFingerprint,
    src/core/16-results-file.js:104-119), computed by the library's host code."""
    import ctypes as C

    from . import hip_solver as hs
    L = hs.load(variant)
    xs, ys, zs = (np.ascontiguousarray(a, np.float64) for a in (mesh.xs, mesh.ys, mesh.zs))
    eg = np.ascontiguousarray(mesh.e_group, np.int32)
    ep = np.ascontiguousarray(mesh.e_part, np.int8)
    es = np.ascontiguousarray(mesh.e_scale, np.float64)
    out = C.create_string_buffer(15)
    f64 = C.POINTER(C.c_double)
    rc = L.fcbga_mesh_fingerprint(int(mesh.nn), int(mesh.ne), xs.ctypes.data_as(f64), len(xs),
                                  ys.ctypes.data_as(f64), len(ys), zs.ctypes.data_as(f64), len(zs),
                                  eg.ctypes.data_as(C.POINTER(C.c_int32)), ep.ctypes.data_as(C.POINTER(C.c_int8)),
                                  es.ctypes.data_as(f64), out)
    if rc != 0:
        raise hs.SolverError(hs.last_error(L))
    return out.value.decode("ascii")


def results_chunk_steps(ndof: int) -> int:
    """Steps of the thermal path per message when increments travel between the
    page and a worker (resultsChunkSteps, src/core/16-results-file.js:134)."""
    return max(1, RESULTS_FILE_CHUNK_BYTES // (4 * int(ndof)))


def _present(d: dict) -> dict:
    """The keys of d whose value is not None (JavaScript drops undefined ones)."""
    return {k: v for k, v in d.items() if v is not None}


def export_analysis_state(cfg, opts: dict, fingerprint: str, ndof: int, steps, du=None) -> dict:
    """What a results file needs to bring an analysis back without solving
    (exportAnalysisState, src/core/16-results-file.js:121-132): configuration and
    run options, the mesh fingerprint, the plan of the thermal path and (unless du
    is None) every step's FP32 increment."""
    st = _present({"cfg": cfg, "preset": opts.get("preset"), "evaluations": opts.get("evaluations"),
                   "precond": opts.get("precond")})
    st["meshOverride"] = _present({"hmin": opts.get("hmin"), "hmax": opts.get("hmax"), "cap": opts.get("cap")})
    st["fingerprint"] = fingerprint
    st["nDof"] = int(ndof)
    st["plan"] = [[s["Ta"], s["Tb"], s["stage"], s["branch"]] for s in steps]
    if du is not None:
        st["du"] = [np.ascontiguousarray(d, np.float32) for d in du]
    return st


def export_device_path(model, cfg, opts: dict, mesh, steps, with_steps=True, variant="hip") -> dict:
    """exportAnalysisState over a path solved on the device (Model.path_solve):
    the increments are read back in single precision."""
    du = [model.path_du(k) for k in range(len(steps))] if with_steps else None
    return export_analysis_state(cfg, opts, mesh_fingerprint(mesh, variant), model.ndof, steps, du)


def check_restorable(st: dict, ndof: int, fingerprint: str, steps):
    """The checks of restoreAnalysisState and Analysis.restorePath
    (src/core/16-results-file.js:141-148, src/core/11-analysis.js:220-227)."""
    if st["nDof"] != ndof or st["fingerprint"] != fingerprint:
        raise ResultsFileError("this version of the tool meshes the saved configuration differently")
    plan = st["plan"]
    same = len(steps) == len(plan) and all(
        s["Ta"] == p[0] and s["Tb"] == p[1] and s["stage"] == p[2] and s["branch"] == p[3] for s, p in zip(steps, plan))
    if not same:
        raise ResultsFileError("this version of the tool plans the thermal path of the saved configuration differently")
    du = st.get("du") or []
    for k in range(len(steps)):
        if k >= len(du) or du[k] is None or len(du[k]) != ndof:
            raise ResultsFileError("the thermal path in the results file is incomplete")


def restore_device_path(model, st: dict, mesh, steps, path: dict, temps_all, variant="hip"):
    """restoreAnalysisState onto the device: the saved increments replace the
    solves (fcbga_path_load) and the fields are summed on the device as finishPath
    sums them.  Returns (fields, u_cpi, u_ddi)."""
    from .path import finish_path_device
    check_restorable(st, model.ndof, mesh_fingerprint(mesh, variant), steps)
    model.path_reset()
    for k in range(len(steps)):
        model.path_load(k, st["du"][k])
    return finish_path_device(model, steps, path, temps_all)
