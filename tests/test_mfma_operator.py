"""The fine operator's element products on the matrix cores (FCBGA_ELEM_OP=mfma:
per-tile element groups sorted by type, v_mfma_{f32,f64}_16x16x4 on the full
element tables, node rows summed from LDS).  The host-emulation build runs the
same lists and index rules with every local index checked.

Bar: the operator (unscaled K x) against the oracle's assembled matvec on
the free DOFs to 1e-13 of its largest entry and against the node-row
operator to 1e-14, for mirror-symmetric and coupled materials;
solves against the reference's golden solutions (box, plate, bare-die mesh,
the layered package mesh with superposed two-element cells) as the default
operator is held."""
import numpy as np
import pytest

import package_case as pc
from conftest import coupled_step, golden_f64, golden_meta, golden_spec
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec, post

VARIANTS = ["hostemu", pytest.param("hip", marks=pytest.mark.gpu)]


def _models(monkeypatch, m, variant):
    monkeypatch.setenv("FCBGA_ELEM_OP", "valu")
    A = hs.Model(m, variant=variant)
    monkeypatch.setenv("FCBGA_ELEM_OP", "mfma")
    B = hs.Model(m, variant=variant)
    return A, B


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("name", ["small_box", "plate_16", "cfg1"])
def test_operator_matches_oracle_and_node_rows(oracle, monkeypatch, variant, name):
    spec = golden_spec(name)
    m = meshspec.build_mesh(spec)
    A, B = _models(monkeypatch, m, variant)
    for D, s0, born in (meshspec.group_step_data(spec, spec["Ta"], spec["Tb"], 0), coupled_step(spec)):
        st = hs.Step(spec["Ta"], spec["Tb"], 0, D, s0, born)
        x = np.random.default_rng(7).standard_normal(m.ndof)
        x[m.constraint_dofs] = 0
        ya, yb = A.apply(st, x), B.apply(st, x)
        yo = oracle.System(m, D, s0, born).matvec(x)
        free = np.ones(m.ndof, bool)
        free[m.constraint_dofs] = False
        assert np.abs(yb[free] - yo[free]).max() <= 1e-13 * np.abs(yo).max()
        assert np.abs(yb - ya).max() <= 1e-14 * np.abs(ya).max()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_solves_match_reference(monkeypatch, variant, prec):
    monkeypatch.setenv("FCBGA_ELEM_OP", "mfma")
    for name, bar in (("small_box", 1e-9), ("plate_16", 1e-7)):
        spec = golden_spec(name)
        m = meshspec.build_mesh(spec)
        u, rec = hs.Model(m, variant=variant).solve_step(hs.step_for_spec(spec), tol=1e-12, precision=prec)
        uref = golden_f64(name + "_u.f64.gz")
        assert rec["converged"] and np.abs(u - uref).max() <= bar * np.abs(uref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_cfg1_matches_reference(monkeypatch, prec):
    monkeypatch.setenv("FCBGA_ELEM_OP", "mfma")
    spec = golden_spec("cfg1")
    m = meshspec.build_mesh(spec)
    u, rec = hs.Model(m, variant="hip").solve_step(hs.step_for_spec(spec), tol=1e-10, precision=prec)
    uref = golden_f64("cfg1_u.f64.gz")
    assert rec["converged"] and np.abs(u - uref).max() <= 1e-6 * np.abs(uref).max()
    w = post.warpage_jeita(m, u, meshspec.bga_sites(**spec["bga"]))
    wref = golden_meta("cfg1")["warpage"]["signed"]
    assert abs(w["signed"] - wref) <= 1e-3 * abs(wref)


@pytest.mark.parametrize("variant", VARIANTS)
def test_layered_package_mesh_with_two_element_cells(monkeypatch, variant):
    """The bare-die package mesh holds cells with two superposed elements
    (the slot-1 groups) and layered sections."""
    monkeypatch.setenv("FCBGA_ELEM_OP", "mfma")
    g, m = pc.load("bare")
    assert (m.cell_elems.reshape(-1, 2)[:, 1] >= 0).any()
    u, rec = hs.Model(m, variant=variant).solve_step(pc.step(g, 1), tol=1e-10, precision=hs.PREC_FP64)
    uref = np.array(g["steps"][1]["u"])
    assert rec["converged"] and np.abs(u - uref).max() <= 1e-6 * np.abs(uref).max()
