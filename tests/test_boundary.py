"""The C-ABI library: exported symbols, host-side set-up helpers and the
triangle precompute, checked against the oracle (CPU only, no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from conftest import ROOT, golden
from gpudem_amd import engine, scenes


def _declared():
    txt = open(os.path.join(ROOT, "include", "gdem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gdem_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(engine.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(engine.SYMBOLS) == names


def test_material_matches_oracle_bitwise():
    spec = scenes.polydisperse_bed(n=500)
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    s = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params())
    assert np.array_equal(m, s.get("m"))
    assert np.array_equal(k, s.get("k"))
    assert np.array_equal(e, s.get("etac"))


def test_bbox_rule():
    verts = scenes.box_mesh((-0.05, -0.05, 0.0), (0.05, 0.05, 0.3))
    got = engine.bbox_rule(verts, [-0.01, 0.01, 0.1, 0.4, 0.1, 0.2])
    # min/max over mesh and inlet, each times 1.1 (a positive minimum moves up)
    assert np.allclose(got, [-0.055, 0.055, -0.055, 0.44, 0.0, 0.33], rtol=0, atol=1e-15)
    assert np.array_equal(got, scenes.bbox_rule(verts, [-0.01, 0.01, 0.1, 0.4, 0.1, 0.2]))


def test_params_default():
    p = engine.Params()
    assert engine.lib().gdem_params_default(0.002, ctypes.byref(p)) == 0
    assert p.cell == 0.002 * 4 and p.skin == 0.002 and p.refresh == 0.001


@pytest.mark.parametrize("name", ["simpleBox.stl", "simpleBox60deg.stl", "boxZigzag.stl",
                                  "simpleBoxWithHole.stl", "spiral.stl", "chute"])
def test_mesh_precompute_matches_oracle_bitwise(name):
    if name == "chute":
        v, counts = scenes.helical_chute_mesh(ntri=2000)[0], None
    else:
        v, counts = scenes.read_ascii_stl(golden(name))
    g = engine.mesh_precompute(v, counts)
    m = oracle.Mesh(v, counts)
    P = np.stack([m.get("mx"), m.get("my"), m.get("mz")], 1)
    i0, i1, i2 = m.get("i0"), m.get("i1"), m.get("i2")
    assert np.array_equal(g["ids"], np.stack([i0, i1, i2, m.get("te0"), m.get("te1"), m.get("te2")], 1))
    assert np.array_equal(g["p0"], P[i0]) and np.array_equal(g["p1"], P[i1]) and np.array_equal(g["p2"], P[i2])
    assert np.array_equal(g["n"], np.stack([m.get("nx"), m.get("ny"), m.get("nz")], 1))
    assert np.array_equal(g["d"], m.get("d"))
    for f in ("e01", "e02", "e12"):
        assert np.array_equal(g[f], np.stack([m.get(f + "x"), m.get(f + "y"), m.get(f + "z")], 1))
    for f in ("d00", "d01", "d11", "denom", "d00inv", "d11inv", "d22inv"):
        assert np.array_equal(g[f], m.get(f)), f
    assert np.array_equal(g["order"], m.get("sortedIndex"))


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(engine, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(RuntimeError):
        engine.lib()
