"""The reference-shaped C++ caller (include/gdem_compat.h, driven by
gpudem_amd/compat_demo.cpp the way main.cu drives its device path) on the
100-sphere lattice drop (cell list of cell 2r rebuilt every step, 20,000
steps, output every 1,000), checked bit for bit against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT
from gpudem_amd import GDEM_FP64, engine, scenes

DEMO = os.path.join(ROOT, "gpudem_amd", "gdem_compat_demo")


def write_stl(path, verts):
    with open(path, "w") as f:
        f.write("solid walls\n")
        for t in np.asarray(verts).reshape(-1, 3, 3):
            f.write("facet normal 0 0 0\n outer loop\n")
            for p in t:
                f.write(f"  vertex {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
            f.write(" endloop\nendfacet\n")
        f.write("endsolid walls\n")


def write_input(path, spec, steps, out_every, precision):
    p = spec.params
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    v = np.zeros((spec.n, 3)) if spec.v is None else spec.v
    with open(path, "w") as f:
        head = [spec.n, steps, out_every, precision, p["dt"], p["mu"], *p["g"], *p["bbox"], p["cell"], p["skin"],
                p["refresh"]]
        f.write(" ".join(f"{h:.17g}" if isinstance(h, float) else str(h) for h in head) + "\n")
        for i in range(spec.n):
            row = [*spec.x[i], *v[i], 0.0, 0.0, 0.0, spec.r[i], m[i], k[i], e[i]]
            f.write(" ".join(f"{q:.17g}" for q in row) + " 1\n")


def test_demo_is_built_and_links_the_library():
    assert os.path.exists(DEMO), "build() compiles gpudem_amd/compat_demo.cpp"
    out = subprocess.run(["ldd", DEMO], capture_output=True, text=True).stdout
    assert "libgdem.so" in out


@pytest.mark.gpu
def test_compat_caller_matches_oracle_on_lattice_drop(tmp_path):
    spec = scenes.lattice_drop()  # cell list, cell 2r, rebuilt every step
    steps, every = 20000, 1000
    write_stl(tmp_path / "walls.stl", spec.verts)
    write_input(tmp_path / "in.txt", spec, steps, every, GDEM_FP64)
    print(f"compat caller: lattice drop n={spec.n}, {steps} steps", flush=True)
    r = subprocess.run([DEMO, str(tmp_path / "in.txt"), str(tmp_path / "walls.stl"), str(tmp_path / "out.txt")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "out.txt")
    orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params())
    orc.step(steps)
    assert np.array_equal(got[:, 0:3], orc.get("x").reshape(-1, 3))
    assert np.array_equal(got[:, 3:6], orc.get("v").reshape(-1, 3))
    assert np.array_equal(got[:, 6:9], orc.get("angv").reshape(-1, 3))
    assert np.array_equal(got[:, 9].astype(np.int32), orc.get("isActive"))
    assert np.any(np.abs(got[:, 3:6]) > 0)
