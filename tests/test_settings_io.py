"""demSettings.json loader and output writers (host code, no GPU).

The loader is checked against an independent reading of the reference's
case files (Python's json after normalising the "1000." literals cJSON
accepts), the set-up formulas of main.cu and ParticleSystem.cu, and an
independent restatement of the shuffle inlet's random draws (std::mt19937
seeded with 1 -- numpy's MT19937 with the same initialisation -- and
libstdc++'s two-draw uniform_real_distribution<double>).  The writers are
checked byte for byte against the layouts of solver_output.cu.
"""
import json
import math
import os
import re
import shutil
import struct

import numpy as np
import pytest

from conftest import GOLDEN
from gpudem_amd import engine, scenes

CASES = sorted(f for f in os.listdir(os.path.join(GOLDEN, "cases")) if f.endswith(".json"))


def lenient_json(path):
    text = open(path).read()
    text = re.sub(r"(\d)\.(\s*[,}\]\n])", r"\1.0\2", text)
    return json.loads(text)


def make_case(tmp_path, name, mutate=None):
    """A case folder: the fixture settings plus the STL files they name."""
    d = tmp_path / name.replace(".json", "")
    (d / "geometry").mkdir(parents=True)
    js = open(os.path.join(GOLDEN, "cases", name)).read()
    if mutate:
        js = mutate(js)
    (d / "demSettings.json").write_text(js)
    for stl in ("simpleBox.stl", "spiral.stl"):
        shutil.copy(os.path.join(GOLDEN, stl), d / "geometry" / stl)
    return str(d / "demSettings.json")


@pytest.mark.parametrize("name", CASES)
def test_reference_cases_load(tmp_path, name):
    path = make_case(tmp_path, name)
    ref = lenient_json(path)
    st = engine.load_settings(path)
    p0 = ref["particleTypes"]["particle0"]
    assert st["n"] == ref["inlet"]["numParticle"]
    assert st["radius"] == p0["radius"] and st["cor"] == p0["CoR"] and st["mu"] == p0["mu"]
    assert st["k"] == p0["k"] and st["young"] == 0.0  # legacy stiffness
    assert st["g"] == [ref["gravity"]["x"], ref["gravity"]["y"], ref["gravity"]["z"]]
    sh = ref["inlet"]["shuffle"]
    assert st["inlet_box"] == [sh["minx"], sh["maxx"], sh["miny"], sh["maxy"], sh["minz"], sh["maxz"]]
    r, dens, k = p0["radius"], p0["density"], p0["k"]
    m = dens * 3.14 * r * r * r * 4. / 3.
    assert st["m"] == m
    lc = math.log(p0["CoR"])
    assert st["etaconst"] == -2. * lc * math.sqrt(k / (3.1415 * 3.1415 + lc * lc))
    dt0 = 2. * 3.14159265359 * math.sqrt(m / k) / ref["others"]["dtFactor"]
    out_step = math.floor(ref["others"]["outputTiming"] / dt0)
    assert st["out_step"] == out_step
    assert st["dt"] == ref["others"]["outputTiming"] / out_step
    assert st["steps"] == int(ref["others"]["endTime"] / st["dt"])
    assert st["brute_on"] == 0
    assert os.path.exists(st["walls"][0])


def test_young_modulus_and_case_insensitive_keys(tmp_path):
    def mutate(js):
        js = js.replace('"k":1.0e6', '"youngmodulus":2.5e7').replace('"CoR"', '"cor"')
        return js.replace('"gpuOn"', '"gpuon"') + "\n"
    path = make_case(tmp_path, "100SimpleBox.json", mutate)
    st = engine.load_settings(path)
    r = st["radius"]
    assert st["young"] == 2.5e7
    assert st["k"] == 4. / 3. * (2.5e7 * 0.5) * math.sqrt(r * 0.5) * math.sqrt(0.05 * r)
    assert st["cor"] == 0.3


def mt19937_uniform(seed, count):
    """`count` draws of std::uniform_real_distribution<double>(0, 1) on
    std::mt19937(seed) as libstdc++ forms them: two 32-bit outputs r1, r2,
    (r1 + r2 * 2^32) / 2^64 in double arithmetic."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    raw = bg.random_raw(2 * count).astype(np.float64)
    out = (raw[0::2] + raw[1::2] * 4294967296.0) / 18446744073709551616.0
    return np.minimum(out, np.nextafter(1.0, 0.0))


def test_mt19937_restatement_matches_the_standard():
    # [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 (seed 5489) yields 4123659995
    bg = np.random.MT19937(0)
    bg._legacy_seeding(5489)
    assert int(bg.random_raw(10000)[-1]) == 4123659995


def test_shuffle_inlet_matches_independent_draws(tmp_path):
    path = make_case(tmp_path, "100SimpleBox.json")
    st = engine.load_settings(path)
    x, v = engine.settings_particles(path, st["n"])
    b, r = st["inlet_box"], st["radius"]
    u = iter(mt19937_uniform(1, 200000))
    want = []
    for i in range(st["n"]):
        while True:
            p = np.array([next(u) * (b[1] - b[0]) + b[0], next(u) * (b[3] - b[2]) + b[2],
                          next(u) * (b[5] - b[4]) + b[4]])
            if all(float(np.dot(p - q, p - q)) >= 4 * r * r for q in want):
                want.append(p)
                break
    assert np.array_equal(x, np.array(want))
    assert np.all(v == 0.0)


def ref_reader(s):
    """csv.h's parse_float, restated: the reference's CSV number conversion."""
    neg = s.startswith("-")
    s = s.lstrip("+-")
    mant, _, exp = s.partition("e")
    ip, _, fp = mant.partition(".")
    x = 0.0
    for ch in ip:
        x = x * 10 + int(ch)
    pos = 1.0
    for ch in fp:
        pos /= 10
        x += int(ch) * pos
    if exp:
        e = int(exp)
        if e:
            base = 0.1 if e < 0 else 10.0
            e = abs(e)
            while e != 1:
                if e % 2 == 0:
                    base = base * base
                    e //= 2
                else:
                    x *= base
                    e -= 1
            x *= base
    return -x if neg else x


def test_file_inlet_both_forms(tmp_path):
    csv = tmp_path / "init.csv"
    rows = [(0.5, -1.25, 3e-2, 0.0, -1.0, 2.5), (-0.0625, 0.75, 1.5e1, 0.25, 0.0, -0.5)]
    with open(csv, "w") as f:
        f.write("vz, y,x ,z,vx,vy,extra\n")
        for x, y, z, vx, vy, vz in rows:
            f.write(f"{vz},{y},{x},{z},{vx},{vy},7\n")
    for form in ('"file": "init.csv"', '"file": {"name": "init.csv"}'):
        def mutate(js, form=form):
            js = js.replace('"inputMode":"shuffle"', '"inputMode":"file"')
            js = js.replace('"numParticle":200', '"numParticle":2')
            return js.replace('"file":"initParticle.txt"', form)
        path = make_case(tmp_path / form[10:14].strip('"{ '), "100SimpleBox.json", mutate)
        shutil.copy(csv, os.path.join(os.path.dirname(path), "init.csv"))
        st = engine.load_settings(path)
        assert st["input_mode"] == 1 and st["n"] == 2
        x, v = engine.settings_particles(path, 2)
        want = np.array([[ref_reader(str(q)) for q in r] for r in rows])
        assert np.array_equal(x, want[:, :3])
        assert np.array_equal(v, want[:, 3:])
        assert st["inlet_box"] == [want[:, 0].min(), want[:, 0].max(), want[:, 1].min(), want[:, 1].max(),
                                   want[:, 2].min(), want[:, 2].max()]


def test_csv_numbers_follow_the_reference_reader(tmp_path):
    # 0.1 accumulated as 1 * (1/10): exact here; 0.123 as 1/10 + 2/100 + 3/1000
    # with the powers formed by repeated division -- not strtod's rounding
    vals = ["0.123", "1.1", "3.14159", "2.5e-3", "7e2", "-0.3"]
    csv = tmp_path / "c.csv"
    with open(csv, "w") as f:
        f.write("x,y,z,vx,vy,vz\n")
        f.write(",".join(vals) + "\n")

    def mutate(js):
        js = js.replace('"inputMode":"shuffle"', '"inputMode":"file"').replace('"numParticle":200', '"numParticle":1')
        return js.replace('"file":"initParticle.txt"', '"file":"c.csv"')
    path = make_case(tmp_path, "100SimpleBox.json", mutate)
    shutil.copy(csv, os.path.join(os.path.dirname(path), "c.csv"))
    x, v = engine.settings_particles(path, 1)
    got = list(x[0]) + list(v[0])
    assert got == [ref_reader(s) for s in vals]


def test_frame_bin_layout(tmp_path):
    x = np.array([[1.0, 2.0, 3.0], [0.1, 0.2, 0.3], [-4.0, 5.5, 6.25]])
    r = np.array([0.5, 0.25, 0.125])
    act = np.array([1, 0, 1], np.int32)
    engine.write_frame_bin(str(tmp_path), 1234, x, r, act)
    blob = open(tmp_path / "result_000000001234.bin", "rb").read()
    want = struct.pack("<i", 2) + np.array([x[0], x[2]], np.float32).tobytes() + \
        np.array([r[0], r[2]], np.float32).tobytes()
    assert blob == want


def test_text_writers(tmp_path):
    d = str(tmp_path / "out")
    engine.output_init(d)
    engine.output_init(d)  # exists already: fine
    engine.write_pid_header(d, 1)
    engine.write_pid_row(d, 20, 0.005, 1, [1.0, 2.0, 3.5], [0.1, -0.2, 0.0], [0.0, -9.81, 0.0],
                         [1.0, 0.0, -1.0], [0.0, 0.0, 0.0], 1)
    lines = open(os.path.join(d, "pid1.txt")).read().splitlines()
    assert lines[0].startswith("0_time(s),1_pid,2_x(m)") and lines[0].endswith("17_isActive")
    assert lines[1] == ("0.100000 1 1.000000 2.000000 3.500000 0.100000 -0.200000 0.000000 0.000000 "
                        "-9.810000 0.000000 1.000000 0.000000 -1.000000 0.000000 0.000000 0.000000 1")
    engine.write_initial_pos(d, [[0.5, 0.25, -0.125], [1.0, 2.0, 3.0]], [0.01, 0.02])
    assert open(os.path.join(d, "initialPos.csv")).read() == \
        "1 1 0.020000 1000. 0.500000 0.250000 -0.125000\n2 1 0.040000 1000. 1.000000 2.000000 3.000000\n"


def test_scene_from_settings(tmp_path):
    path = make_case(tmp_path, "100SimpleBox.json")
    spec, st = scenes.from_settings(path)
    assert spec.n == st["n"] and spec.params["dt"] == st["dt"]
    assert spec.params["cell"] == 4 * st["radius"] and spec.params["skin"] == st["radius"]
    assert list(spec.params["bbox"]) == list(scenes.bbox_rule(spec.verts, st["inlet_box"]))
    assert spec.tris_per_file == [len(spec.verts)]
