"""solver_output frames compacted on the device (write_frame_bin,
solver_output.cu:259-316): Scene.frame() and gdem_write_frame_compact give
the bytes the host writer gives from a full download."""
import numpy as np
import pytest

from gpudem_amd import GDEM_FP64, GDEM_MIXED, Scene, engine, scenes


def host_frame(x, r, active, lf):
    keep = np.asarray(active) == 1
    return (np.asarray(x)[keep] * lf).astype(np.float32), (np.asarray(r)[keep] * lf).astype(np.float32)


def test_compact_writer_matches_host_writer(tmp_path):
    rng = np.random.default_rng(4)
    n = 300
    x = rng.uniform(-1.0, 1.0, (n, 3))
    r = rng.uniform(0.01, 0.02, n)
    act = (rng.uniform(size=n) < 0.7).astype(np.int32)
    for lf in (1.0, 1000.0):
        a, b = tmp_path / f"host{lf}", tmp_path / f"dev{lf}"
        a.mkdir()
        b.mkdir()
        engine.write_frame_bin(str(a), 77, x, r, act, lf)
        engine.write_frame_compact(str(b), 77, *host_frame(x, r, act, lf))
        name = "result_000000000077.bin"
        assert (a / name).read_bytes() == (b / name).read_bytes()
    c = tmp_path / "empty"
    c.mkdir()
    engine.write_frame_compact(str(c), 1, np.empty((0, 3), np.float32), np.empty(0, np.float32))
    assert (c / "result_000000000001.bin").read_bytes() == np.int32(0).tobytes()
    with pytest.raises(ValueError):
        engine.write_frame_compact(str(c), 2, np.zeros((3, 3), np.float32), np.zeros(2, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_device_frame_equals_host_compaction(ctx, precision, tmp_path):
    # 50,000 spheres, 30 % inactive from the start: the compaction spans many
    # workgroups and skips ids everywhere
    spec = scenes.box_scene(9, 50000)
    rng = np.random.default_rng(11)
    active = (rng.uniform(size=spec.n) >= 0.3).astype(np.int32)
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    print(f"box n={spec.n}, {int(active.sum())} active, precision {precision}", flush=True)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision), active=active)
    sc.step(300)
    # the rebuild at creation sorted the slots by cell, away from id order
    order = sc.debug_order()["order"]
    assert not np.array_equal(order, np.arange(spec.n))
    d = sc.download()
    for lf in (1.0, 1000.0):
        xyz, r = sc.frame(lf)
        hx, hr = host_frame(d["x"], spec.r, d["active"], lf)
        assert xyz.shape == hx.shape and np.array_equal(xyz.view(np.uint32), hx.view(np.uint32))
        assert np.array_equal(r.view(np.uint32), hr.view(np.uint32))
        a, b = tmp_path / f"h{lf}", tmp_path / f"d{lf}"
        a.mkdir()
        b.mkdir()
        engine.write_frame_bin(str(a), 300, d["x"], spec.r, d["active"], lf)
        engine.write_frame_compact(str(b), 300, xyz, r)
        name = "result_000000000300.bin"
        assert (a / name).read_bytes() == (b / name).read_bytes()
    sc.close()


@pytest.mark.gpu
def test_device_frame_drops_particles_that_left(ctx):
    spec = scenes.out_of_bounds_flight()
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=GDEM_FP64), v=spec.v)
    xyz, r = sc.frame()
    assert len(r) == spec.n
    sc.step(8000)  # three of the six have crossed the box (steps 6558, 7043 and 7559)
    d = sc.download()
    assert d["active"].sum() == 3
    xyz, r = sc.frame()
    hx, hr = host_frame(d["x"], spec.r, d["active"], 1.0)
    assert np.array_equal(xyz, hx) and np.array_equal(r, hr)
    sc.close()
