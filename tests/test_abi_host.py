"""CPU tests: the C-ABI library and the host-side model (no GPU needed).

* the built library loads, exports every function include/iactrace_amd.h
  declares, and its struct sizes match the ctypes mirrors;
* invalid descriptors are rejected with status codes before any HIP call;
* YAML parsing, grouping, hex lattice detection and the telescope generators.
"""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import _lib, configs
from iactrace_amd.sensors import build_lookup_table, detect_hex_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "iactrace_amd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(iact_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib._SIGNATURES), set(names) - set(_lib._SIGNATURES)
    assert L.iact_abi_version() == _lib.IACT_ABI_VERSION


def test_struct_sizes_match():
    _lib.check_abi()
    assert _lib.lib().iact_struct_size(99) == -1


def _status(rc):
    return rc, _lib.lib().iact_last_error().decode()


def _valid_square_scene():
    sc = _lib.IactScene()
    s = configs.workload("C1").sensor.to_abi(None)
    sc.sensor = s
    return sc


def test_error_paths_without_gpu():
    L = _lib.lib()
    Lc = _lib.IactLaunch()
    rc, msg = _status(L.iact_trace_fwd(None, None, None, 0, 0, 0, None, None, 0, None, None,
                                       None, None))
    assert rc == _lib.IACT_ERR_INVALID and "null scene" in msg
    sc = _valid_square_scene()
    sc.sensor.kind = 7
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "sensor kind" in msg
    sc = _valid_square_scene()
    sc.facets.n_facets = -1
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "negative" in msg
    sc = _valid_square_scene()
    sc.facets.n_facets, sc.facets.n_samples = 3, 0
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "no samples" in msg
    sc = _valid_square_scene()
    sc.facets.n_facets, sc.facets.n_samples = 3, 4
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "null facet buffer" in msg
    sc = _valid_square_scene()
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 5, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "null sources" in msg
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 9, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "source type" in msg
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 5, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "mode" in msg
    sc.n_surfaces = 99
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "n_surfaces" in msg
    sc = _valid_square_scene()
    sc.n_surfaces = 1
    sc.surfaces[0].n_aspheric = 9
    rc, _ = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                     None, None, None, None))
    assert rc == _lib.IACT_ERR_UNSUPPORTED
    sc = _valid_square_scene()
    sc.obstructions.n_cylinders = 2
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "obstruction" in msg
    sc = _valid_square_scene()
    sc.sensor.width = 5
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "does not match" in msg
    hexs = configs.workload("C3").sensor.to_abi(None)
    sc = _lib.IactScene()
    sc.sensor = hexs
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "lookup table" in msg
    sc = _valid_square_scene()
    sc.n_stages = 1
    sc.stages[0] = 1
    rc, msg = _status(L.iact_trace_fwd(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0,
                                       None, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "stages" in msg
    # sampler
    D = _lib.IactSampleDesc()
    D.n_facets, D.n_samples, D.aperture = 2, 8, 5
    rc, msg = _status(L.iact_sample_facets(ctypes.byref(D), ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), None))
    assert rc == _lib.IACT_ERR_INVALID
    D.aperture = _lib.APERTURE_POLYGON
    D.offsets = 8
    D.vertices = 8
    D.n_vertices = 11
    rc, msg = _status(L.iact_sample_facets(ctypes.byref(D), ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), None))
    assert rc == _lib.IACT_ERR_UNSUPPORTED and "vertices" in msg
    assert L.iact_sample_facets(None, None, None, None, None, None) == _lib.IACT_ERR_INVALID
    # sensor accumulate
    s = configs.workload("C1").sensor.to_abi(None)
    rc, msg = _status(L.iact_sensor_accumulate(ctypes.byref(s), None, None, None, -1, None, None,
                                               None))
    assert rc == _lib.IACT_ERR_INVALID


def test_batched_and_fit_error_paths_without_gpu():
    L = _lib.lib()
    sc = _valid_square_scene()
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), None, 0, None, None, 0, None,
                                               None))
    assert rc == _lib.IACT_ERR_INVALID and "batch" in msg
    B = _lib.IactBatch()
    B.n_instances = 2
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), ctypes.byref(B), 0, None, None,
                                               0, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "offsets" in msg
    off = np.array([0, 3, 1], np.int64)
    B.rotations, B.source_offsets = 8, 8
    B.host_source_offsets = off.ctypes.data_as(ctypes.c_void_p)
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), ctypes.byref(B), 0, None, None,
                                               0, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "decrease" in msg
    off = np.array([0, 3, 5], np.int64)
    B.host_source_offsets = off.ctypes.data_as(ctypes.c_void_p)
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), ctypes.byref(B), 0, None, None,
                                               0, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "sources" in msg
    B.sources, B.values = 8, 8
    Lc = _lib.IactLaunch()
    Lc.deterministic = 1
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), ctypes.byref(B), 0,
                                               ctypes.byref(Lc), None, 0, None, None))
    assert rc == _lib.IACT_ERR_UNSUPPORTED and "deterministic" in msg
    need = L.iact_batch_workspace_size(ctypes.byref(sc), ctypes.byref(B), None)
    assert need >= 3 * 8 and need % 256 == 0
    rc, msg = _status(L.iact_trace_batched_fwd(ctypes.byref(sc), ctypes.byref(B), 0, None,
                                               ctypes.c_void_p(256), need - 1, None, None))
    assert rc == _lib.IACT_ERR_WORKSPACE
    rc, msg = _status(L.iact_trace_batched_bwd(ctypes.byref(sc), ctypes.byref(B), 3, None, None,
                                               None, 0, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "source type" in msg
    # loss / optimiser helpers
    rc, msg = _status(L.iact_l2_loss_grad(None, None, 4, 2, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID
    assert L.iact_l2_loss_grad(None, None, 4, 0, None, None, None) == _lib.IACT_OK
    rc, msg = _status(L.iact_adam_step(None, None, None, None, None, 4, 1e-3, 0.9, 0.999, 1e-8,
                                       None, None))
    assert rc == _lib.IACT_ERR_INVALID and "step" in msg
    rc, msg = _status(L.iact_adam_step(ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8),
                                       ctypes.c_void_p(8), ctypes.c_void_p(8), 4, 1e-3, 1.5,
                                       0.999, 1e-8, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "betas" in msg
    h = ctypes.c_void_p()
    rc, msg = _status(L.iact_fit_graph_create(ctypes.byref(sc), None, 0, None, None, 0, None,
                                              None, None, None, None, None, None, None, None,
                                              1e-3, ctypes.byref(h)))
    assert rc == _lib.IACT_ERR_INVALID and h.value is None
    assert L.iact_graph_launch(None, None) == _lib.IACT_ERR_INVALID
    assert L.iact_graph_destroy(None) == _lib.IACT_OK


def test_autotune_error_paths_and_key():
    L = _lib.lib()
    sc = _valid_square_scene()
    rc, msg = _status(L.iact_autotune(ctypes.byref(sc), None, None, 0, 0, 0, None, None, 0, None,
                                      1, None, None, None))
    assert rc == _lib.IACT_ERR_INVALID and "result" in msg
    best = _lib.IactLaunch()
    rc, msg = _status(L.iact_autotune(ctypes.byref(sc), None, None, 0, 0, _lib.MODE_DEBUG, None,
                                      None, 0, ctypes.c_void_p(8), 1, ctypes.byref(best), None,
                                      None))
    assert rc == _lib.IACT_ERR_INVALID and "image and response" in msg
    from iactrace_amd.autotune import tuning_key
    wl = configs.workload("C2", n_sources=8, n_samples=4)
    from iactrace_amd.mirrors import group_mirrors
    tel = ia.Telescope(group_mirrors(wl.mirrors), [], [wl.sensor])
    k1 = tuning_key(tel, 3000, arch="gfx950:256cu")
    k2 = tuning_key(tel, 4000, arch="gfx950:256cu")
    k3 = tuning_key(tel, 5000, arch="gfx950:256cu")
    assert k1 == k2 != k3 and '"sources": 4096' in k2


def test_pack_sources_and_c5_workload():
    s, v, off = ia.pack_sources([np.zeros((2, 3)), np.zeros((0, 3)), np.ones((3, 3))],
                                [np.ones(2), np.ones(0), np.ones(3)])
    assert s.shape == (5, 3) and v.dtype == np.float32 and off.tolist() == [0, 2, 2, 5]
    with pytest.raises(ValueError):
        ia.pack_sources([np.zeros((2, 3))], [np.ones(3)])
    bw = configs.batch_workload(64)
    n = np.diff(bw.offsets)
    assert len(n) == 64 and n.min() >= 256 and n.max() <= 16384
    assert bw.n_rays == 875 * 1024 * int(bw.offsets[-1])
    assert 4e11 < bw.n_rays < 6e11


def test_error_text_is_thread_local_and_raised():
    with pytest.raises(_lib.IactError) as e:
        _lib.check(_lib.lib().iact_trace_fwd(None, None, None, 0, 0, 0, None, None, 0, None,
                                             None, None, None), "iact_trace_fwd")
    assert "null scene" in str(e.value)


def test_render_requires_gpu_tensors():
    tel = ia.Telescope(ia.group_mirrors([ia.Mirror([0, 0, 0], [0, 0, 0],
                                                   ia.AsphericSurface(0.0, 0.0),
                                                   ia.DiskAperture(1.0))]),
                       [], [configs.workload("C1").sensor])
    with pytest.raises(RuntimeError):
        ia.render(tel, torch.zeros((1, 3)), torch.ones(1), "parallel")


# ------------------------------------------------------------------ host model
def test_yaml_roundtrip_and_grouping(tmp_path):
    b = configs.TelescopeConfigBuilder("T")
    b.add_mirror_template("sph", 1 / 30.0, 0.0, [])
    b.add_mirror_template("asph", 0.05, -1.05, [1e-4, 2e-6])
    v = configs.hex_facet_vertices(0.6)
    for i in range(5):
        b.add_mirror_polygon(f"M{i}", "sph", [i, 0, 0], [1, 2, 0], v)
    for i in range(3):
        b.add_mirror_circular(f"D{i}", "sph", [0, i, 0], [0, 0, 0], 0.3)
    b.add_mirror_circular("S", "asph", [0, 0, 6], [180, 0, 0], 0.5, stage=1, offset=[0.1, 0])
    b.add_obstruction_cylinder("c", [0, 0, 0], [0, 0, 1], 0.1)
    b.add_obstruction_box("b", [1, 1, 1], [0, 0, 0])
    b.add_obstruction_sphere("s", [0, 0, 3], 0.2)
    b.add_square_sensor_array("sq", [0, 0, 15], [0, 0, 0], 10, 8, [-1, 1, -0.8, 0.8], 0.01)
    b.add_hexagon_sensor_array("hx", [0, 0, 15], [0, 0, 0], *configs.hex_camera(100, 0.05).T)
    f = tmp_path / "t.yaml"
    b.save(f)
    tel = ia.load_telescope(str(f), None)
    info = tel.get_info()
    assert info["n_mirrors"] == 9 and info["mirrors_by_stage"] == {0: 8, 1: 1}
    assert [type(g).__name__ for g in tel.mirror_groups] == [
        "AsphericDiskMirrorGroup", "AsphericPolygonMirrorGroup", "AsphericDiskMirrorGroup"]
    assert tel.mirror_groups[2].offsets.tolist() == [[pytest.approx(0.1), 0.0]]
    assert [g.kind for g in tel.obstruction_groups] == ["cylinder", "box", "sphere"]
    assert tel.sensors[0].edge_width == pytest.approx(0.01)
    assert tel.sensors[1].n_pixels == 100
    box = tel.obstruction_groups[1].table()
    assert box.tolist() == [[0, 0, 0, 1, 1, 1]]
    cyl = tel.obstruction_groups[0].table()
    assert np.allclose(cyl[0], [0, 0, 0, 0, 0, 1, 1, 0.1])


def test_unknown_types_raise():
    from iactrace_amd.yaml_loader import _parse_aperture, _parse_obstruction, _parse_sensor
    with pytest.raises(ValueError):
        _parse_aperture({"type": "ellipse"})
    with pytest.raises(ValueError):
        _parse_obstruction({"type": "torus"})
    with pytest.raises(ValueError):
        _parse_sensor({"type": "round"})


@pytest.mark.skipif(not os.path.exists("/root/reference/configs/HESS/CT3.yaml"),
                    reason="reference configs not present")
def test_reference_ct3_config_parses():
    tel = ia.load_telescope("/root/reference/configs/HESS/CT3.yaml", None)
    assert tel.get_mirror_count() == 380
    assert tel.get_obstruction_count() == 33
    hx = tel.sensors[0]
    assert hx.n_pixels == 960 and hx.lookup_table.shape == (41, 36)
    assert hx.edge_width == pytest.approx(0.0004)


def test_hex_grid_detection_and_table():
    for rot in (0.0, 0.25, 1.0):
        c = configs.hex_camera(500, 0.05)
        cs, sn = np.cos(rot), np.sin(rot)
        c = (c @ np.array([[cs, sn], [-sn, cs]])).astype(np.float32)
        size, ang, off = detect_hex_grid(c)
        assert size == pytest.approx(0.05 / np.sqrt(3), rel=1e-5)
        assert np.isclose(np.mod(ang - rot, np.pi / 3), 0, atol=1e-5) or \
            np.isclose(np.mod(ang - rot, np.pi / 3), np.pi / 3, atol=1e-5)
        tab, qmin, rmin = build_lookup_table(c, size, ang, off)
        assert sorted(tab[tab >= 0].tolist()) == list(range(500))


def test_generators():
    dc = configs.davies_cotton_dish()
    assert len(dc) == 382
    p = np.stack([m.position.numpy() for m in dc])
    assert np.allclose(p[:, 2], 15 - np.sqrt(225 - (p[:, 0] ** 2 + p[:, 1] ** 2)), atol=1e-5)
    par = configs.parabolic_dish()
    assert len(par) == 875
    cam = configs.hex_camera(2048)
    assert len(cam) == 2048
    wl = configs.workload("C4")
    assert wl.n_rays == 8192 * 2048
    s = configs.parallel_grid(4, 3, 0.01)
    assert np.allclose(np.linalg.norm(s, axis=1), 1)


def test_operations_units():
    g = ia.group_mirrors([ia.Mirror([0, 0, 0], [1.0, 2.0, 0.0], ia.AsphericSurface(0.0, 0.0),
                                    ia.DiskAperture(1.0)) for _ in range(50)])
    tel = ia.Telescope(g, [], [])
    t2 = tel.apply_roughness(24.0)
    assert float(t2.mirror_groups[0].perturbation_scale[0]) == pytest.approx(
        24.0 * np.pi / (180 * 3600), rel=1e-6)
    t3 = tel.apply_misalignment_to_group(0, 3600.0, 0.0, 7)
    d = (t3.mirror_groups[0].rotations - tel.mirror_groups[0].rotations).numpy()
    assert np.allclose(d[:, 1:], 0) and 0.5 < d[:, 0].std() < 1.5  # 1 deg sigma
    with pytest.raises(IndexError):
        tel.remove_sensor(0)
    with pytest.raises(IndexError):
        tel.remove_obstruction(0)
