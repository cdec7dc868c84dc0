"""The reference's CT3 / CT5 geometry fixtures (tests/golden/hess_ct*.npz,
written by tests/golden/make_hess_geometry.py from configs/HESS/*.yaml) load
through the YAML loader's config path with the counts the files hold (CPU)."""
import numpy as np

import iactrace_amd as ia
from iactrace_amd.yaml_loader import build_telescope
from tests.golden import hess


def test_ct3():
    tel = build_telescope(hess.config("CT3"), None)
    (g,) = tel.mirror_groups
    assert isinstance(g, ia.AsphericDiskMirrorGroup) and len(g) == 380
    assert np.allclose(g.radii.numpy(), 0.3) and g.curvature == np.float32(0.0333)
    assert [len(o) for o in tel.obstruction_groups] == [33]
    cam, lid = tel.sensors
    assert isinstance(cam, ia.HexagonalSensor) and cam.n_pixels == 960
    assert cam.edge_width == np.float32(0.0004)
    assert isinstance(lid, ia.SquareSensor) and (lid.width, lid.height) == (1536, 1024)


def test_ct5():
    tel = build_telescope(hess.config("CT5"), None)
    (g,) = tel.mirror_groups
    assert isinstance(g, ia.AsphericPolygonMirrorGroup) and len(g) == 876
    v = g.vertices[0].numpy()
    # the files' hexagons are clockwise (quirk Q12; harmless at stage 0)
    assert (v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[1, 1] - v[0, 1]) * (v[2, 0] - v[0, 0]) < 0
    assert [len(o) for o in tel.obstruction_groups] == [270, 1]
    assert [s.n_pixels for s in tel.sensors] == [1764, 2048, 1501 * 1431]
