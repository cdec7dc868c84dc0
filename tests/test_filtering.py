"""Trainable-parameter selection by path pattern (CPU): the pattern rules of
the reference's paths.py, trainable/frozen partitions of a telescope, and the
torch parameters / rebuild round trip."""
import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs
from iactrace_amd.filtering import leaves, match_path, parameters


@pytest.mark.parametrize("path,pattern,expected", [
    ("mirror_groups.[0].rotations", "mirror_groups.*.rotations", True),
    ("mirror_groups.[3].rotations", "**.rotations", True),
    ("mirror_groups.[0].positions", "**.rotations", False),
    ("mirror_groups.[0].rotations", "mirror_groups.**", True),
    ("mirror_groups.[0].rotations", "mirror_groups.rotations", False),
    ("mirror_groups.[0].rotations", "mirror_groups.*", False),
    ("sensors.[0].position", "sensors.*.position", True),
    ("a", "a.*", True),               # leftover '*' components match an exhausted path
    ("a", "a.**", True),
    ("a.b", "a.**.b", True),
    ("a.x.y.b", "a.**.b", True),
    ("a.x.y.c", "a.**.b", False),
    ("a.b", "**", True),
])
def test_match_path(path, pattern, expected):
    assert match_path(path, pattern) is expected


def _tel():
    wl = configs.workload("C3", n_sources=2, n_samples=4)
    groups = ia.group_mirrors(wl.mirrors)
    return ia.Telescope(groups, ia.group_obstructions(wl.obstructions), [wl.sensor])


def test_make_filter_partition_and_structure(capsys):
    tel = _tel()
    paths = [p for p, _ in leaves(tel)]
    assert "mirror_groups.[0].rotations" in paths and "sensors.[0].position" in paths
    f = ia.make_filter(tel, trainable="mirror_groups.*.rotations")
    assert f["mirror_groups.[0].rotations"] and not f["mirror_groups.[0].positions"]
    f = ia.make_filter(tel, frozen=["**.points", "sensors.**"])
    assert not f["mirror_groups.[0].points"] and not f["sensors.[0].position"]
    assert f["mirror_groups.[0].rotations"]
    assert all(ia.make_filter(tel).values())            # default: everything trainable
    with pytest.raises(ValueError):
        ia.make_filter(tel, trainable="**", frozen="**")
    tr, fr = ia.partition(tel, trainable=lambda path, leaf: path.endswith("positions"))
    assert set(tr) == {"mirror_groups.[0].positions"} and len(fr) == len(paths) - 1
    assert list(ia.trainable_params(tel, "**.aspheric")) == ["mirror_groups.[0].aspheric"]
    ia.show_structure(tel)
    assert "mirror_groups.[0].rotations" in capsys.readouterr().out


def test_parameters_rebuild_roundtrip():
    tel = _tel()
    params, rebuild = parameters(tel, trainable=["mirror_groups.*.rotations",
                                                 "mirror_groups.*.positions"])
    assert len(params) == 2 and all(p.requires_grad and p.is_leaf for p in params)
    opt = torch.optim.Adam(params, lr=0.1)
    (params[0].sum() + 2 * params[1].sum()).backward()
    opt.step()
    t2 = rebuild(params)
    g0, g2 = tel.mirror_groups[0], t2.mirror_groups[0]
    assert torch.allclose(g2.rotations.detach(), g0.rotations + -0.1, atol=1e-5)
    assert torch.allclose(g2.positions.detach(), g0.positions - 0.1, atol=1e-5)
    assert g2.points is g0.points
    with pytest.raises(ValueError):
        parameters(tel, trainable="sensors.*.position")
