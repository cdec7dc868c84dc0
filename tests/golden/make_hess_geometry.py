"""Writes tests/golden/hess_ct3.npz and hess_ct5.npz: the geometry of the
reference's two real telescopes (configs/HESS/CT3.yaml, CT5.yaml) as numeric
arrays, so GPU tests can rebuild them where the reference is absent.

Run in the build container (it reads the YAML files under /root/reference):
    python tests/golden/make_hess_geometry.py
The arrays are the parsed numbers of the files (mirror poses, apertures and
surface templates, obstruction primitives, sensor frames, pixel centres and
edge widths); tests/golden/hess.py turns them back into the loader's config
dict.
"""
import os

import numpy as np
import yaml

SRC = "/root/reference/configs/HESS"
HERE = os.path.dirname(os.path.abspath(__file__))


def flatten(cfg):
    out = {}
    tnames = list(cfg["mirror_templates"])
    out["template_curvature"] = np.array(
        [cfg["mirror_templates"][t]["surface"]["curvature"] for t in tnames], np.float64)
    out["template_conic"] = np.array(
        [cfg["mirror_templates"][t]["surface"]["conic"] for t in tnames], np.float64)
    K = max([len(cfg["mirror_templates"][t]["surface"]["aspheric"]) for t in tnames] + [0])
    asph = np.zeros((len(tnames), K))
    nasph = np.zeros(len(tnames), np.int64)
    for i, t in enumerate(tnames):
        a = cfg["mirror_templates"][t]["surface"]["aspheric"]
        asph[i, :len(a)] = a
        nasph[i] = len(a)
    out["template_aspheric"], out["template_n_aspheric"] = asph, nasph
    ms = cfg["mirrors"]
    out["mirror_template"] = np.array([tnames.index(m["template"]) for m in ms], np.int64)
    out["mirror_position"] = np.array([m["position"] for m in ms], np.float64)
    out["mirror_orientation"] = np.array([m["orientation"] for m in ms], np.float64)
    out["mirror_stage"] = np.array([m.get("stage", 0) for m in ms], np.int64)
    out["mirror_offset"] = np.array([m.get("offset", [0.0, 0.0]) for m in ms], np.float64)
    nv = max([len(m["aperture"].get("vertices", [])) for m in ms] + [1])
    radius = np.full(len(ms), np.nan)
    verts = np.full((len(ms), nv, 2), np.nan)
    nverts = np.zeros(len(ms), np.int64)
    for i, m in enumerate(ms):
        ap = m["aperture"]
        if ap["type"] == "circular":
            radius[i] = ap["radius"]
        else:
            v = np.asarray(ap["vertices"], np.float64)
            verts[i, :len(v)] = v
            nverts[i] = len(v)
    out["mirror_radius"], out["mirror_vertices"], out["mirror_n_vertices"] = radius, verts, nverts
    obs = cfg.get("obstructions", [])
    cyl = [o for o in obs if o["type"] == "cylinder"]
    box = [o for o in obs if o["type"] == "box"]
    assert len(cyl) + len(box) == len(obs), "only cylinders and boxes in these files"
    out["cylinder_p1"] = np.array([o["p1"] for o in cyl], np.float64).reshape(-1, 3)
    out["cylinder_p2"] = np.array([o["p2"] for o in cyl], np.float64).reshape(-1, 3)
    out["cylinder_r"] = np.array([o["r"] for o in cyl], np.float64)
    out["box_p1"] = np.array([o["p1"] for o in box], np.float64).reshape(-1, 3)
    out["box_p2"] = np.array([o["p2"] for o in box], np.float64).reshape(-1, 3)
    sens = cfg["sensors"]
    out["n_sensors"] = np.array(len(sens))
    for i, s in enumerate(sens):
        p = f"sensor{i}_"
        out[p + "hexagonal"] = np.array(s["type"] == "hexagonal")
        out[p + "position"] = np.array(s["position"], np.float64)
        out[p + "orientation"] = np.array(s["orientation"], np.float64)
        out[p + "edge_width"] = np.array(s.get("edge_width", 0.0), np.float64)
        if s["type"] == "hexagonal":
            out[p + "centers"] = np.stack([s["centers_x"], s["centers_y"]], 1).astype(np.float64)
        else:
            out[p + "width"] = np.array(s["width"])
            out[p + "height"] = np.array(s["height"])
            out[p + "bounds"] = np.array(s["bounds"], np.float64)
    return out


if __name__ == "__main__":
    for name in ("CT3", "CT5"):
        with open(os.path.join(SRC, f"{name}.yaml")) as f:
            cfg = yaml.safe_load(f)
        arrs = flatten(cfg)
        path = os.path.join(HERE, f"hess_{name.lower()}.npz")
        np.savez_compressed(path, **arrs)
        print(path, os.path.getsize(path), "bytes")
