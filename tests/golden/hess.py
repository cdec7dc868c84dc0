"""The reference's CT3 / CT5 telescopes from tests/golden/hess_ct*.npz
(written by make_hess_geometry.py), as the YAML loader's config dict."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def config(name):
    z = np.load(os.path.join(HERE, f"hess_{name.lower()}.npz"))
    tn = [f"t{i}" for i in range(len(z["template_curvature"]))]
    cfg = {"telescope": {"name": name}, "mirror_templates": {}, "mirrors": [],
           "obstructions": [], "sensors": []}
    for i, t in enumerate(tn):
        k = int(z["template_n_aspheric"][i])
        cfg["mirror_templates"][t] = {"surface": {
            "curvature": float(z["template_curvature"][i]), "conic": float(z["template_conic"][i]),
            "aspheric": [float(a) for a in z["template_aspheric"][i, :k]]}}
    for i in range(len(z["mirror_template"])):
        if np.isnan(z["mirror_radius"][i]):
            n = int(z["mirror_n_vertices"][i])
            ap = {"type": "polygon", "vertices": z["mirror_vertices"][i, :n].tolist()}
        else:
            ap = {"type": "circular", "radius": float(z["mirror_radius"][i])}
        cfg["mirrors"].append({"template": tn[int(z["mirror_template"][i])],
                               "position": z["mirror_position"][i].tolist(),
                               "orientation": z["mirror_orientation"][i].tolist(),
                               "stage": int(z["mirror_stage"][i]),
                               "offset": z["mirror_offset"][i].tolist(), "aperture": ap})
    for p1, p2, r in zip(z["cylinder_p1"], z["cylinder_p2"], z["cylinder_r"]):
        cfg["obstructions"].append({"type": "cylinder", "p1": p1.tolist(), "p2": p2.tolist(),
                                    "r": float(r)})
    for p1, p2 in zip(z["box_p1"], z["box_p2"]):
        cfg["obstructions"].append({"type": "box", "p1": p1.tolist(), "p2": p2.tolist()})
    for i in range(int(z["n_sensors"])):
        p = f"sensor{i}_"
        s = {"position": z[p + "position"].tolist(), "orientation": z[p + "orientation"].tolist(),
             "edge_width": float(z[p + "edge_width"])}
        if bool(z[p + "hexagonal"]):
            s.update(type="hexagonal", centers_x=z[p + "centers"][:, 0].tolist(),
                     centers_y=z[p + "centers"][:, 1].tolist())
        else:
            s.update(type="square", width=int(z[p + "width"]), height=int(z[p + "height"]),
                     bounds=z[p + "bounds"].tolist())
        cfg["sensors"].append(s)
    return cfg
