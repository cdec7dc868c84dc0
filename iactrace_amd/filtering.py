"""Selecting trainable arrays of a telescope by path pattern
(iactrace/utils/filtering.py:14-112, iactrace/utils/paths.py:18-56).

Every array of a telescope has a dotted path, with list positions written
``[i]``: ``mirror_groups.[0].rotations``, ``obstruction_groups.[1].p1``,
``sensors.[0].position``.  Patterns select paths component by component:
``*`` matches one component, ``**`` any number (including none), anything
else must match exactly -- ``'mirror_groups.*.rotations'``, ``'**.rotations'``,
``'mirror_groups.**'``.  A spec is a pattern, a list of patterns, or a
callable ``(path, array) -> bool``.

``make_filter`` / ``partition`` / ``trainable_params`` / ``show_structure``
keep the reference's meaning (trainable OR frozen, default everything
trainable); they return dictionaries keyed by path.  ``parameters`` is the
torch side: leaf tensors with ``requires_grad`` for the selected mirror-group
arrays, and a function that rebuilds the telescope around them, so any torch
optimizer drives the fit (the render's gradients flow to stage-0 rotations
and positions and to every group's aspheric coefficients).
"""
from __future__ import annotations

import dataclasses
from typing import Any, Callable, Sequence, Union

import numpy as np
import torch

__all__ = ["leaves", "match_path", "make_filter", "partition", "trainable_params",
           "show_structure", "parameters"]

PathSpec = Union[str, Sequence[str], Callable[[str, Any], bool]]


def _is_array(v):
    return isinstance(v, (torch.Tensor, np.ndarray))


def _fields(obj):
    if dataclasses.is_dataclass(obj):
        return [(f.name, getattr(obj, f.name)) for f in dataclasses.fields(obj)]
    return list(vars(obj).items())


def leaves(tel):
    """(path, array) for every array of the telescope, in a fixed order."""
    out = []
    for part in ("mirror_groups", "obstruction_groups", "sensors"):
        for i, obj in enumerate(getattr(tel, part) or []):
            for name, v in _fields(obj):
                if _is_array(v):
                    out.append((f"{part}.[{i}].{name}", v))
    return out


def _match(parts, pat):
    if not pat:
        return not parts
    if not parts:
        # the reference's rule: leftover '*' components (or a lone '**') match
        return pat == ["**"] or all(p == "*" for p in pat)
    head, rest = pat[0], pat[1:]
    if head == "**":
        return not rest or any(_match(parts[i:], rest) for i in range(len(parts) + 1))
    return (head == "*" or head == parts[0]) and _match(parts[1:], rest)


def match_path(path, pattern):
    """Glob match of dotted components ('*' one component, '**' any number),
    with the reference's semantics (paths.py:18-56)."""
    return _match(path.split("."), pattern.split("."))


def _spec_fn(spec):
    if callable(spec):
        return spec
    pats = [spec] if isinstance(spec, str) else list(spec)
    return lambda path, leaf: any(match_path(path, p) for p in pats)


def make_filter(tel, trainable: PathSpec | None = None, frozen: PathSpec | None = None):
    """{path: bool} -- True for trainable arrays."""
    if trainable is not None and frozen is not None:
        raise ValueError("Specify trainable OR frozen, not both")
    if trainable is not None:
        fn = _spec_fn(trainable)
        return {p: bool(fn(p, v)) for p, v in leaves(tel)}
    fn = _spec_fn(frozen if frozen is not None else [])
    return {p: not fn(p, v) for p, v in leaves(tel)}


def partition(tel, trainable: PathSpec | None = None, frozen: PathSpec | None = None):
    """({path: array} trainable, {path: array} frozen)."""
    keep = make_filter(tel, trainable, frozen)
    tr, fr = {}, {}
    for p, v in leaves(tel):
        (tr if keep[p] else fr)[p] = v
    return tr, fr


def trainable_params(tel, spec: PathSpec):
    """{path: array} of the arrays ``spec`` selects."""
    return partition(tel, trainable=spec)[0]


def show_structure(tel):
    """Print every array path with its shape and dtype."""
    print("Model structure:")
    for p, v in leaves(tel):
        print(f"  {p}: {tuple(v.shape)} {v.dtype}")


def parameters(tel, trainable: PathSpec | None = None, frozen: PathSpec | None = None):
    """(params, rebuild): leaf tensors (requires_grad) for the selected
    mirror-group arrays, and ``rebuild(params) -> Telescope`` using them.
    Selecting an array outside the mirror groups raises (the render has no
    gradient for obstruction or sensor arrays)."""
    from .telescope import Telescope
    sel = [p for p, keep in make_filter(tel, trainable, frozen).items() if keep]
    params, where = [], []
    for p in sel:
        part, idx, name = p.split(".")
        if part != "mirror_groups":
            raise ValueError(f"{p}: only mirror-group arrays can be trained")
        v = getattr(tel.mirror_groups[int(idx[1:-1])], name)
        if not isinstance(v, torch.Tensor) or not v.is_floating_point():
            continue
        params.append(v.detach().clone().requires_grad_(True))
        where.append((int(idx[1:-1]), name))

    def rebuild(ps):
        groups = list(tel.mirror_groups)
        for (gi, name), t in zip(where, ps):
            groups[gi] = groups[gi].replace(**{name: t})
        out = Telescope(groups, tel.obstruction_groups, tel.sensors, tel.name)
        out.launch = tel.launch
        return out

    return params, rebuild
