"""Symbolic parameter paths (reference tree_utils.py:34-181).

``component.params.focal_length`` (and ``ray.params.x``,
``descanner.params.descan_error.pxo_pxi``, ``detector.params.pixel_size[0]``)
builds a ``PathBuilder``: a path from a root object (a model element or the
input ray) through attributes and items to a parameter.  ``run_with_grads``
takes such paths to select the variables it differentiates.

Paths resolve against plain Python objects (dataclasses, NamedTuples, tuples,
lists, dicts), not jax pytrees, and they are keyed by the names the user
wrote (the reference's ``original`` keys).
"""
from __future__ import annotations

import dataclasses
import numbers
from typing import Any


def _is_namedtuple(obj) -> bool:
    return isinstance(obj, tuple) and hasattr(obj, "_asdict") and hasattr(obj, "_fields")


class PathBuilder:
    """A parameter path: a root object and the steps taken from it, each an
    attribute name or an item key.  Attribute access and indexing extend the
    path; `_value()` walks it, `_path()` is (root, key, key, ...)."""

    __slots__ = ("_origin", "_steps")

    def __init__(self, origin: Any, steps: tuple = ()):
        object.__setattr__(self, "_origin", origin)
        object.__setattr__(self, "_steps", steps)
        # the container the last step reads from must be one a path can enter
        holder = _walk(origin, steps[:-1])
        if steps and not (dataclasses.is_dataclass(holder) or isinstance(holder, (list, tuple, dict))):
            raise NotImplementedError(f"a parameter path cannot enter {type(holder).__name__}")

    def __getattr__(self, name: str):
        if name.startswith("__"):
            raise AttributeError(name)
        return PathBuilder(self._origin, self._steps + ((True, name),))

    def __getitem__(self, key):
        return PathBuilder(self._origin, self._steps + ((False, key),))

    def _root(self):
        return self._origin

    def _value(self):
        return _walk(self._origin, self._steps)

    def _path(self) -> tuple:
        return (self._origin,) + tuple(key for _, key in self._steps)

    def __repr__(self):
        return f"PathBuilder({type(self._origin).__name__}, {self._path()[1:]})"


def _walk(obj, steps):
    """The object reached from obj by (by_attribute, key) steps."""
    for by_attr, key in steps:
        obj = getattr(obj, key) if by_attr else obj[key]
    return obj


class HasParamsMixin:
    """``params``: a copy of the object whose fields are PathBuilders to them
    (reference tree_utils.py:171-181); ``new_with`` copies with replacements."""

    def new_with(self, **kwargs):
        return dataclasses.replace(self, **kwargs)

    @property
    def params(self):
        builders = {f.name: PathBuilder(self, ((True, f.name),)) for f in dataclasses.fields(self) if f.init}
        obj = object.__new__(type(self))
        for name, b in builders.items():
            object.__setattr__(obj, name, b)
        return obj


# ---------------------------------------------------------------------------
# helpers for run_with_grads
# ---------------------------------------------------------------------------
def get_at(obj, keys):
    for k in keys:
        if dataclasses.is_dataclass(obj) or (_is_namedtuple(obj) and isinstance(k, str)):
            obj = getattr(obj, k)
        else:
            obj = obj[k]
    return obj


def replace_at(obj, keys, fn):
    """Copy of `obj` with the value at `keys` replaced by fn(value)."""
    if not keys:
        return fn(obj)
    k, rest = keys[0], keys[1:]
    if dataclasses.is_dataclass(obj):
        return dataclasses.replace(obj, **{k: replace_at(getattr(obj, k), rest, fn)})
    if _is_namedtuple(obj):
        if isinstance(k, int):
            k = obj._fields[k]
        return obj._replace(**{k: replace_at(getattr(obj, k), rest, fn)})
    if isinstance(obj, (list, tuple)):
        items = list(obj)
        items[k] = replace_at(items[k], rest, fn)
        return type(obj)(items)
    if isinstance(obj, dict):
        out = dict(obj)
        out[k] = replace_at(out[k], rest, fn)
        return out
    raise NotImplementedError(f"cannot replace inside {type(obj).__name__}")


def scalar_leaves(obj, prefix=()):
    """[(keys, value)] of every scalar parameter under obj (numbers and
    single-element tensors), depth first in field order."""
    import numpy as np
    import torch

    if isinstance(obj, bool):
        return []
    if isinstance(obj, numbers.Number):
        return [(prefix, obj)]
    if isinstance(obj, (torch.Tensor, np.ndarray)):
        return [(prefix, obj)]
    if dataclasses.is_dataclass(obj):
        out = []
        for f in dataclasses.fields(obj):
            if f.init:
                out += scalar_leaves(getattr(obj, f.name), prefix + (f.name,))
        return out
    if _is_namedtuple(obj):
        out = []
        for name in obj._fields:
            out += scalar_leaves(getattr(obj, name), prefix + (name,))
        return out
    if isinstance(obj, (list, tuple)):
        out = []
        for i, v in enumerate(obj):
            out += scalar_leaves(v, prefix + (i,))
        return out
    if isinstance(obj, dict):
        out = []
        for k, v in obj.items():
            out += scalar_leaves(v, prefix + (k,))
        return out
    return []
