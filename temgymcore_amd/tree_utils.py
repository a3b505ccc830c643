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
from typing import Any, Literal


def _is_namedtuple(obj) -> bool:
    return isinstance(obj, tuple) and hasattr(obj, "_asdict") and hasattr(obj, "_fields")


class PathBuilder:
    """One step of a parameter path: `key` of `parent`, fetched by attribute or item."""

    def __init__(self, parent: Any, key: int | str, getter: Literal["attr", "item"]):
        self._parent = parent
        self._original_key = key
        self._getter = getter
        true_parent = self._resolve_parent()
        if not (dataclasses.is_dataclass(true_parent) or isinstance(true_parent, (list, tuple, dict))):
            raise NotImplementedError(f"Unknown key type for cls {type(true_parent)}")

    def __getattr__(self, name: str):
        if name.startswith("__"):
            raise AttributeError(name)
        return type(self)(self, name, "attr")

    def __getitem__(self, idx):
        return type(self)(self, idx, "item")

    def _resolve_root(self):
        return self._parent._resolve_root() if isinstance(self._parent, PathBuilder) else self._parent

    def _resolve_parent(self):
        return self._parent._resolve() if isinstance(self._parent, PathBuilder) else self._parent

    def _resolve(self):
        get_from = self._resolve_parent()
        if self._getter == "attr":
            return getattr(get_from, self._original_key)
        if self._getter == "item":
            return get_from[self._original_key]
        raise ValueError(f"Unknown get with {self._getter}")

    def _build(self, children: tuple = (), original: bool = True) -> tuple:
        """(root, key, key, ...) of this path."""
        this = (self._original_key,) + tuple(children)
        if not isinstance(self._parent, PathBuilder):
            return (self._parent,) + this
        return self._parent._build(this, original=original)

    def __repr__(self):
        root, *keys = self._build()
        return f"PathBuilder({type(root).__name__}, {tuple(keys)})"


class HasParamsMixin:
    """``params``: a copy of the object whose fields are PathBuilders to them
    (reference tree_utils.py:171-181); ``new_with`` copies with replacements."""

    def new_with(self, **kwargs):
        return dataclasses.replace(self, **kwargs)

    @property
    def params(self):
        builders = {f.name: PathBuilder(self, f.name, "attr") for f in dataclasses.fields(self) if f.init}
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
