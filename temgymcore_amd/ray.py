"""Ray: the SoA ray state of the drop-in API (reference ray.py:8-166).

Fields are Python scalars (uniform over the bundle) or torch tensors (one
value per ray, any broadcastable shapes).  Instances are immutable; use
``derive()`` for modified copies.  ``_one`` is the homogeneous carrier used
for offsets in 5x5 ray-transfer matrices.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from ._program import RAY_FIELDS, _is_array
from .tree_utils import HasParamsMixin


@dataclasses.dataclass(frozen=True)
class Ray(HasParamsMixin):
    """Parametric ray with positions, slopes, z, and pathlength.

    Mirrors reference ray.py:8-45: x, y (m), dx, dy (rad, paraxial slopes),
    z (m), pathlength (m), _one (homogeneous coordinate, default 1.0).
    """

    x: object
    y: object
    dx: object
    dy: object
    z: object
    pathlength: object
    _one: object = 1.0

    @classmethod
    def origin(cls):
        """Ray(x=0, y=0, dx=0, dy=0, z=0, pathlength=0) (reference ray.py:47-56)."""
        return cls(*((0.0,) * 6))

    @property
    def size(self) -> int:
        """1 for scalar rays, else the common element count (reference ray.py:58-77)."""
        sizes = set()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            sizes.add(int(v.numel()) if isinstance(v, torch.Tensor) else int(np.size(v)))
        assert len(sizes) == 1
        return tuple(sizes)[0]

    def __getitem__(self, arg):
        """Index every field (reference ray.py:79-93)."""
        params = {f.name: getattr(self, f.name)[arg] for f in dataclasses.fields(self)}
        return type(self)(**params)

    def to_ray(self):
        return self

    def item(self):
        """Single-element ray -> Python scalars (reference ray.py:98-117)."""
        params = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            params[f.name] = v.item() if hasattr(v, "item") and hasattr(v, "ndim") else v
        return type(self)(**params)

    def to_vector(self):
        """Every field as an at-least-1-D tensor (reference ray.py:119-125)."""
        params = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if isinstance(v, torch.Tensor):
                params[f.name] = torch.atleast_1d(v)
            elif isinstance(v, np.ndarray):
                params[f.name] = torch.atleast_1d(torch.from_numpy(np.ascontiguousarray(v)))
            else:
                params[f.name] = torch.atleast_1d(torch.as_tensor(v, dtype=torch.float64))
        return type(self)(**params)

    def derive(self, x=None, y=None, dx=None, dy=None, z=None, pathlength=None) -> "Ray":
        """Copy with selected fields replaced; `_one` is kept (reference ray.py:127-166)."""
        return Ray(
            x=x if x is not None else self.x,
            y=y if y is not None else self.y,
            dx=dx if dx is not None else self.dx,
            dy=dy if dy is not None else self.dy,
            z=z if z is not None else self.z,
            pathlength=pathlength if pathlength is not None else self.pathlength,
            _one=self._one,
        )

    def __repr__(self):
        def fmt(v):
            if _is_array(v) and v.ndim > 0:
                return f"{type(v).__name__}{tuple(v.shape)}"
            return repr(v)
        inner = ", ".join(f"{n}={fmt(getattr(self, n))}" for n in RAY_FIELDS if n != "_one")
        return f"{type(self).__name__}({inner})"
