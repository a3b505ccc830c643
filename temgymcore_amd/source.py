"""Initial ray bundles (reference source.py:11-188, utils.py:46-205).

Host-side generators of slope/position samples; the resulting Ray holds
device tensors (float64 by default, as numpy produces them).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from ._types import CoordsXY
from .ray import Ray
from .tree_utils import HasParamsMixin


def multi_cumsum_inplace(values: np.ndarray, partitions: np.ndarray, start: float) -> None:
    """Cumulative sums restarted at every partition (reference utils.py:46-80)."""
    part_idx = 0
    current_part_len = partitions[part_idx]
    part_count = 0
    values[0] = start
    for v_idx in range(1, values.size):
        if current_part_len == part_count:
            part_count = 0
            part_idx += 1
            current_part_len = partitions[part_idx]
            values[v_idx] = start
        else:
            values[v_idx] += values[v_idx - 1]
            part_count += 1


def concentric_rings(num_points_approx: int, radius: float) -> np.ndarray:
    """Deterministic (y, x) samples on concentric rings (reference utils.py:117-175)."""
    num_rings = max(1, int(np.floor((-1 + np.sqrt(1 + 4 * num_points_approx / np.pi)) / 2)))
    num_points_kth_ring = np.round(2 * np.pi * np.arange(1, num_rings + 1)).astype(int)
    num_rings = num_points_kth_ring.size
    points_per_unit = num_points_approx / num_points_kth_ring.sum()
    points_per_ring = np.round(num_points_kth_ring * points_per_unit).astype(int)
    radii = np.linspace(0, radius, num_rings + 1, endpoint=True)[1:]
    div_angle = 2 * np.pi / points_per_ring
    params = np.stack((radii, div_angle), axis=0)
    all_params = np.repeat(params, points_per_ring.tolist(), axis=-1)
    multi_cumsum_inplace(all_params[1, :], points_per_ring, 0.0)
    all_radii = all_params[0, :]
    all_angles = all_params[1, :]
    return np.stack((all_radii * np.sin(all_angles), all_radii * np.cos(all_angles)), axis=-1)


def random_coords(num: int, rng: np.random.Generator | None = None) -> np.ndarray:
    """Uniform points in the unit disc by rejection (reference utils.py:178-205)."""
    rng = np.random.default_rng() if rng is None else rng
    while True:
        yx = rng.uniform(-1, 1, size=(max(1, int(num * 1.28)), 2))
        mask = np.sqrt((yx**2).sum(axis=1)) < 1
        if mask.sum() > 0:
            return yx[mask, :]


def fibonacci_spiral(nb_samples: int, radius: float, alpha=2):
    """Fibonacci-spiral disc sampling (reference utils.py:297-325)."""
    ga = np.pi * (3.0 - np.sqrt(5.0))
    np_boundary = np.round(alpha * np.sqrt(nb_samples))
    ii = np.arange(nb_samples)
    rr = np.where(
        ii > nb_samples - (np_boundary + 1),
        radius,
        radius * np.sqrt((ii + 0.5) / (nb_samples - 0.5 * (np_boundary + 1))),
    )
    rr[0] = 0.0
    phi = ii * ga
    return rr * np.cos(phi), rr * np.sin(phi)


class Source(HasParamsMixin):
    """Base class of ray sources; calling one is a no-op (reference source.py:11-79)."""

    z: float

    def __call__(self, ray):
        return ray

    def generate_array(self, num: int, random: bool = False) -> np.ndarray:
        raise NotImplementedError

    def make_rays(self, num: int, random: bool = False, device=None, dtype=torch.float64) -> Ray:
        r = self.generate_array(num, random=random)
        if r.shape[0] == 1:
            return Ray(x=float(r[0, 0]), y=float(r[0, 1]), dx=float(r[0, 2]), dy=float(r[0, 3]),
                       z=self.z, pathlength=0.0)
        if device is None:
            device = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")

        def col(i):
            return torch.as_tensor(np.ascontiguousarray(r[:, i]), dtype=dtype, device=device)

        return Ray(x=col(0), y=col(1), dx=col(2), dy=col(3), z=self.z, pathlength=0.0)


@dataclasses.dataclass(frozen=True)
class PointSource(Source):
    """Point source with semi-convergence angle (reference source.py:82-135)."""

    z: float
    semi_conv: float
    offset_xy: CoordsXY = CoordsXY(x=0.0, y=0.0)

    def generate_array(self, num: int, random: bool = False) -> np.ndarray:
        dyx = random_coords(num) * self.semi_conv if random else concentric_rings(num, self.semi_conv)
        dy, dx = dyx.T
        r = np.zeros((dx.size, 5), dtype=np.float64)
        r[:, 0] += self.offset_xy[0]
        r[:, 1] += self.offset_xy[1]
        r[:, 2] = dx
        r[:, 3] = dy
        r[:, 4] = 1.0
        return r


@dataclasses.dataclass(frozen=True)
class ParallelBeam(Source):
    """Parallel beam filling a disc of `radius` (reference source.py:138-188)."""

    z: float
    radius: float
    offset_xy: CoordsXY = (0.0, 0.0)

    def generate_array(self, num: int, random: bool = False) -> np.ndarray:
        yx = random_coords(num) * self.radius if random else concentric_rings(num, self.radius)
        y, x = yx.T
        r = np.zeros((x.size, 5), dtype=np.float64)
        r[:, 0] = x + self.offset_xy[0]
        r[:, 1] = y + self.offset_xy[1]
        r[:, 4] = 1.0
        return r
