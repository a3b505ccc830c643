"""Initial ray bundles (reference source.py:11-188, utils.py:46-205, :297-325).

``make_rays`` on a GPU device generates the bundle there (tg_source_rings /
tg_source_disc): only the O(sqrt N) ring table is built on the host.  On the
CPU the numpy generators below produce the same arrays (float64 by default,
as numpy produces them).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from ._types import CoordsXY
from .ray import Ray
from .tree_utils import HasParamsMixin

_P, _I32, _I64, _D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_lib.register_signatures({
    "tg_source_rings": (ctypes.c_int, [_P, _P, _P, _I32, _P, _I32, _I64, _I32, _D, _D, _I32, _P, _P, _P, _P, _P]),
    "tg_source_fibonacci": (ctypes.c_int, [_I64, _D, _D, _D, _D, _I32, _P, _P, _P]),
    "tg_source_disc": (ctypes.c_int, [_I64, ctypes.c_uint64, _D, _I32, _D, _D, _I32, _P, _P, _P, _P, _P]),
})
POINT_SOURCE, PARALLEL_BEAM = 0, 1


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


def _ring_counts(num_points_approx: int, radius: float):
    """(points per ring, ring radii, angle steps) of concentric_rings."""
    num_rings = max(1, int(np.floor((-1 + np.sqrt(1 + 4 * num_points_approx / np.pi)) / 2)))
    num_points_kth_ring = np.round(2 * np.pi * np.arange(1, num_rings + 1)).astype(int)
    num_rings = num_points_kth_ring.size
    points_per_unit = num_points_approx / num_points_kth_ring.sum()
    points_per_ring = np.round(num_points_kth_ring * points_per_unit).astype(int)
    radii = np.linspace(0, radius, num_rings + 1, endpoint=True)[1:]
    with np.errstate(divide="ignore"):
        div_angle = 2 * np.pi / points_per_ring
    return points_per_ring, radii, div_angle


def ring_table(num_points_approx: int, radius: float):
    """Table of concentric_rings for tg_source_rings: (radii, steps,
    ring_first, seg_first, n).

    seg_first lists where multi_cumsum_inplace restarts its sum: its counter
    first restarts after len_0 + 1 points and then every len_k + 1 points, so
    restart k sits at ring boundary k plus k."""
    points_per_ring, radii, div_angle = _ring_counts(num_points_approx, radius)
    n = int(points_per_ring.sum())
    ring_first = np.concatenate([[0], np.cumsum(points_per_ring)[:-1]]).astype(np.int64)
    seg = [0]
    for length in points_per_ring:
        nxt = seg[-1] + int(length) + 1
        if nxt >= n:
            break
        seg.append(nxt)
    return radii.astype(np.float64), div_angle.astype(np.float64), ring_first, np.array(seg, np.int64), n


def _dtype_code(dtype) -> int:
    if dtype == torch.float64:
        return _lib.TG_F64
    if dtype == torch.float32:
        return _lib.TG_F32
    raise ValueError(f"ray sources produce float32 or float64, not {dtype}")


def device_rings(num_points_approx: int, radius: float, kind: int, offset_xy, device, dtype=torch.float64):
    """(x, y, dx, dy) device tensors of a ring source generated on the GPU."""
    radii, steps, ring_first, seg_first, n = ring_table(num_points_approx, radius)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    tr, ts, trf, tsf = t(radii), t(steps), t(ring_first), t(seg_first)
    out = [torch.empty(n, dtype=dtype, device=device) for _ in range(4)]
    st = _lib.load().tg_source_rings(tr.data_ptr(), ts.data_ptr(), trf.data_ptr(), len(radii), tsf.data_ptr(),
                                     len(seg_first), n, kind, float(offset_xy[0]), float(offset_xy[1]),
                                     _dtype_code(dtype), *[o.data_ptr() for o in out], _lib.stream_ptr(device))
    _lib.check(st, "tg_source_rings")
    return out


def device_disc(num: int, scale: float, kind: int, offset_xy, device, dtype=torch.float64, seed: int | None = None):
    """(x, y, dx, dy) device tensors of `num` uniform disc samples generated on the GPU."""
    if seed is None:
        seed = int(np.random.default_rng().integers(0, 2**63))
    out = [torch.empty(num, dtype=dtype, device=device) for _ in range(4)]
    st = _lib.load().tg_source_disc(num, seed, float(scale), kind, float(offset_xy[0]), float(offset_xy[1]),
                                    _dtype_code(dtype), *[o.data_ptr() for o in out], _lib.stream_ptr(device))
    _lib.check(st, "tg_source_disc")
    return out


def concentric_rings(num_points_approx: int, radius: float) -> np.ndarray:
    """Deterministic (y, x) samples on concentric rings (reference utils.py:117-175)."""
    points_per_ring, radii, div_angle = _ring_counts(num_points_approx, radius)
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


def fibonacci_spiral(nb_samples: int, radius: float, alpha=2, device=None, dtype=torch.float64):
    """Fibonacci-spiral disc sampling (reference utils.py:297-325): (x, y)
    numpy arrays, or device tensors generated on the GPU when `device` is one."""
    ga = np.pi * (3.0 - np.sqrt(5.0))
    np_boundary = np.round(alpha * np.sqrt(nb_samples))
    if device is not None and torch.device(device).type == "cuda":
        device = torch.device(device)
        band = float(nb_samples - (np_boundary + 1))
        denom = float(nb_samples - 0.5 * (np_boundary + 1))
        x, y = (torch.empty(nb_samples, dtype=dtype, device=device) for _ in range(2))
        st = _lib.load().tg_source_fibonacci(nb_samples, float(radius), band, denom, float(ga), _dtype_code(dtype),
                                             x.data_ptr(), y.data_ptr(), _lib.stream_ptr(device))
        _lib.check(st, "tg_source_fibonacci")
        return x, y
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

    # (kind, disc radius) of the device generators; None: host only
    _device_kind = None

    def _disc_radius(self) -> float:
        raise NotImplementedError

    def make_rays(self, num: int, random: bool = False, device=None, dtype=torch.float64,
                  seed: int | None = None) -> Ray:
        """Ray bundle of about `num` rays (reference source.py:58-79).

        On a GPU device the bundle is generated there: deterministic rings
        give the reference's points; random=True gives exactly `num` uniform
        disc samples from a Philox stream (`seed`; the reference's unseeded
        rejection sampling returns a random count near `num`)."""
        if device is None:
            device = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
        device = torch.device(device)
        if device.type == "cuda" and self._device_kind is not None and (random or ring_table(num, 1.0)[4] > 1):
            off = self.offset_xy
            if random:
                x, y, dx, dy = device_disc(max(1, int(num)), self._disc_radius(), self._device_kind, off, device,
                                           dtype, seed)
            else:
                x, y, dx, dy = device_rings(num, self._disc_radius(), self._device_kind, off, device, dtype)
            return Ray(x=x, y=y, dx=dx, dy=dy, z=self.z, pathlength=0.0)
        r = self.generate_array(num, random=random)
        if r.shape[0] == 1:
            return Ray(x=float(r[0, 0]), y=float(r[0, 1]), dx=float(r[0, 2]), dy=float(r[0, 3]),
                       z=self.z, pathlength=0.0)

        def col(i):
            return torch.as_tensor(np.ascontiguousarray(r[:, i]), dtype=dtype, device=device)

        return Ray(x=col(0), y=col(1), dx=col(2), dy=col(3), z=self.z, pathlength=0.0)


@dataclasses.dataclass(frozen=True)
class PointSource(Source):
    """Point source with semi-convergence angle (reference source.py:82-135)."""

    z: float
    semi_conv: float
    offset_xy: CoordsXY = CoordsXY(x=0.0, y=0.0)
    _device_kind = POINT_SOURCE

    def _disc_radius(self) -> float:
        return float(self.semi_conv)

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
    _device_kind = PARALLEL_BEAM

    def _disc_radius(self) -> float:
        return float(self.radius)

    def generate_array(self, num: int, random: bool = False) -> np.ndarray:
        yx = random_coords(num) * self.radius if random else concentric_rings(num, self.radius)
        y, x = yx.T
        r = np.zeros((x.size, 5), dtype=np.float64)
        r[:, 0] = x + self.offset_xy[0]
        r[:, 1] = y + self.offset_xy[1]
        r[:, 4] = 1.0
        return r
