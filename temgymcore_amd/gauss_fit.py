"""Batched Gaussian-decomposition fits on the HIP kernels.

The reference fits one instance at a time: a GaussianRay bundle evaluated on
a detector grid with ``evaluate_gaussian_input_image`` (gaussian.py:372-427)
and the intensity loss of examples/fit_three_gaussians.ipynb (cell 2),

    loss(rays) = mean((|u(rays)|^2 - |u_target|^2)^2),

differentiated with ``jax.grad`` (cell 3).  ``GaussianFitBatch`` runs many
such instances of uneven size (Gaussian counts and detector shapes may all
differ) in one pass of six kernel launches: the forward fields, every
instance's loss, and the gradient of each loss with respect to every
parameter row, from a hand-written adjoint.  ``fit_loss`` wraps it as a
differentiable torch function.

Parameter rows are the [n, 16] float64 rows of ``gaussian.pack_params``:
x, y, dx, dy, pathlength, amp_re, amp_im, waist_x, waist_y, radius_x,
radius_y, wavelength, theta, 0, 0, 0.  Gradient columns follow the same
order (the complex amplitude as its real and imaginary parts; the three
unused columns are zero).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .gaussian import GaussianRay, _pix2m, pack_params

NPARAM = 16


class FitInstance(ctypes.Structure):
    _fields_ = [("gauss_begin", ctypes.c_int64), ("n_gauss", ctypes.c_int64), ("pix_begin", ctypes.c_int64),
                ("height", ctypes.c_int32), ("width", ctypes.c_int32), ("pix2m", ctypes.c_double * 6)]


class FitPlan(ctypes.Structure):
    _fields_ = [("n_inst", ctypes.c_int32), ("n_items", ctypes.c_int32), ("n_loss_blocks", ctypes.c_int32),
                ("n_moment_blocks", ctypes.c_int32), ("n_gauss", ctypes.c_int64), ("n_assigned", ctypes.c_int64),
                ("n_pix", ctypes.c_int64), ("cut", ctypes.c_float), ("tile_shape", ctypes.c_int32)] + [
        (name, ctypes.c_int64) for name in (
            "off_inst", "off_items", "off_loss_blocks", "off_order", "off_gauss_inst", "off_coef", "off_unit",
            "off_partial", "off_weight", "off_loss_part", "off_moments", "table_bytes", "workspace_bytes")]


_P = ctypes.c_void_p
_lib.register_signatures({
    "tg_gauss_fit_plan": (ctypes.c_int, [ctypes.POINTER(FitInstance), ctypes.c_int32, ctypes.c_int64,
                                         ctypes.c_double, ctypes.POINTER(FitPlan)]),
    "tg_gauss_fit_upload": (ctypes.c_int, [ctypes.POINTER(FitPlan), ctypes.POINTER(FitInstance), _P, _P]),
    "tg_gauss_fit_run": (ctypes.c_int, [ctypes.POINTER(FitPlan), _P, _P, _P, _P, _P, _P, _P]),
})


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


class GaussianFitBatch:
    """Fixed geometry of a batch of fit instances: one detector grid and one
    Gaussian count per instance.  Parameters and targets change per call.

    grids:  Grid objects (e.g. ``Detector``) with ``shape`` and the pixel map
    counts: Gaussians per instance; instance i owns parameter rows
            [gauss_offsets[i], gauss_offsets[i] + counts[i])
    cutoff: as ``evaluate_gaussian_input_image``: None keeps every
            contribution above exp(-80) of a Gaussian's peak; 4.0 is a 4-sigma cut.
    """

    def __init__(self, grids: Sequence, counts: Sequence[int], cutoff: float | None = None):
        if len(grids) != len(counts) or len(grids) == 0:
            raise ValueError("GaussianFitBatch: need one Gaussian count per grid (and at least one grid)")
        lib = _lib.load()
        self.device = _lib.require_device()
        self.grids = list(grids)
        self.counts = [int(c) for c in counts]
        if any(c < 0 for c in self.counts):
            raise ValueError("GaussianFitBatch: negative Gaussian count")
        self.shapes = [(int(g.shape[0]), int(g.shape[1])) for g in self.grids]
        self.gauss_offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        npix = [h * w for h, w in self.shapes]
        self.pix_offsets = np.concatenate([[0], np.cumsum(npix)]).astype(np.int64)
        self.n_inst = len(self.grids)
        self.n_gauss = int(self.gauss_offsets[-1])
        self.n_pix = int(self.pix_offsets[-1])
        inst = (FitInstance * self.n_inst)()
        for i, (g, (h, w)) in enumerate(zip(self.grids, self.shapes)):
            inst[i].gauss_begin = int(self.gauss_offsets[i])
            inst[i].n_gauss = self.counts[i]
            inst[i].pix_begin = int(self.pix_offsets[i])
            inst[i].height, inst[i].width = h, w
            inst[i].pix2m[:] = [float(v) for v in _pix2m(g)]
        self._inst = inst
        self.plan = FitPlan()
        nsigma = 0.0 if cutoff is None else float(cutoff)
        _lib.check(lib.tg_gauss_fit_plan(inst, self.n_inst, self.n_gauss, nsigma, ctypes.byref(self.plan)),
                   "tg_gauss_fit_plan")
        # +256 so the workspace can start on a 256-byte boundary
        self._ws_raw = torch.empty(int(self.plan.workspace_bytes) + 256, dtype=torch.uint8, device=self.device)
        off = (-self._ws_raw.data_ptr()) % 256
        self.workspace = self._ws_raw[off:off + int(self.plan.workspace_bytes)]
        _lib.check(lib.tg_gauss_fit_upload(ctypes.byref(self.plan), inst, self.workspace.data_ptr(),
                                           _lib.stream_ptr(self.device)), "tg_gauss_fit_upload")
        gi = np.full(self.n_gauss, -1, dtype=np.int64)
        for i, c in enumerate(self.counts):
            gi[self.gauss_offsets[i]:self.gauss_offsets[i] + c] = i
        self.gauss_instance = torch.from_numpy(gi).to(self.device)

    # -- packing helpers -------------------------------------------------
    def pack(self, rays: Sequence[GaussianRay]) -> torch.Tensor:
        """[n_gauss, 16] parameter rows of one GaussianRay bundle per instance."""
        if len(rays) != self.n_inst:
            raise ValueError(f"pack: expected {self.n_inst} bundles, got {len(rays)}")
        rows = []
        for i, r in enumerate(rays):
            p = pack_params(r.to_vector(), self.device) if self.counts[i] else torch.zeros(
                (0, NPARAM), dtype=torch.float64, device=self.device)
            if p.shape[0] != self.counts[i]:
                raise ValueError(f"pack: instance {i} has {p.shape[0]} Gaussians, the batch expects {self.counts[i]}")
            rows.append(p)
        return torch.cat(rows, 0).contiguous()

    def images(self, packed: torch.Tensor) -> list[torch.Tensor]:
        """Views [H, W] of a packed per-pixel array, one per instance."""
        return [packed[self.pix_offsets[i]:self.pix_offsets[i + 1]].view(*self.shapes[i])
                for i in range(self.n_inst)]

    def pack_images(self, images: Sequence[torch.Tensor], dtype=torch.float32) -> torch.Tensor:
        """Concatenate per-instance [H, W] images into the packed pixel layout."""
        if len(images) != self.n_inst:
            raise ValueError("pack_images: one image per instance")
        flat = []
        for i, im in enumerate(images):
            if tuple(im.shape) != self.shapes[i]:
                raise ValueError(f"pack_images: instance {i} image shape {tuple(im.shape)} != {self.shapes[i]}")
            flat.append(im.reshape(-1).to(device=self.device, dtype=dtype))
        return torch.cat(flat).contiguous()

    # -- kernels -----------------------------------------------------------
    def _check(self, t, name, shape, dtype):
        if t is None:
            return
        if not (t.is_cuda and t.dtype == dtype and t.is_contiguous() and tuple(t.shape) == shape):
            raise ValueError(f"{name}: expected a contiguous {dtype} device tensor of shape {shape}, "
                             f"got {t.dtype} {tuple(t.shape)} on {t.device}")

    def run(self, params: torch.Tensor, target: torch.Tensor | None = None, field: torch.Tensor | None = None,
            loss: torch.Tensor | None = None, grad: torch.Tensor | None = None) -> None:
        """One pass into caller-owned outputs (graph-capturable: no allocation).

        params [n_gauss, 16] f64; target [n_pix] f32 intensities; field
        [n_pix] complex64; loss [n_inst] f64; grad [n_gauss, 16] f64."""
        self._check(params, "params", (self.n_gauss, NPARAM), torch.float64)
        self._check(target, "target", (self.n_pix,), torch.float32)
        self._check(field, "field", (self.n_pix,), torch.complex64)
        self._check(loss, "loss", (self.n_inst,), torch.float64)
        self._check(grad, "grad", (self.n_gauss, NPARAM), torch.float64)
        if (loss is not None or grad is not None) and target is None:
            raise ValueError("run: loss and grad need target intensities")
        lib = _lib.load()
        _lib.check(lib.tg_gauss_fit_run(ctypes.byref(self.plan), self.workspace.data_ptr(), params.data_ptr(),
                                        _ptr(target), _ptr(field), _ptr(loss), _ptr(grad),
                                        _lib.stream_ptr(self.device)), "tg_gauss_fit_run")

    def field(self, params: torch.Tensor) -> torch.Tensor:
        """Packed complex64 fields [n_pix] (evaluate_gaussian_input_image per instance)."""
        out = torch.empty(self.n_pix, dtype=torch.complex64, device=self.device)
        self.run(params, field=out)
        return out

    def intensity(self, params: torch.Tensor) -> torch.Tensor:
        """Packed |u|^2 [n_pix] f32, e.g. the target of a fit."""
        return self.field(params).abs().square()

    def loss_and_grad(self, params: torch.Tensor, target: torch.Tensor):
        """(loss [n_inst] f64, d sum(loss) / d params [n_gauss, 16] f64)."""
        loss = torch.empty(self.n_inst, dtype=torch.float64, device=self.device)
        grad = torch.empty((self.n_gauss, NPARAM), dtype=torch.float64, device=self.device)
        self.run(params, target=target, loss=loss, grad=grad)
        return loss, grad


class _FitLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, batch, target):
        loss, grad = batch.loss_and_grad(params.detach().contiguous(), target)
        ctx.save_for_backward(grad)
        ctx.batch = batch
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        gi = ctx.batch.gauss_instance
        w = torch.where(gi >= 0, grad_loss[gi.clamp(min=0)], torch.zeros((), dtype=grad.dtype, device=grad.device))
        return grad * w[:, None], None, None


def fit_loss(params: torch.Tensor, batch: GaussianFitBatch, target: torch.Tensor) -> torch.Tensor:
    """Per-instance losses mean((|u_i|^2 - T_i)^2) [n_inst], differentiable with
    respect to ``params`` through the hand-written adjoint."""
    return _FitLoss.apply(params, batch, target)


def evaluate_gaussian_input_images(rays: Sequence[GaussianRay], grids: Sequence,
                                   cutoff: float | None = None) -> list[torch.Tensor]:
    """``evaluate_gaussian_input_image`` for many (bundle, grid) pairs at once."""
    dev = _lib.require_device()
    rows = [pack_params(r.to_vector(), dev) for r in rays]
    batch = GaussianFitBatch(grids, [p.shape[0] for p in rows], cutoff)
    return [im.clone() for im in batch.images(batch.field(torch.cat(rows, 0).contiguous()))]


# ---------------------------------------------------------------------------
# the synthetic batched-fit workload
# ---------------------------------------------------------------------------
def workload_tables(n_inst: int = 256, seed_counts: int = 5, seed_sides: int = 6, n_min: int = 3,
                    n_max: int = 2000, side_min: int = 128, side_max: int = 512):
    """Gaussian counts round(exp(U(ln n_min, ln n_max))) (seed 5) and square
    detector sides uniform in {side_min..side_max} (seed 6)."""
    counts = np.rint(np.exp(np.random.default_rng(seed_counts).uniform(np.log(n_min), np.log(n_max), n_inst)))
    sides = np.random.default_rng(seed_sides).integers(side_min, side_max + 1, n_inst)
    return counts.astype(np.int64), sides.astype(np.int64)


def workload_params(counts, sides, seed: int = 7, pixel: float = 1e-6, wavelength: float = 2.51e-12,
                    tilt: float = 2e-3, waist=(5e-6, 25e-6)):
    """Seeded parameter rows [sum(counts), 16] (numpy f64): centres uniform in
    the inner 80% of each detector, waists U(waist), theta U(-pi/2, pi/2),
    |amplitude| U(0.2, 1) with phase U(-pi, pi), tilts U(-tilt, tilt), flat
    wavefronts, zero pathlength (the ranges of fit_three_gaussians.ipynb)."""
    rng = np.random.default_rng(seed)
    rows = []
    for n, s in zip(counts, sides):
        n = int(n)
        half = 0.5 * s * pixel
        p = np.zeros((n, NPARAM))
        p[:, 0] = rng.uniform(-0.8 * half, 0.8 * half, n)
        p[:, 1] = rng.uniform(-0.8 * half, 0.8 * half, n)
        p[:, 2:4] = rng.uniform(-tilt, tilt, (n, 2))
        amp = rng.uniform(0.2, 1.0, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
        p[:, 5], p[:, 6] = amp.real, amp.imag
        p[:, 7:9] = rng.uniform(waist[0], waist[1], (n, 2))
        p[:, 9:11] = np.inf
        p[:, 11] = wavelength
        p[:, 12] = rng.uniform(-np.pi / 2, np.pi / 2, n)
        rows.append(p)
    return np.concatenate(rows, 0) if rows else np.zeros((0, NPARAM))


def perturb_params(params, seed: int = 8, pixel: float = 1e-6):
    """Target parameters: centres moved by N(0, 1 pixel), waists and
    |amplitude| scaled by 1 + N(0, 0.05), theta moved by N(0, 0.05)."""
    rng = np.random.default_rng(seed)
    p = np.array(params, dtype=np.float64, copy=True)
    n = p.shape[0]
    p[:, 0:2] += rng.normal(0, pixel, (n, 2))
    p[:, 7:9] *= 1 + rng.normal(0, 0.05, (n, 2))
    p[:, 5:7] *= (1 + rng.normal(0, 0.05, n))[:, None]
    p[:, 12] += rng.normal(0, 0.05, n)
    return p


def workload(n_inst: int = 256, pixel: float = 1e-6, wavelength: float = 2.51e-12, cutoff: float | None = None,
             **table_kw):
    """The batched-fit workload: (batch, params [G, 16] f64 device, target
    intensities [n_pix] f32 device from the perturbed parameters)."""
    from .components import Detector

    counts, sides = workload_tables(n_inst, **table_kw)
    grids = [Detector(z=0.0, pixel_size=(pixel, pixel), shape=(int(s), int(s))) for s in sides]
    batch = GaussianFitBatch(grids, counts, cutoff)
    p = workload_params(counts, sides, pixel=pixel, wavelength=wavelength)
    params = torch.from_numpy(p).to(batch.device)
    tgt = torch.from_numpy(perturb_params(p, pixel=pixel)).to(batch.device)
    target = batch.intensity(tgt)
    return batch, params, target
