"""4D-STEM shadow-image model on the HIP kernels (forward + descan/geometry fit).

The model is the reference's component chain for one scan position
(reference components.py:218-406, run.py:85-116):

    point source (x = y = 0, slopes on an n x n grid inside the cone)
    -> Scanner(z=0, scan_pos, scan_tilt)          scan tilt = pivot * scan_pos
    -> ScanGrid(z=defocus)                        the sample plane
    -> TiltDescanner(z=defocus, 4x4 error map)    == Descanner for zero tilt columns
    -> Detector(z=defocus + L_c)

``STEMModel.components(c)`` returns exactly that list (scan positions as
[S, 1] tensors, ray slopes as [1, J] tensors) for the generic ``run_to_end``
path; ``stem_forward`` and ``descan_fit_loss`` run the specialised kernels
that evaluate it for all scan positions and camera lengths in one launch.

Parameter vector (float64, 18 + n_cam): descan map A (16, row-major; rows are
the x, y, dx, dy offsets, columns spx, spy, stx, sty), scan step (m), scan
rotation (degrees), camera lengths (m).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from .components import Component, Detector, ScanGrid, Scanner
from .ray import Ray


class StemGeometry(ctypes.Structure):
    _fields_ = [
        ("scan_h", ctypes.c_int32), ("scan_w", ctypes.c_int32),
        ("n_side", ctypes.c_int32), ("n_cam", ctypes.c_int32),
        ("semi_conv", ctypes.c_double), ("defocus", ctypes.c_double), ("pivot", ctypes.c_double),
        ("offsets", ctypes.c_double * 4),
        ("img_h", ctypes.c_int32), ("img_w", ctypes.c_int32),
        ("img_m", ctypes.c_double * 6),
        ("det_h", ctypes.c_int32), ("det_w", ctypes.c_int32),
        ("det_m", ctypes.c_double * 6),
    ]


_P = ctypes.c_void_p
_lib.register_signatures({
    "tg_stem_forward": (ctypes.c_int, [ctypes.POINTER(StemGeometry), _P, _P, _P, _P, _P, _P]),
    "tg_stem_fit_workspace": (ctypes.c_int64, [ctypes.POINTER(StemGeometry)]),
    "tg_stem_fit": (ctypes.c_int, [ctypes.POINTER(StemGeometry), _P, _P, _P, _P, _P]),
    "tg_stem_backproject": (ctypes.c_int, [ctypes.POINTER(StemGeometry), _P, _P, ctypes.c_int32, _P, _P, _P]),
    "tg_stem_fit_terms": (ctypes.c_int, [ctypes.POINTER(StemGeometry), _P, _P, _P, _P]),
    "tg_stem_fit_noalloc_max_bytes": (ctypes.c_int64, []),
    "tg_stem_fit_noalloc": (ctypes.c_int, [ctypes.POINTER(StemGeometry), _P, _P, _P, _P]),
})

MAX_CAM = 16


@dataclasses.dataclass(frozen=True)
class TiltDescanner(Component):
    """Descanner with a full 4x4 error map of (spx, spy, stx, sty).

    x += (A[0].v + off[0] - spx), y += (A[1].v + off[1] - spy),
    dx += (A[2].v + off[2] - stx), dy += (A[3].v + off[3] - sty), v = (spx, spy, stx, sty).
    With the tilt columns of A zero this is the reference Descanner
    (components.py:320-372) with DescanError(pxo_pxi=A[0][0], pxo_pyi=A[0][1], ...).
    """

    z: float
    scan_pos_x: object
    scan_pos_y: object
    scan_tilt_x: object = 0.0
    scan_tilt_y: object = 0.0
    descan_matrix: tuple = ((0.0,) * 4,) * 4
    descan_offsets: tuple = (0.0, 0.0, 0.0, 0.0)

    def __call__(self, ray: Ray):
        A, o = self.descan_matrix, self.descan_offsets
        spx, spy, stx, sty = self.scan_pos_x, self.scan_pos_y, self.scan_tilt_x, self.scan_tilt_y

        def row(r):
            return spx * A[r][0] + spy * A[r][1] + stx * A[r][2] + sty * A[r][3] + o[r]

        return ray.derive(
            x=ray.x + (row(0) - spx) * ray._one,
            y=ray.y + (row(1) - spy) * ray._one,
            dx=ray.dx + (row(2) - stx) * ray._one,
            dy=ray.dy + (row(3) - sty) * ray._one,
        )


@dataclasses.dataclass
class STEMModel:
    """Geometry and parameters of the 4D-STEM shadow-image model."""

    scan_shape: tuple = (64, 64)
    scan_step: float = 1e-10
    scan_rotation: float = 12.0
    n_side: int = 64
    semi_conv: float = 0.5e-3
    defocus: float = 200e-9
    pivot: float = 1e4
    descan: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((4, 4)))
    descan_offsets: tuple = (0.0, 0.0, 0.0, 0.0)
    camera_lengths: tuple = (0.5,)
    det_shape: tuple = (256, 256)
    det_pixel: float = 2.5e-6
    sample_shape: tuple = (512, 512)
    sample_pixel: float = 2e-11

    # -- derived ---------------------------------------------------------------
    @property
    def n_scan(self) -> int:
        return int(self.scan_shape[0] * self.scan_shape[1])

    @property
    def n_rays(self) -> int:
        return int(self.n_side * self.n_side)

    @property
    def n_cam(self) -> int:
        return len(self.camera_lengths)

    def params(self) -> np.ndarray:
        """The float64 parameter vector (descan 16, step, rotation, camera lengths)."""
        return np.concatenate([np.asarray(self.descan, dtype=np.float64).reshape(16),
                               [self.scan_step, self.scan_rotation],
                               np.asarray(self.camera_lengths, dtype=np.float64)])

    def with_params(self, p) -> "STEMModel":
        p = np.asarray(p, dtype=np.float64)
        return dataclasses.replace(self, descan=p[:16].reshape(4, 4).copy(), scan_step=float(p[16]),
                                   scan_rotation=float(p[17]), camera_lengths=tuple(float(v) for v in p[18:]))

    def scan_grid(self) -> ScanGrid:
        return ScanGrid(z=self.defocus, pixel_size=(self.scan_step, self.scan_step), shape=tuple(self.scan_shape),
                        rotation=self.scan_rotation)

    def sample_grid(self) -> Detector:
        return Detector(z=self.defocus, pixel_size=(self.sample_pixel, self.sample_pixel),
                        shape=tuple(self.sample_shape))

    def detector(self, c: int = 0) -> Detector:
        return Detector(z=self.defocus + self.camera_lengths[c], pixel_size=(self.det_pixel, self.det_pixel),
                        shape=tuple(self.det_shape))

    def slopes(self) -> np.ndarray:
        """f32 slopes of one grid axis: (k - (n-1)/2) * 2*semi_conv/n."""
        k = np.arange(self.n_side, dtype=np.float64)
        return ((k - 0.5 * (self.n_side - 1)) * (2.0 * self.semi_conv / self.n_side)).astype(np.float32)

    def scan_positions(self) -> tuple[np.ndarray, np.ndarray]:
        """(spx, spy) f64 for flattened scan index s = row * scan_w + col."""
        yy, xx = np.meshgrid(np.arange(self.scan_shape[0]), np.arange(self.scan_shape[1]), indexing="ij")
        spx, spy = self.scan_grid().pixels_to_metres((yy.ravel(), xx.ravel()))
        return np.asarray(spx, dtype=np.float64), np.asarray(spy, dtype=np.float64)

    def geometry(self) -> StemGeometry:
        """The kernels' geometry struct (a fresh copy per call).

        Building it takes ~50-90 us of Python (two pixel-transform matrices),
        more than a forward launch, so it is cached on the instance and keyed
        by every field it depends on."""
        key = (tuple(int(v) for v in self.scan_shape), int(self.n_side), float(self.semi_conv),
               float(self.defocus), float(self.pivot), tuple(float(v) for v in self.descan_offsets),
               tuple(float(v) for v in self.camera_lengths), tuple(int(v) for v in self.det_shape),
               float(self.det_pixel), tuple(int(v) for v in self.sample_shape), float(self.sample_pixel))
        cached = self.__dict__.get("_geometry_cache")
        if cached is None or cached[0] != key:
            cached = (key, self._build_geometry())
            self.__dict__["_geometry_cache"] = cached
        return StemGeometry.from_buffer_copy(cached[1])

    def _build_geometry(self) -> StemGeometry:
        if not 1 <= self.n_cam <= MAX_CAM:
            raise ValueError(f"1..{MAX_CAM} camera lengths supported, got {self.n_cam}")
        if self.n_side % 2:
            raise ValueError("n_side must be even")
        g = StemGeometry()
        g.scan_h, g.scan_w = int(self.scan_shape[0]), int(self.scan_shape[1])
        g.n_side, g.n_cam = int(self.n_side), int(self.n_cam)
        g.semi_conv, g.defocus, g.pivot = float(self.semi_conv), float(self.defocus), float(self.pivot)
        for k in range(4):
            g.offsets[k] = float(self.descan_offsets[k])
        g.img_h, g.img_w = int(self.sample_shape[0]), int(self.sample_shape[1])
        M = self.sample_grid().metres_to_pixels_mat
        for k, v in enumerate((M[0, 0], M[0, 1], M[0, 2], M[1, 0], M[1, 1], M[1, 2])):
            g.img_m[k] = float(v)
        g.det_h, g.det_w = int(self.det_shape[0]), int(self.det_shape[1])
        M = self.detector(0).metres_to_pixels_mat
        for k, v in enumerate((M[0, 0], M[0, 1], M[0, 2], M[1, 0], M[1, 1], M[1, 2])):
            g.det_m[k] = float(v)
        return g

    # -- the same model as a reference component list -----------------------------
    def rays(self, device=None) -> Ray:
        """Point-source rays: x = y = 0, slopes [1, J] (row-major over (dy, dx))."""
        device = _lib.require_device() if device is None else device
        t = self.slopes()
        dy, dx = np.meshgrid(t, t, indexing="ij")
        return Ray(x=0.0, y=0.0, dx=torch.from_numpy(dx.reshape(1, -1).copy()).to(device),
                   dy=torch.from_numpy(dy.reshape(1, -1).copy()).to(device), z=0.0, pathlength=0.0)

    def components(self, c: int = 0, device=None) -> list:
        device = _lib.require_device() if device is None else device
        spx, spy = self.scan_positions()
        spx_t = torch.from_numpy(spx.reshape(-1, 1)).to(device)
        spy_t = torch.from_numpy(spy.reshape(-1, 1)).to(device)
        stx_t = torch.from_numpy((self.pivot * spx).reshape(-1, 1)).to(device)
        sty_t = torch.from_numpy((self.pivot * spy).reshape(-1, 1)).to(device)
        A = tuple(tuple(float(v) for v in row) for row in np.asarray(self.descan, dtype=np.float64))
        return [
            Scanner(z=0.0, scan_pos_x=spx_t, scan_pos_y=spy_t, scan_tilt_x=stx_t, scan_tilt_y=sty_t),
            self.scan_grid(),
            TiltDescanner(z=self.defocus, scan_pos_x=spx_t, scan_pos_y=spy_t, scan_tilt_x=stx_t,
                          scan_tilt_y=sty_t, descan_matrix=A, descan_offsets=tuple(self.descan_offsets)),
            self.detector(c),
        ]


def _ptr(t):
    return None if t is None else t.data_ptr()


def params_tensor(model: STEMModel, device=None) -> torch.Tensor:
    device = _lib.require_device() if device is None else device
    return torch.from_numpy(model.params()).to(device)


def stem_forward(model: STEMModel, params: torch.Tensor | None = None, sample: torch.Tensor | None = None,
                 coords: bool = True, values: bool = False, image4d: bool = False, out=None):
    """Run the forward model for all (camera length, scan position, ray).

    Returns a dict with `coords` [n_cam, S, J, 2] f32, `values` [S, J] f32 and
    `image4d` [S, det_h, det_w] f32 (camera length 0) as requested.
    """
    device = _lib.require_device()
    lib = _lib.load()
    params = params_tensor(model, device) if params is None else params
    _check_params(model, params)
    S, J, C = model.n_scan, model.n_rays, model.n_cam
    res = {} if out is None else out
    shapes = {"coords": (C, S, J, 2), "values": (S, J), "image4d": (S,) + tuple(model.det_shape)}
    for name, wanted in (("coords", coords), ("values", values), ("image4d", image4d)):
        if not wanted:
            continue
        if name not in res:
            res[name] = torch.empty(shapes[name], dtype=torch.float32, device=device)
        else:
            # the kernel writes through the raw pointer: a reused buffer must
            # have exactly the shape and layout it assumes
            t = res[name]
            if (not isinstance(t, torch.Tensor) or tuple(t.shape) != shapes[name] or t.dtype != torch.float32
                    or not t.is_cuda or not t.is_contiguous()):
                raise ValueError(f"out[{name!r}] must be a contiguous device float32 tensor of shape {shapes[name]}")
    if image4d:
        # the kernel adds every ray's deposit into the image: it starts from
        # zero on every call (a fill kernel, so the call stays graph-capturable)
        res["image4d"].zero_()
    if (values or image4d) and sample is None:
        raise ValueError("a sample image is needed for values / image4d")
    if sample is not None:
        if tuple(sample.shape) != tuple(model.sample_shape) or sample.dtype != torch.float32 or not sample.is_cuda:
            raise ValueError(f"sample must be a device float32 tensor of shape {tuple(model.sample_shape)}")
        sample = sample.contiguous()
    g = model.geometry()
    st = lib.tg_stem_forward(ctypes.byref(g), params.data_ptr(), _ptr(sample), _ptr(res.get("coords") if coords else None),
                             _ptr(res.get("values") if values else None),
                             _ptr(res.get("image4d") if image4d else None), _lib.stream_ptr(device))
    _lib.check(st, "tg_stem_forward")
    return res


def backproject(model: STEMModel, data4d: torch.Tensor, params: torch.Tensor | None = None, cam: int = 0,
                weights: bool = False):
    """Shadow-image back-projection of a 4D-STEM dataset onto the sample plane.

    data4d: [S, det_h, det_w] f32 device tensor (one detector image per scan
    position, camera length `cam`).  Returns (image, weight): f64
    [img_h, img_w] sums over every ray of the detector value it lands on
    times the bilinear weights at its sample-plane position (the adjoint of
    the forward's 4D deposit), and the sums of those weights (or None).
    image / weight is the mean back-projected value per sample pixel; it
    lines up across scan positions only when the model's geometry is right."""
    device = _lib.require_device()
    params = params_tensor(model, device) if params is None else params
    _check_params(model, params)
    shape = (model.n_scan,) + tuple(model.det_shape)
    if tuple(data4d.shape) != shape or data4d.dtype != torch.float32 or not data4d.is_cuda:
        raise ValueError(f"data4d must be a device float32 tensor of shape {shape}")
    data4d = data4d.contiguous()
    image = torch.zeros(tuple(model.sample_shape), dtype=torch.float64, device=device)
    weight = torch.zeros_like(image) if weights else None
    g = model.geometry()
    st = _lib.load().tg_stem_backproject(ctypes.byref(g), params.data_ptr(), data4d.data_ptr(), int(cam),
                                         image.data_ptr(), _ptr(weight), _lib.stream_ptr(device))
    _lib.check(st, "tg_stem_backproject")
    return image, weight


def _check_params(model, params):
    if not (isinstance(params, torch.Tensor) and params.is_cuda and params.dtype == torch.float64
            and params.is_contiguous() and params.numel() == 18 + model.n_cam):
        raise ValueError(f"params must be a contiguous device float64 tensor of {18 + model.n_cam} entries")


class STEMFitWorkspace:
    """Device buffers of one fit configuration (reused across calls and graph replays)."""

    def __init__(self, model: STEMModel, device=None):
        device = _lib.require_device() if device is None else device
        self.model = model
        self.geometry = model.geometry()
        self.np = 18 + model.n_cam + 1
        n = int(_lib.load().tg_stem_fit_workspace(ctypes.byref(self.geometry))) // 8
        self.workspace = torch.empty(n, dtype=torch.float64, device=device)
        # its head [S, np] (row s: scan position s's contributions), exposed
        # transposed: partials[t, s] is scan position s's share of entry t
        self.partials = self.workspace[: self.np * model.n_scan].view(model.n_scan, self.np).t()
        self.out = torch.empty(self.np, dtype=torch.float64, device=device)

    def run(self, params: torch.Tensor, targets: torch.Tensor) -> torch.Tensor:
        """Enqueue the fused forward+adjoint; returns the [grad..., loss] vector."""
        m = self.model
        _check_params(m, params)
        shape = (m.n_cam, m.n_scan, m.n_rays, 2)
        if tuple(targets.shape) != shape or targets.dtype != torch.float32 or not targets.is_cuda \
                or not targets.is_contiguous():
            raise ValueError(f"targets must be a contiguous device float32 tensor of shape {shape}")
        st = _lib.load().tg_stem_fit(ctypes.byref(self.geometry), params.data_ptr(), targets.data_ptr(),
                                     self.workspace.data_ptr(), self.out.data_ptr(),
                                     _lib.stream_ptr(targets.device))
        _lib.check(st, "tg_stem_fit")
        return self.out

    def terms(self, params: torch.Tensor, targets: torch.Tensor) -> torch.Tensor:
        """Enqueue the per-scan-position terms only; returns an [np, S] view
        of the workspace's [S, np] rows: row t, column s is scan position s's
        share of run()'s entry t, and the last row is the loss per scan
        position, a residual map."""
        m = self.model
        _check_params(m, params)
        shape = (m.n_cam, m.n_scan, m.n_rays, 2)
        if tuple(targets.shape) != shape or targets.dtype != torch.float32 or not targets.is_cuda \
                or not targets.is_contiguous():
            raise ValueError(f"targets must be a contiguous device float32 tensor of shape {shape}")
        st = _lib.load().tg_stem_fit_terms(ctypes.byref(self.geometry), params.data_ptr(), targets.data_ptr(),
                                           self.workspace.data_ptr(), _lib.stream_ptr(targets.device))
        _lib.check(st, "tg_stem_fit_terms")
        return self.partials


class _DescanFitLoss(torch.autograd.Function):
    """loss = sum |detector(params) - targets|^2 with the hand-written adjoint.

    The kernel computes loss and gradient in one pass over the targets; the
    backward multiplies the stored gradient by the incoming cotangent."""

    @staticmethod
    def forward(ctx, params, targets, ws: STEMFitWorkspace):
        out = ws.run(params.detach(), targets).clone()
        ctx.save_for_backward(out[:-1])
        return out[-1]

    @staticmethod
    def backward(ctx, grad_loss):
        (g,) = ctx.saved_tensors
        return grad_loss * g, None, None


def descan_fit_loss(model: STEMModel, params: torch.Tensor, targets: torch.Tensor,
                    workspace: STEMFitWorkspace | None = None) -> torch.Tensor:
    """Differentiable L2 descan/geometry-fit loss (scalar tensor)."""
    ws = STEMFitWorkspace(model, params.device) if workspace is None else workspace
    return _DescanFitLoss.apply(params, targets, ws)


def reference_workload(n_cam: int = 1, seed_descan: int = 1, seed_sample: int = 2, **overrides):
    """The seeded 4D-STEM configuration: 64x64 scan (0.1 nm, 12 deg), 64x64 rays
    in a 0.5 mrad cone, defocus 200 nm, descan map ~ N(0, 1e-3), 256^2 detector
    (2.5 um pixels), 512^2 random-phase sample (0.02 nm pixels).  With
    n_cam == 1 the camera length is 0.5 m, else linspace(0.1, 1.0, n_cam)."""
    cams = (0.5,) if n_cam == 1 else tuple(np.linspace(0.1, 1.0, n_cam))
    descan = np.random.default_rng(seed_descan).normal(0.0, 1e-3, (4, 4))
    model = STEMModel(descan=descan, camera_lengths=cams, **overrides)
    sample = np.random.default_rng(seed_sample).uniform(-np.pi, np.pi, model.sample_shape).astype(np.float32)
    return model, sample


def perturbed(model: STEMModel, seed: int = 3, rel: float = 1e-3) -> STEMModel:
    """Every parameter scaled by (1 + rel * N(0, 1)) (target generation)."""
    p = model.params()
    return model.with_params(p * (1.0 + rel * np.random.default_rng(seed).normal(size=p.shape)))


# ---------------------------------------------------------------------------
# calibration loop (the consumer of the fused fit kernel)
# ---------------------------------------------------------------------------
FIT_GROUPS = {
    "descan": slice(0, 16),
    "scan_step": slice(16, 17),
    "scan_rotation": slice(17, 18),
    "camera_lengths": slice(18, None),
}
# typical size of a parameter whose start value is 0 (descan entries are
# dimensionless map coefficients, ~1e-3)
_ZERO_SCALE = {"descan": 1e-3, "scan_step": 1e-10, "scan_rotation": 1.0, "camera_lengths": 0.1}


@dataclasses.dataclass
class CalibrationResult:
    model: STEMModel           # the model with the fitted parameters
    params: np.ndarray         # fitted parameter vector (descan 16, step, rotation, camera lengths)
    loss: float                # final loss (sum of squared detector residuals, m^2)
    initial_loss: float
    evaluations: int           # fused loss + gradient launches
    history: list              # loss at every evaluation


def calibrate(model: STEMModel, targets: torch.Tensor, fit=("descan", "scan_step", "scan_rotation", "camera_lengths"),
              max_iter: int = 200, history_size: int = 30, params0=None) -> CalibrationResult:
    """Fit descan map, scan step, scan rotation and camera lengths of `model`
    to measured detector coordinates (the descan-error calibration over
    camera lengths that the reference describes, README.md:13-16).

    targets: [n_cam, S, J, 2] f32 device tensor of measured detector (x, y).
    fit: which parameter groups move (FIT_GROUPS); the rest stay at the
    model's values (or params0).  Each step evaluates the loss and its full
    gradient in one pass of the fused fit kernel.  The optimiser is L-BFGS
    with a strong-Wolfe line search over parameters normalised by their
    start values (or a typical size where they start at 0), on the loss
    divided by its start value.

    Detector coordinates are linear in the scan positions, so with the
    descan map free the scan step and rotation are not determined by them
    (a change of either is absorbed by the map), and neither are map columns
    that are proportional for this scan (position and tilt columns when the
    tilt is pivot x position): L-BFGS then returns one minimiser of that
    valley.  Fit step and rotation with the map held, or the map with them held."""
    dev = targets.device
    p_start = np.asarray(model.params() if params0 is None else params0, dtype=np.float64)
    idx = []
    scale = np.zeros_like(p_start)
    for name in fit:
        if name not in FIT_GROUPS:
            raise ValueError(f"unknown parameter group {name!r}; choose from {tuple(FIT_GROUPS)}")
        sl = FIT_GROUPS[name]
        ids = list(range(len(p_start)))[sl]
        idx += ids
        for i in ids:
            scale[i] = abs(p_start[i]) if p_start[i] != 0 else _ZERO_SCALE[name]
    idx = sorted(set(idx))
    ws = STEMFitWorkspace(model, dev)
    p0 = torch.from_numpy(p_start).to(dev)
    sel = torch.tensor(idx, dtype=torch.long, device=dev)
    sc = torch.from_numpy(scale[idx]).to(dev)
    delta = torch.zeros(len(idx), dtype=torch.float64, device=dev, requires_grad=True)

    def params_of(d):
        return p0.index_add(0, sel, sc * d)

    with torch.no_grad():
        loss0 = float(ws.run(p0, targets)[-1])
    norm = loss0 if loss0 > 0 else 1.0
    opt = torch.optim.LBFGS([delta], lr=1.0, max_iter=max_iter, history_size=history_size,
                            line_search_fn="strong_wolfe", tolerance_grad=1e-15, tolerance_change=1e-15)
    history = []

    def closure():
        opt.zero_grad()
        loss = descan_fit_loss(model, params_of(delta), targets, ws) / norm
        loss.backward()
        history.append(float(loss.detach()) * norm)
        return loss

    opt.step(closure)
    with torch.no_grad():
        p = params_of(delta).cpu().numpy()
        final = float(ws.run(params_of(delta), targets)[-1])
    return CalibrationResult(model=model.with_params(p), params=p, loss=final, initial_loss=loss0,
                             evaluations=len(history), history=history)
