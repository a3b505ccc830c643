"""Gaussian beamlets (reference gaussian.py:13-427).

``make_gaussian_image`` and ``evaluate_gaussian_input_image`` keep the
reference signatures; ``batch_size`` is accepted and ignored (the kernels
tile the detector themselves).  Both return a complex64 device tensor of the
grid's shape.  ``cutoff`` (in units of the beamlet's envelope: contributions
below exp(-cutoff^2) of a beamlet's peak amplitude are skipped) defaults to
None, which keeps everything down to exp(-80) -- below f32 resolution, i.e.
the reference's uncut sum.  ``cutoff=4.0`` is the "4 sigma" cut.

The per-beamlet 5x5 ray-transfer matrix comes from ``transfer.jacobian`` of
``run_to_end`` over the model (reference gaussian.py:234-238 takes
``jax.jacobian``), evaluated once per distinct input-plane z -- exact for
models built from the affine elements.  With a Biprism the matrix depends
on each beamlet's central ray; it is then taken per beamlet by forward-mode
tangents on the GPU (``_abcd_traced``).

Both images are differentiable: when a GaussianRay field (a torch tensor with
requires_grad) or an explicit ``abcd`` tensor requires a gradient, the field
is produced by an autograd function whose backward is the hand-written
adjoint ``tg_beamlet_backward`` (the reference differentiates these images
with jax.grad).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from .ray import Ray

NPARAM = 16


class BeamletCoef(ctypes.Structure):
    _fields_ = [
        ("phi", ctypes.c_double * 12),
        ("k_re", ctypes.c_float), ("k_im", ctypes.c_float),
        ("cx", ctypes.c_float), ("cy", ctypes.c_float),
        ("alpha", ctypes.c_float), ("beta", ctypes.c_float), ("gamma", ctypes.c_float),
        ("cut", ctypes.c_float),
        ("box_w", ctypes.c_float), ("box_h", ctypes.c_float),
    ]


_P = ctypes.c_void_p
_lib.register_signatures({
    "tg_beamlet_coefficients": (ctypes.c_int, [_P, ctypes.c_int64, _P, ctypes.c_int64, ctypes.c_int32, _P,
                                               ctypes.c_double, _P, _P]),
    "tg_beamlet_sum_workspace": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "tg_beamlet_sum": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _P, _P, ctypes.c_int64,
                                      _P]),
    "tg_beamlet_backward_workspace": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "tg_beamlet_backward": (ctypes.c_int, [_P, ctypes.c_int64, _P, ctypes.c_int64, ctypes.c_int32, _P, ctypes.c_double,
                                           ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P, ctypes.c_int64, _P]),
})

MODE_PROPAGATED = 0
MODE_INPUT = 1


# ---- closed forms (reference gaussian.py:13-32, :92-110) -----------------------

def w_z(w0, z, z_r):
    return w0 * np.sqrt(1 + (z / z_r) ** 2)


def zR(w0, wavelength):
    return (np.pi * w0**2) / wavelength


def R(z, z_r):
    if abs(z) < 1e-10:
        return np.inf
    return z * (1 + (z_r / z) ** 2)


def gaussian_beam(x, y, q_inv, k, offset_x=0, offset_y=0):
    return np.exp(1j * k * ((x + offset_x) ** 2 + (y + offset_y) ** 2) / 2 * q_inv)


def Qinv_ABCD(Qinv, A, B, C, D):
    """solve(A + B Qinv, C + D Qinv) (reference gaussian.py:92-96)."""
    return np.linalg.solve(A + B @ Qinv, C + D @ Qinv)


def q_inv(z, w0, wl):
    z_r = zR(w0, wl)
    if abs(z) < 1e-10:
        return 1j * wl / (np.pi * w0**2)
    return -1.0 / R(z, z_r) + 1j * wl / (np.pi * w_z(w0, z, z_r) ** 2)


def decompose_Q_inv(Q_inv, wavelength, eps=1e-12):
    """Beam parameters (waist_x, waist_y, r_x, r_y, theta) of 2x2 complex
    Q_inv matrices [..., 2, 2] (reference gaussian.py:35-89).

    The principal axes are the eigenvectors of the symmetrised imaginary part
    (a proper rotation), the larger waist first; on each axis
    Im(1/q) = wavelength / (pi w^2) gives the waist (inf where |Im| <= eps)
    and the radius is 1 / Re(1/q) (inf where |Re| <= eps).  As in the
    reference, that radius is the negative of GaussianRay's radii_of_curv
    convention (1/q = -1/R + i ...): rebuilding from it keeps |field| and
    mirrors the curvature.  theta is the first axis' angle, taken in
    (-pi/2, pi/2] (an eigenvector's sign is arbitrary; the ellipse is the
    same).  numpy arrays or tensors in, numpy arrays out."""
    Q = np.asarray(_host(Q_inv), dtype=np.complex128)
    S = Q.imag
    S = 0.5 * (S + np.swapaxes(S, -1, -2))
    _, V = np.linalg.eigh(S)
    V = V.copy()

    def proper(V):
        flip = np.where(np.linalg.det(V) < 0, -1.0, 1.0)
        V[..., :, 1] *= flip[..., None]
        return V

    V = proper(V)
    Qd = np.swapaxes(V, -1, -2) @ Q @ V
    qd = np.stack([Qd[..., 0, 0], Qd[..., 1, 1]], axis=-1)

    def waists_of(im):
        with np.errstate(divide="ignore"):
            return np.sqrt(np.where(np.abs(im) > eps, np.abs(wavelength / (np.pi * im)), np.inf))

    swap = waists_of(qd.imag)[..., 0] < waists_of(qd.imag)[..., 1]
    qd = np.where(swap[..., None], qd[..., ::-1], qd)
    V = proper(np.where(swap[..., None, None], V[..., :, ::-1], V))
    waists = waists_of(qd.imag)
    with np.errstate(divide="ignore"):
        radii = np.where(np.abs(qd.real) > eps, 1.0 / qd.real, np.inf)
    theta = np.arctan2(V[..., 1, 0], V[..., 0, 0])
    theta = np.where(theta > np.pi / 2, theta - np.pi, np.where(theta <= -np.pi / 2, theta + np.pi, theta))
    return waists[..., 0], waists[..., 1], radii[..., 0], radii[..., 1], theta


# ---- small algebra helpers (reference gaussian.py:180-222, :340-369) ------------------

def _einsum(spec, *ops):
    if any(isinstance(o, torch.Tensor) for o in ops):
        return torch.einsum(spec, *[torch.as_tensor(o) for o in ops])
    return np.einsum(spec, *ops)


def matrix_vector_mul(M, v):
    """M @ v for one 2x2 M and 2-vector v (reference gaussian.py:180-187)."""
    return _einsum("ij,j->i", M, v)


def matrix_matrix_mul(M1, M2):
    """M1 @ M2 for 2x2 matrices (reference gaussian.py:190-197)."""
    return _einsum("ij,jk->ik", M1, M2)


def matrix_quadratic_mul(v, M):
    """v^T M v (reference gaussian.py:200-207)."""
    return _einsum("i,ij,j->", v, M, v)


def matrix_linear_mul(v, M, w):
    """v^T M w_n for the rows w_n of w [np, 2]: shape [np] (reference
    gaussian.py:210-218)."""
    return _einsum("i,ij,nj->n", v, M, w)


def matrix_matrix_matrix_mul(M1, M2, M3):
    """M1_n M2_n M3_n^T over a batch [n, 2, 2] (reference gaussian.py:221-222)."""
    return _einsum("nij,njk,npk->nip", M1, M2, M3)


def _leading(xs):
    if isinstance(xs, dict):
        return len(next(iter(xs.values())))
    if isinstance(xs, (tuple, list)):
        return len(xs[0])
    return len(xs)


def _take(xs, i):
    if isinstance(xs, dict):
        return {k: v[i] for k, v in xs.items()}
    if isinstance(xs, (tuple, list)):
        return type(xs)(v[i] for v in xs)
    return xs[i]


def map_reduce(f, reducer, init, xs, *, batch_size: int | None = None):
    """The fold reducer(... reducer(reducer(init, f(x_0)), f(x_1)) ..., f(x_n-1))
    over the leading axis of xs (an array/tensor, or a tuple, list or dict of
    them with one leading length) (reference gaussian.py:340-369, lax.scan;
    its batch_size only groups f's evaluations, the fold order is the same,
    so it is accepted and ignored here)."""
    acc = init
    for i in range(_leading(xs)):
        acc = reducer(acc, f(_take(xs, i)))
    return acc


# ---- GaussianRay -------------------------------------------------------------------

@dataclasses.dataclass(frozen=True, kw_only=True)
class GaussianRay(Ray):
    """A Ray carrying a Gaussian beamlet (reference gaussian.py:113-177)."""

    amplitude: object
    waist_xy: object
    radii_of_curv: object
    wavelength: object
    theta: object

    def derive(self, **updates):
        """Copy with fields replaced; values may be callables of self."""
        vals = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        for k, v in updates.items():
            vals[k] = v(self) if callable(v) else v
        return GaussianRay(**vals)

    def to_ray(self) -> Ray:
        return Ray(x=self.x, y=self.y, dx=self.dx, dy=self.dy, z=self.z, pathlength=self.pathlength, _one=self._one)

    @property
    def q_inv(self):
        """(1/q_x, 1/q_y) per principal axis (numpy complex128)."""
        w = np.asarray(_host(self.waist_xy), dtype=np.float64).reshape(-1, 2)
        r = np.asarray(_host(self.radii_of_curv), dtype=np.float64).reshape(-1, 2)
        wl = np.asarray(_host(self.wavelength), dtype=np.float64)
        im = 1j * wl[..., None] / (np.pi * w**2) if wl.ndim else 1j * wl / (np.pi * w**2)
        with np.errstate(divide="ignore"):
            q = np.where(np.isinf(r), im, -1.0 / r + im)
        return q[:, 0], q[:, 1]

    @property
    def Q_inv(self):
        """R(theta) diag(1/qx, 1/qy) R(theta)^T, shape (n, 2, 2)."""
        qx, qy = self.q_inv
        th = np.broadcast_to(np.asarray(_host(self.theta), dtype=np.float64), qx.shape)
        c, s = np.cos(th), np.sin(th)
        Rm = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)
        D = np.zeros(qx.shape + (2, 2), dtype=np.complex128)
        D[..., 0, 0] = qx
        D[..., 1, 1] = qy
        return Rm @ D @ np.swapaxes(Rm, -1, -2)


def _host(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return v


# ---- device plumbing -----------------------------------------------------------------

def _col(v, n, device, dtype=torch.float64):
    t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
    t = t.to(device=device)
    if t.is_complex():
        raise TypeError("complex value in a real beamlet parameter")
    return t.to(dtype).reshape(-1).expand(n) if t.numel() == 1 else t.to(dtype).reshape(n)


def pack_params(rays: GaussianRay, device) -> torch.Tensor:
    """[n, 16] float64 parameter rows of the beamlets (see include/temgym_hip.h)."""
    x = rays.x if isinstance(rays.x, torch.Tensor) else torch.as_tensor(np.asarray(rays.x))
    n = int(x.numel())
    amp = rays.amplitude
    amp = amp if isinstance(amp, torch.Tensor) else torch.as_tensor(np.asarray(amp))
    amp = amp.to(device).reshape(-1)
    if amp.numel() == 1:
        amp = amp.expand(n)
    amp_re = amp.real.to(torch.float64) if amp.is_complex() else amp.to(torch.float64)
    amp_im = amp.imag.to(torch.float64) if amp.is_complex() else torch.zeros_like(amp_re)
    w = rays.waist_xy if isinstance(rays.waist_xy, torch.Tensor) else torch.as_tensor(np.asarray(rays.waist_xy))
    r = rays.radii_of_curv if isinstance(rays.radii_of_curv, torch.Tensor) else torch.as_tensor(
        np.asarray(rays.radii_of_curv))
    w = w.to(device=device, dtype=torch.float64).reshape(-1, 2).expand(n, 2)
    r = r.to(device=device, dtype=torch.float64).reshape(-1, 2).expand(n, 2)
    cols = [
        _col(rays.x, n, device), _col(rays.y, n, device), _col(rays.dx, n, device), _col(rays.dy, n, device),
        _col(rays.pathlength, n, device), amp_re, amp_im, w[:, 0], w[:, 1], r[:, 0], r[:, 1],
        _col(rays.wavelength, n, device), _col(rays.theta, n, device),
    ]
    zeros = torch.zeros(n, dtype=torch.float64, device=device)
    cols += [zeros] * (NPARAM - len(cols))
    return torch.stack(cols, dim=1).contiguous()


def _pix2m(grid) -> np.ndarray:
    T = grid.pixels_to_metres_mat
    return np.array([T[0, 0], T[0, 1], T[0, 2], T[1, 0], T[1, 1], T[1, 2]], dtype=np.float64)


_JAC_FIELDS = ("x", "y", "dx", "dy", "_one")


def _abcd_traced(rays: GaussianRay, model, device, n: int) -> torch.Tensor:
    """Per-beamlet 5x5 Jacobians of run_to_end at every beamlet's central ray
    (reference gaussian.py:234-238: vmap(jacobian(run_to_end))), for models
    whose matrix depends on the ray (Biprism).  Forward-mode tangents: one
    fused launch per input field gives one column of every beamlet's matrix."""
    from .run import run_with_grads

    def col(v):
        t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, dtype=np.float64))
        t = t.detach().to(device=device, dtype=torch.float64).reshape(-1)
        return t.expand(n).contiguous() if t.numel() == 1 else t.reshape(n).contiguous()

    r = rays.to_ray()
    cr = Ray(x=col(r.x), y=col(r.y), dx=col(r.dx), dy=col(r.dy), z=col(r.z), pathlength=col(r.pathlength),
             _one=torch.ones(n, dtype=torch.float64, device=device))
    _, grads = run_with_grads(cr, list(model), [getattr(cr.params, f) for f in _JAC_FIELDS])
    J = torch.empty((n, 5, 5), dtype=torch.float64, device=device)
    for j, fin in enumerate(_JAC_FIELDS):
        g = grads[(cr, fin)]
        for i, fout in enumerate(_JAC_FIELDS):
            v = getattr(g, fout)
            J[:, i, j] = v.reshape(-1) if isinstance(v, torch.Tensor) and v.numel() == n else float(v)
    return J


def _abcd_per_beamlet(rays: GaussianRay, model, device, n: int):
    """(abcd tensor [m, 5, 5] f64, stride) of the model for each beamlet's z
    (or for each beamlet, when the model has ray-dependent elements)."""
    from .components import Deflector, Descanner, Detector, Lens, Plane, Rotator, ScanGrid, Scanner
    from .source import Source
    from .transfer import ray_matrix

    # elements whose map is affine in (x, y, dx, dy, 1): one matrix serves every
    # beamlet; any other element (Biprism, AberratedLensKrivanek, user
    # components) gets each beamlet's own Jacobian
    affine = (Deflector, Descanner, Detector, Lens, Plane, Rotator, ScanGrid, Scanner)
    if not all(type(c) in affine or isinstance(c, Source) for c in model):
        if n == 0:
            return torch.zeros((1, 5, 5), dtype=torch.float64, device=device), 0
        return _abcd_traced(rays, model, device, n), 25
    z = rays.z if isinstance(rays.z, torch.Tensor) else torch.as_tensor(np.asarray(rays.z, dtype=np.float64))
    z = z.reshape(-1).to(torch.float64)
    if n == 0:
        return torch.zeros((1, 5, 5), dtype=torch.float64, device=device), 0
    uniq, inv = torch.unique(z.cpu(), return_inverse=True)
    mats = np.stack([ray_matrix(model, Ray(0.0, 0.0, 0.0, 0.0, float(zv), 0.0)) for zv in uniq.tolist()])
    if len(mats) == 1:
        return torch.from_numpy(mats).to(device), 0
    idx = inv.reshape(-1).expand(n) if inv.numel() == 1 else inv.reshape(-1)
    return torch.from_numpy(mats[idx.numpy()]).to(device).contiguous(), 25


def _check_beamlet_inputs(params: torch.Tensor, mode: int, abcd: torch.Tensor | None, abcd_stride: int) -> None:
    """The kernels read params [n, 16] and, in propagated mode, n (stride 25)
    or one (stride 0) 5x5 matrices through raw pointers: refuse anything else."""
    if not (isinstance(params, torch.Tensor) and params.is_cuda and params.dtype == torch.float64
            and params.is_contiguous() and params.dim() == 2 and params.shape[1] == NPARAM):
        raise ValueError(f"beamlet params must be a contiguous device float64 tensor of shape [n, {NPARAM}]")
    if mode != MODE_PROPAGATED:
        return
    n = int(params.shape[0])
    need = 25 if abcd_stride == 0 else 25 * n
    if abcd_stride not in (0, 25):
        raise ValueError(f"abcd_stride must be 0 (one shared matrix) or 25 (one per beamlet), got {abcd_stride}")
    if not (isinstance(abcd, torch.Tensor) and abcd.is_cuda and abcd.dtype == torch.float64
            and abcd.is_contiguous() and abcd.device == params.device and abcd.numel() >= need):
        raise ValueError(f"propagated beamlets need a contiguous device float64 abcd of at least {need} entries "
                         f"on the params device")


def beamlet_coefficients(params: torch.Tensor, grid, mode: int, abcd: torch.Tensor | None = None,
                         abcd_stride: int = 0, cutoff: float | None = None) -> torch.Tensor:
    """Run the f64 coefficient kernel; returns a uint8 tensor of n packed records."""
    _check_beamlet_inputs(params, mode, abcd, abcd_stride)
    lib = _lib.load()
    n = int(params.shape[0])
    out = torch.empty(n * ctypes.sizeof(BeamletCoef), dtype=torch.uint8, device=params.device)
    m = _pix2m(grid)
    st = lib.tg_beamlet_coefficients(params.data_ptr(), n, abcd.data_ptr() if abcd is not None else None,
                                     abcd_stride, mode, m.ctypes.data, float(cutoff or 0.0), out.data_ptr(),
                                     _lib.stream_ptr(params.device))
    _lib.check(st, "tg_beamlet_coefficients")
    return out


class BeamletSum:
    """Reusable buffers of one (n, detector shape) summation."""

    def __init__(self, n: int, shape, device):
        lib = _lib.load()
        self.n = int(n)
        self.shape = (int(shape[0]), int(shape[1]))
        self.ws_bytes = int(lib.tg_beamlet_sum_workspace(self.n, *self.shape))
        self.workspace = torch.empty(max(self.ws_bytes, 16), dtype=torch.uint8, device=device)
        self.field = torch.empty(self.shape + (2,), dtype=torch.float32, device=device)

    def run(self, coef: torch.Tensor) -> torch.Tensor:
        need = int(_lib.load().tg_beamlet_sum_workspace(self.n, *self.shape))
        if need > self.workspace.numel():  # the slice count was (re)tuned
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.field.device)
        self.ws_bytes = max(need, 16)
        st = _lib.load().tg_beamlet_sum(coef.data_ptr(), self.n, self.shape[0], self.shape[1],
                                         self.field.data_ptr(), self.workspace.data_ptr(), self.ws_bytes,
                                         _lib.stream_ptr(self.field.device))
        _lib.check(st, "tg_beamlet_sum")
        return torch.view_as_complex(self.field)


def beamlet_field(params: torch.Tensor, grid, mode: int, abcd: torch.Tensor | None = None, abcd_stride: int = 0,
                  cutoff: float | None = None) -> torch.Tensor:
    """complex64 field [H, W] of the beamlets params [n, 16] on `grid`."""
    params = params.detach().contiguous()
    abcd = None if abcd is None else abcd.detach().contiguous()
    coef = beamlet_coefficients(params, grid, mode, abcd, abcd_stride, cutoff)
    return BeamletSum(params.shape[0], grid.shape, params.device).run(coef).clone()


def beamlet_backward(params: torch.Tensor, grid, mode: int, grad_field: torch.Tensor,
                     abcd: torch.Tensor | None = None, abcd_stride: int = 0, cutoff: float | None = None,
                     want_abcd: bool = False):
    """Adjoint of ``beamlet_field``: (dL/dparams [n, 16] f64, dL/dT [n, 25] f64
    or None) for the field cotangent grad_field = dL/dRe + i dL/dIm."""
    lib = _lib.load()
    params = params.detach().contiguous()
    abcd = None if abcd is None else abcd.detach().contiguous()
    _check_beamlet_inputs(params, mode, abcd, abcd_stride)
    n = int(params.shape[0])
    H, W = int(grid.shape[0]), int(grid.shape[1])
    g = grad_field.detach().to(device=params.device, dtype=torch.complex64).resolve_conj().contiguous()
    if tuple(g.shape) != (H, W):
        raise ValueError(f"beamlet_backward: grad_field shape {tuple(g.shape)} != grid shape {(H, W)}")
    grad = torch.empty((n, NPARAM), dtype=torch.float64, device=params.device)
    gabcd = None
    if want_abcd and mode == MODE_PROPAGATED:
        gabcd = torch.empty((n, 25), dtype=torch.float64, device=params.device)
    ws_bytes = int(lib.tg_beamlet_backward_workspace(n, H, W))
    raw = torch.empty(ws_bytes + 256, dtype=torch.uint8, device=params.device)
    off = (-raw.data_ptr()) % 256
    m = _pix2m(grid)
    st = lib.tg_beamlet_backward(params.data_ptr(), n, abcd.data_ptr() if abcd is not None else None, abcd_stride,
                                 mode, m.ctypes.data, float(cutoff or 0.0), H, W,
                                 torch.view_as_real(g).data_ptr(), grad.data_ptr(),
                                 gabcd.data_ptr() if gabcd is not None else None, raw.data_ptr() + off, ws_bytes,
                                 _lib.stream_ptr(params.device))
    _lib.check(st, "tg_beamlet_backward")
    return grad, gabcd


class _BeamletField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, abcd, grid, mode, abcd_stride, cutoff):
        ctx.save_for_backward(params, abcd)
        ctx.grid, ctx.mode, ctx.abcd_stride, ctx.cutoff = grid, mode, abcd_stride, cutoff
        return beamlet_field(params, grid, mode, abcd, abcd_stride, cutoff)

    @staticmethod
    def backward(ctx, grad_field):
        params, abcd = ctx.saved_tensors
        want_abcd = abcd is not None and ctx.needs_input_grad[1]
        gp, ga = beamlet_backward(params, ctx.grid, ctx.mode, grad_field, abcd, ctx.abcd_stride, ctx.cutoff,
                                  want_abcd=want_abcd)
        g_abcd = None
        if want_abcd:
            ga = ga.view(-1, 5, 5)
            g_abcd = ga.sum(0, keepdim=True) if ctx.abcd_stride == 0 else ga
            g_abcd = g_abcd.reshape(abcd.shape).to(abcd.dtype)
        return gp.to(params.dtype), g_abcd, None, None, None, None


def _field(params, grid, mode, abcd, stride, cutoff):
    needs = torch.is_grad_enabled() and (params.requires_grad or (abcd is not None and abcd.requires_grad))
    if needs:
        return _BeamletField.apply(params, abcd, grid, mode, stride, cutoff)
    return beamlet_field(params, grid, mode, abcd, stride, cutoff)


def _user_abcd(abcd, n, device):
    a = abcd if isinstance(abcd, torch.Tensor) else torch.as_tensor(np.asarray(abcd, dtype=np.float64))
    a = a.to(device=device, dtype=torch.float64)
    if a.shape == (5, 5) or a.shape == (1, 5, 5):
        return a.reshape(1, 5, 5), 0
    if a.shape == (n, 5, 5):
        return a, 25
    raise ValueError(f"abcd: expected [5, 5] or [n, 5, 5], got {tuple(a.shape)}")


def make_gaussian_image(gaussian_rays: GaussianRay, model, batch_size=128, cutoff: float | None = None, *,
                        abcd=None):
    """Complex field of all beamlets after `model` on its last element (a Grid).

    ``abcd`` ([5, 5] or [n, 5, 5]) replaces the model's ray-transfer matrix,
    e.g. a tensor built from differentiable lens parameters."""
    from .grid import Grid

    rays = gaussian_rays
    assert isinstance(rays, GaussianRay)
    grid = model[-1]
    assert isinstance(grid, Grid)
    device = _lib.require_device()
    params = pack_params(rays.to_vector(), device)
    n = int(params.shape[0])
    if abcd is None:
        abcd, stride = _abcd_per_beamlet(rays, model, device, n)
    else:
        abcd, stride = _user_abcd(abcd, n, device)
    return _field(params, grid, MODE_PROPAGATED, abcd, stride, cutoff)


def evaluate_gaussian_input_image(gaussian_rays: GaussianRay, grid, batch_size=128, cutoff: float | None = None):
    """Complex input-plane field of all beamlets sampled on `grid`."""
    rays = gaussian_rays
    assert isinstance(rays, GaussianRay)
    device = _lib.require_device()
    params = pack_params(rays.to_vector(), device)
    return _field(params, grid, MODE_INPUT, None, 0, cutoff)
