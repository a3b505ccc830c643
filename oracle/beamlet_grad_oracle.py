"""Derivatives of single-beamlet fields by finite differences (numpy complex128).

Test infrastructure only (see oracle/__init__.py): product code never imports it.

The field of one beamlet is the reference's _beam_field (gaussian.py:276-316,
restated in beamlet_oracle.beam_field) for the propagated image, or
_input_beam_field (gaussian.py:402-407) for the input plane.  For a real loss
L with field cotangent g = dL/dRe(u) + i dL/dIm(u),

    dL/dtheta = Re sum_pixels conj(g) du_b/dtheta,

and grad_fd() evaluates du_b/dtheta by Richardson-extrapolated central
differences, for every parameter column and (propagated) ray-transfer entry,
together with S = sum_pixels |conj(g) du_b/dtheta|, the size of the terms
each entry sums.  tests/test_oracle_beamlet_grad.py pins it against the
finite-difference-pinned analytic gradient of gauss_fit_oracle, the
free-space translation identity du/dx0 = -du/dr_x and the amplitude's
closed form.
"""
import numpy as np

from oracle import beamlet_oracle as B

NPARAM = 16
PROPAGATED, INPUT = 0, 1


def one_field(prow, T, mode, r):
    """Field [P] of one beamlet (parameter row prow [16]) at points r [P, 2]."""
    p = np.asarray(prow, dtype=np.float64)
    amp = p[5] + 1j * p[6]
    lam = p[11]
    k = 2 * np.pi / lam
    Q1 = B.ray_Q_inv(p[None, 7:9], p[None, 9:11], lam, np.array([p[12]]))[0]
    r1m, t1m = p[0:2].copy(), p[2:4].copy()
    if mode == INPUT:
        return B.input_field(amp, k * p[4], Q1, r1m, t1m, k, r)
    T = np.asarray(T, dtype=np.float64).reshape(5, 5)
    A, Bm, C, D = T[0:2, 0:2], T[0:2, 2:4], T[2:4, 0:2], T[2:4, 2:4]
    e, f = T[0:2, 4], T[2:4, 4]
    Q2 = B.Qinv_ABCD(Q1, A, Bm, C, D)
    return B.beam_field(amp, k * p[4], Q1, Q2, r1m, t1m, A, Bm, e, f, k, r)


def derivative(fun, x0, target_rel=1e-4, max_rel_step=1e-3):
    """d fun / dx at x0 (fun: float -> complex array), central differences with
    one Richardson step; the step makes fun change by ~target_rel of max|fun|,
    but stays below max_rel_step |x0| (weakly coupled parameters)."""
    u0 = fun(x0)
    scale = np.abs(u0).max()
    if scale == 0:
        return np.zeros_like(u0)
    h = abs(x0) * 1e-6 if x0 != 0 else 1e-9
    for _ in range(4):
        d = np.abs(fun(x0 + h) - u0).max()
        if d == 0:
            h *= 1e3
            continue
        h *= target_rel * scale / d
    if x0 != 0:
        h = min(h, max_rel_step * abs(x0))

    def cd(hh):
        return (fun(x0 + hh) - fun(x0 - hh)) / (2 * hh)

    return (4 * cd(h / 2) - cd(h)) / 3


def grad_fd(params, abcd, mode, r, g):
    """(grad [n, 16], grad_T [n, 25], S [n, 16], S_T [n, 25]) for the
    cotangent g [P] (complex) at points r [P, 2].  abcd: [n, 5, 5] (propagated)."""
    p = np.asarray(params, dtype=np.float64).reshape(-1, NPARAM)
    n = p.shape[0]
    grad, S = np.zeros((n, NPARAM)), np.zeros((n, NPARAM))
    gT, ST = np.zeros((n, 25)), np.zeros((n, 25))
    w = np.conj(np.asarray(g, dtype=np.complex128))
    for b in range(n):
        T = None if mode == INPUT else np.asarray(abcd[b], dtype=np.float64).reshape(25)
        for c in range(13):
            if c in (9, 10) and np.isinf(p[b, c]):
                continue

            def fp(v, c=c):
                q = p[b].copy()
                q[c] = v
                return one_field(q, T, mode, r)

            du = derivative(fp, p[b, c])
            grad[b, c] = np.real(np.sum(w * du))
            S[b, c] = np.sum(np.abs(w * du))
        if mode == INPUT:
            continue
        for e in range(20):

            def ft(v, e=e):
                TT = T.copy()
                TT[e] = v
                return one_field(p[b], TT, mode, r)

            du = derivative(ft, T[e])
            gT[b, e] = np.real(np.sum(w * du))
            ST[b, e] = np.sum(np.abs(w * du))
    return grad, gT, S, ST
