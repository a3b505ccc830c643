"""Krivanek wave aberrations of a thin lens (reference aberrations.py:1-108).

The aberration function of the reference is a sum over (order n, fold m)
terms in polar slope coordinates (alpha, phi):

    W = sum C_nm / (n + 1) * alpha^(n+1) * cos(m (phi - phi_nm))

``W_krivanek``/``grad_W_krivanek`` keep the reference's signatures.

``krivanek_poly`` is the form the lens evaluates per ray.  With
z = ax + i ay (so alpha^m e^(i m phi) = z^m) and s = ax^2 + ay^2, every term is
a polynomial in the slopes, because n + 1 - m is even for every Krivanek term:

    alpha^(n+1) cos(m (phi - phi0)) = s^((n+1-m)/2) Re[z^m e^(-i m phi0)].

Grouping the terms by fold m gives W = Re sum_m z^m P_m(s) with complex
polynomials P_m, and

    dW/dax =  Re sum_m [m z^(m-1) P_m(s)] + 2 ax Re sum_m [z^m P_m'(s)]
    dW/day = -Im sum_m [m z^(m-1) P_m(s)] + 2 ay Re sum_m [z^m P_m'(s)].

Per ray this is additions and multiplications only: no hypot, atan2, sin or
cos (the cos/sin of m phi0 are host constants), and no special case at
alpha = 0, where the reference substitutes 1e-30 for alpha and obtains the
same zero gradient.  The trigonometric and polynomial forms agree to
rounding; terms whose coefficient is exactly 0.0 are left out (the reference
multiplies them in, which only differs for non-finite slopes).
"""
from __future__ import annotations

import dataclasses
import math

from . import _program


@dataclasses.dataclass(frozen=True)
class KrivanekCoeffs:
    """Krivanek aberration coefficients (metres) and azimuths (radians)
    (reference aberrations.py:5-31)."""

    C10: float = 0.0
    C12: float = 0.0
    phi12: float = 0.0
    C21: float = 0.0
    phi21: float = 0.0
    C23: float = 0.0
    phi23: float = 0.0
    C30: float = 0.0
    C32: float = 0.0
    phi32: float = 0.0
    C34: float = 0.0
    phi34: float = 0.0
    C41: float = 0.0
    phi41: float = 0.0
    C43: float = 0.0
    phi43: float = 0.0
    C45: float = 0.0
    phi45: float = 0.0
    C50: float = 0.0
    C52: float = 0.0
    phi52: float = 0.0
    C54: float = 0.0
    phi54: float = 0.0
    C56: float = 0.0
    phi56: float = 0.0


# (order n, fold m, coefficient, azimuth) of every term, as the reference's
# brackets B2..B6 list them (aberrations.py:43-48)
TERMS = (
    (1, 0, "C10", None), (1, 2, "C12", "phi12"),
    (2, 1, "C21", "phi21"), (2, 3, "C23", "phi23"),
    (3, 0, "C30", None), (3, 2, "C32", "phi32"), (3, 4, "C34", "phi34"),
    (4, 1, "C41", "phi41"), (4, 3, "C43", "phi43"), (4, 5, "C45", "phi45"),
    (5, 0, "C50", None), (5, 2, "C52", "phi52"), (5, 4, "C54", "phi54"), (5, 6, "C56", "phi56"),
)


def as_coeffs(p) -> KrivanekCoeffs:
    """KrivanekCoeffs from a KrivanekCoeffs or a dict of its fields."""
    if isinstance(p, KrivanekCoeffs):
        return p
    if isinstance(p, dict):
        return KrivanekCoeffs(**p)
    raise TypeError(f"aberration coefficients must be KrivanekCoeffs or a dict, not {type(p).__name__}")


# ---------------------------------------------------------------------------
# reference API (array functions of alpha, phi)
# ---------------------------------------------------------------------------
def _math(v):
    import numpy as np
    import torch

    if isinstance(v, torch.Tensor):
        return torch
    if isinstance(v, np.ndarray):
        return np
    return math


def krivanek_coeff_brackets(phi, p):
    """(B2, B3, B4, B5, B6): the azimuthal brackets of orders 1..5
    (reference aberrations.py:42-48)."""
    p = as_coeffs(p)
    M = _math(phi)
    B = [0.0] * 5
    for n, m, cname, phname in TERMS:
        C = getattr(p, cname)
        B[n - 1] = B[n - 1] + (C if phname is None else C * M.cos(m * (phi - getattr(p, phname))))
    return tuple(B)


def W_krivanek(alpha, phi, p):
    """Wave aberration W(alpha, phi) in metres (reference aberrations.py:51-61)."""
    B2, B3, B4, B5, B6 = krivanek_coeff_brackets(phi, p)
    a2 = alpha * alpha
    a3 = a2 * alpha
    a4 = a2 * a2
    return 0.5 * a2 * B2 + (a3 / 3.0) * B3 + 0.25 * a4 * B4 + 0.2 * a4 * alpha * B5 + (a3 * a3 / 6.0) * B6


def grad_W_krivanek(alpha_x, alpha_y, p):
    """(dW/dalpha_x, dW/dalpha_y) (reference aberrations.py:64-108), in the
    polynomial form of krivanek_poly."""
    _, gx, gy = krivanek_poly(alpha_x, alpha_y, p)
    return gx, gy


# ---------------------------------------------------------------------------
# per-ray polynomial form
# ---------------------------------------------------------------------------
def _is_zero(c) -> bool:
    return isinstance(c, (int, float)) and not isinstance(c, bool) and c == 0.0


def _add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    return a + b


def _mul(a, b):
    if a is None or b is None:
        return None
    return a * b


def _cmul(a, b):
    """(a_re + i a_im)(b_re + i b_im) on pairs whose parts may be None (= 0)."""
    ar, ai = a
    br, bi = b
    re = _add(_mul(ar, br), None if ai is None or bi is None else -(ai * bi))
    im = _add(_mul(ar, bi), _mul(ai, br))
    return re, im


def _poly(cs: dict, s):
    """sum_q cs[q] s^q by Horner's rule (cs: {q: constant}); None if empty."""
    if not cs:
        return None
    top = max(cs)
    acc = cs[top]
    for q in range(top - 1, -1, -1):
        acc = acc * s
        if q in cs:
            acc = acc + cs[q]
    return acc


def krivanek_poly(ax, ay, p):
    """(W, dW/dax, dW/day) at slopes (ax, ay) by the polynomial form.

    Works on floats, tensors, traced per-ray values and dual numbers alike
    (additions and multiplications only).  Missing parts are returned as 0.0."""
    p = as_coeffs(p)
    # complex coefficients c = C/(n+1) e^(-i m phi0), grouped by fold m and power q of s
    groups: dict[int, tuple[dict, dict]] = {}
    for n, m, cname, phname in TERMS:
        C = getattr(p, cname)
        if _is_zero(C):
            continue
        k = C / (n + 1)
        q = (n + 1 - m) // 2
        re, im = groups.setdefault(m, ({}, {}))
        if phname is None:
            re[q] = k
        else:
            ang = m * getattr(p, phname)
            re[q] = k * _program.cos(ang)
            im[q] = -(k * _program.sin(ang))
    if not groups:
        return 0.0, 0.0, 0.0
    s = ax * ax + ay * ay
    # folds in increasing order; z^m is formed from z^(m-1) when first needed,
    # so at most two powers are live at a time (short register lifetimes in
    # the traced per-ray program)
    zprev, zcur, mcur = None, None, 0  # z^(mcur-1), z^mcur (None: z^0 = 1)

    def times(zm, c):  # z^m c, z^0 = None
        return c if zm is None else _cmul(zm, c)

    W = G_re = G_im = H_re = None
    for m, (re, im) in sorted(groups.items()):
        while mcur < m:
            zprev, zcur = zcur, ((ax, ay) if zcur is None else _cmul(zcur, (ax, ay)))
            mcur += 1
        # each partial sum is completed before the next is started
        P = (_poly(re, s), _poly(im, s))
        W = _add(W, times(zcur, P)[0])
        if m > 0:
            g = times(zprev, P)
            if m > 1:
                g = (_mul(g[0], m), _mul(g[1], m))
            G_re = _add(G_re, g[0])
            G_im = _add(G_im, g[1])
        dre = {q - 1: q * v for q, v in re.items() if q > 0}
        dim = {q - 1: q * v for q, v in im.items() if q > 0}
        if dre or dim:
            H_re = _add(H_re, times(zcur, (_poly(dre, s), _poly(dim, s)))[0])
    dWx = _add(G_re, None if H_re is None else 2 * ax * H_re)
    dWy = _add(None if G_im is None else -G_im, None if H_re is None else 2 * ay * H_re)
    z = 0.0
    return (z if W is None else W), (z if dWx is None else dWx), (z if dWy is None else dWy)
