"""Seeded random cases for the beamlet adjoint (tg_beamlet_backward, propagated
mode) against finite differences of the complex128 oracle field.

Each case draws 1-12 beamlets, a ray-transfer matrix composed of a drift, a
thin lens and a drift of random lengths and focal length (some with offset
columns, some varied per beamlet), a detector of random pixel size and
shape, and a random cotangent with a part correlated with the field.  Bars
as in test_gpu_beamlet_grad.py: every parameter and matrix entry within
1e-4 of the size of the terms it sums, columns within 1e-2 of their largest
entry."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(__file__))
from test_gpu_beamlet_grad import _check, _cotangent, _rays  # noqa: E402

from oracle import beamlet_grad_oracle as BG  # noqa: E402
from oracle import beamlet_oracle as BO  # noqa: E402
from temgymcore_amd.components import Detector  # noqa: E402
from temgymcore_amd.gaussian import MODE_PROPAGATED, beamlet_backward, beamlet_field  # noqa: E402

pytestmark = pytest.mark.gpu


def _drift(d):
    M = np.eye(5)
    M[0, 2] = M[1, 3] = d
    return M


def _random_case(seed):
    rng = np.random.default_rng(7000 + seed)
    n = int(rng.integers(1, 13))
    p = _rays(n, rng, w=float(rng.uniform(3e-8, 8e-8)), spread=float(rng.uniform(0.0, 3e-7)),
              tilt=float(rng.uniform(0.0, 1.5e-5)))
    d1, d2 = (float(v) for v in rng.uniform(2e-4, 1e-3, 2))
    T = _drift(d2)
    if rng.uniform() < 0.6:
        lens = np.eye(5)
        lens[2, 0] = lens[3, 1] = -1.0 / float(rng.uniform(0.5e-3, 3e-3))
        T = T @ lens
    T = T @ _drift(d1)
    if rng.uniform() < 0.4:
        T[0:4, 4] = rng.normal(0.0, 1.0, 4) * np.array([3e-8, 3e-8, 2e-5, 2e-5])
    abcd = np.repeat(T[None], n, 0)
    if rng.uniform() < 0.5 and n > 1:
        abcd[1:, 0, 2] *= rng.uniform(0.9, 1.1, n - 1)
        abcd[1:, 1, 3] = abcd[1:, 0, 2]
    pix = float(rng.uniform(2.5e-8, 4.5e-8))
    shape = (int(rng.integers(16, 100)), int(rng.integers(16, 100)))
    return p, abcd, pix, shape, dict(n=n, lens=bool(T[2, 0] != 0), offsets=bool(np.any(T[0:4, 4])), shape=shape)


@pytest.mark.parametrize("seed", range(8))
def test_random_propagated_backward_vs_oracle(device, seed):
    p, abcd, pix, shape, info = _random_case(seed)
    print(f"[random beamlet adjoint seed={seed} {info}]", flush=True)
    rng = np.random.default_rng(7100 + seed)
    grid = Detector(z=1e-3, pixel_size=(pix, pix), shape=shape)
    r = BO.grid_coords((pix, pix), shape)
    params = torch.from_numpy(p).to(device)
    A = torch.from_numpy(abcd).to(device).contiguous()
    u = beamlet_field(params, grid, MODE_PROPAGATED, A, 25).cpu().numpy()
    g = _cotangent(rng, shape, u)
    gp, ga = beamlet_backward(params, grid, MODE_PROPAGATED, torch.from_numpy(g).to(device), A, 25, want_abcd=True)
    ref, refT, S, ST = BG.grad_fd(p, abcd, BG.PROPAGATED, r, g.reshape(-1).astype(np.complex128))
    _check(gp.cpu().numpy()[:, :13], ref[:, :13], S[:, :13])
    assert np.all(gp.cpu().numpy()[:, 13:] == 0)
    _check(ga.cpu().numpy()[:, :20], refT[:, :20], ST[:, :20])
    assert np.all(ga.cpu().numpy()[:, 20:] == 0)
