"""Seeded random ragged batches for the batched Gaussian fit against the oracle.

Each case draws 1-6 instances with random image shapes (1 x 1 up to
~200 x 200, mostly not multiples of the 64-pixel tile), Gaussian counts
(zero included), tilts from near zero to 2e-3 and occasional finite radii,
and targets from perturbed centres and waists.  Field, per-instance loss and
every gradient row are checked with the bars of test_gpu_gauss_fit.py
(field 2e-5 of max |ref| and 1e-4 rad; loss relative 1e-4; gradient entries
1e-4 of the size of the terms they sum, columns 1e-2 of their largest
entry)."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(__file__))
from test_gpu_gauss_fit import PIX, _batch, _field_ok, _grad_ok, _oracle_instance, _params  # noqa: E402

from oracle import beamlet_oracle as BO  # noqa: E402
from oracle import gauss_fit_oracle as GF  # noqa: E402

pytestmark = pytest.mark.gpu


def _random_batch(seed):
    rng = np.random.default_rng(5000 + seed)
    n_inst = int(rng.integers(1, 7))
    shapes = [(int(rng.integers(1, 200)), int(rng.integers(1, 200))) for _ in range(n_inst)]
    if rng.uniform() < 0.3:
        shapes[int(rng.integers(n_inst))] = (1, 1)
    counts = [int(rng.choice([0, 1, int(rng.integers(2, 41))])) for _ in range(n_inst)]
    tilt = float(10.0 ** rng.uniform(-7, np.log10(2e-3)))
    batch, _ = _batch(shapes, counts)
    p = np.concatenate([_params(n, s, rng, tilt=tilt, finite_radius=bool(rng.uniform() < 0.4))
                        for n, s in zip(counts, shapes)]) if sum(counts) else np.zeros((0, 16))
    tgt = p.copy()
    tgt[:, 0:2] += rng.normal(0, PIX, (len(p), 2))
    tgt[:, 7:9] *= 1 + rng.normal(0, 0.05, (len(p), 2))
    target = np.concatenate([
        np.abs(GF.field(tgt[batch.gauss_offsets[i]:batch.gauss_offsets[i + 1]],
                        BO.grid_coords((PIX, PIX), s))) ** 2 for i, s in enumerate(shapes)]).astype(np.float32)
    return batch, p, target, dict(shapes=shapes, counts=counts, tilt=tilt)


@pytest.mark.parametrize("seed", range(8))
def test_random_ragged_batch_vs_oracle(device, seed):
    batch, p, target, info = _random_batch(seed)
    print(f"[random fit batch seed={seed} {info}]", flush=True)
    params = torch.from_numpy(p).to(device)
    tg = torch.from_numpy(target).to(device)
    field = torch.empty(batch.n_pix, dtype=torch.complex64, device=device)
    loss = torch.empty(batch.n_inst, dtype=torch.float64, device=device)
    grad = torch.empty((batch.n_gauss, 16), dtype=torch.float64, device=device)
    batch.run(params, target=tg, field=field, loss=loss, grad=grad)
    torch.cuda.synchronize()
    f, l, g = field.cpu().numpy(), loss.cpu().numpy(), grad.cpu().numpy()
    for i in range(batch.n_inst):
        u, L, G, S = _oracle_instance(batch, p, target, i)
        _field_ok(f[batch.pix_offsets[i]:batch.pix_offsets[i + 1]], u)
        assert l[i] == pytest.approx(L, rel=1e-4, abs=1e-12), i
        _grad_ok(g[batch.gauss_offsets[i]:batch.gauss_offsets[i + 1]], G, S)
