"""Model drivers (reference run.py:47-267).

``run_iter`` / ``run_to_end`` keep the reference's step structure: before
*every* Source/Component a propagator step over ``component.z - ray.z`` is
inserted, even when that distance is 0 (reference run.py:75-80).

On a vectorised ray:

* ``run_to_end`` traces the whole model into one ray program -> one launch;
* ``run_iter`` runs one launch per step, so each yielded ray is materialised;
* ``run_to_image`` additionally bins the final positions into a detector
  image inside the same launch.

On a scalar ray everything runs as plain Python arithmetic.
"""
from __future__ import annotations

from typing import Any, Callable, Sequence

import numpy as np

from . import _lib, _program
from .components import Component
from .propagator import BasePropagator, FreeSpaceParaxial, Propagator
from .ray import Ray
from .source import Source


def passthrough_transform(component):
    def inner(ray):
        out = component(ray)
        return out, out

    return inner


def jacobian_transform(component):
    """Per-step (ray_out, jacobian) for scalar rays, via forward-mode duals."""
    from .transfer import jacobian

    def inner(ray):
        jac, out = jacobian(component, has_aux_value=True)(ray)
        return out, jac

    return inner


def run_iter(ray: Ray, components: Sequence, transform: Callable = passthrough_transform,
             propagator: BasePropagator = FreeSpaceParaxial()):
    """Yield (step, ray_out) for every propagator and component step."""
    for component in components:
        if isinstance(component, (Source, Component)):
            distance = component.z - ray.z
            propagator_d = propagator.with_distance(distance)
            ray, out = transform(propagator_d)(ray)
            yield propagator_d, out
        ray, out = transform(component)(ray)
        yield component, out


def _trace_model(ray, components, propagator):
    """Apply the model to a (symbolic or scalar) ray, step by step."""
    for component in components:
        if isinstance(component, (Source, Component)):
            distance = component.z - ray.z
            ray = propagator.with_distance(distance)(ray)
        ray = component(ray)
    return ray


def run_to_end(ray: Ray, components: Sequence, propagator: BasePropagator = FreeSpaceParaxial()) -> Ray:
    """Propagate `ray` through all components; one kernel launch for tensor rays."""
    if _program.is_device_ray(ray):
        return _program.apply_fn(lambda r: _trace_model(r, components, propagator), ray)
    return _trace_model(ray, components, propagator)


def plan_run_to_image(ray: Ray, components: Sequence, acc=None, grid=None,
                      propagator: BasePropagator = FreeSpaceParaxial()) -> "_program.Plan":
    """Compile ``run_to_image`` once; ``plan.launch()`` re-runs it (the caller
    zeroes ``plan.acc`` between launches if counts should not accumulate)."""
    from .grid import Grid, _prepare_acc, make_bin_spec

    grid = components[-1] if grid is None else grid
    if not isinstance(grid, Grid):
        raise ValueError("run_to_image needs a Grid (e.g. Detector) to bin into")
    device = _lib.require_device()
    acc = _prepare_acc(grid.shape, acc, device)
    spec = make_bin_spec(grid.metres_to_pixels_mat, grid.shape, acc)
    if not _program.is_device_ray(ray):
        ray = ray.to_vector()
    plan = _program.plan_fn(
        lambda r: _trace_model(r, components, propagator), ray,
        bin_request_fn=lambda r: (spec, r.x, r.y),
    )
    plan.acc = acc
    return plan


def run_to_image(ray: Ray, components: Sequence, acc=None, grid=None,
                 propagator: BasePropagator = FreeSpaceParaxial()):
    """``run_to_end`` + ``grid.into_image`` fused into one launch.

    `grid` defaults to the last component (it must be a Grid, e.g. Detector).
    Returns (ray_out, acc).
    """
    plan = plan_run_to_image(ray, components, acc, grid, propagator)
    return plan.launch(), plan.acc


def plan_run_to_end(ray: Ray, components: Sequence,
                    propagator: BasePropagator = FreeSpaceParaxial()) -> "_program.Plan":
    """Compile ``run_to_end`` for a vectorised ray once; ``plan.launch()`` runs it."""
    return _program.plan_fn(lambda r: _trace_model(r, components, propagator), ray)


def solve_model(ray: Ray, model: Sequence, propagator: BasePropagator = FreeSpaceParaxial()):
    """Per-step 5x5 ABCD matrices along the model (reference run.py:150-179)."""
    from .transfer import custom_jacobian_matrix

    mats = []
    for _, jac in run_iter(ray, model, transform=jacobian_transform, propagator=propagator):
        mats.append(custom_jacobian_matrix(jac))
    return np.array(mats)
