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

import dataclasses
from typing import Any, Callable, Sequence

import numpy as np
import torch

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


def jacobian_and_value(fn, argnums: int = 0):
    """``f(ray, *args) -> (d fn / d ray, fn(ray, *args))`` for a scalar ray, by
    forward-mode duals (reference run.py:17-23, which uses jax.jacobian).
    Only the ray argument (argnums=0) is differentiated."""
    if argnums != 0:
        raise NotImplementedError("jacobian_and_value differentiates the ray argument (argnums=0) only")
    from .transfer import jacobian

    def inner(ray, *args, **kwargs):
        return jacobian(lambda r: fn(r, *args, **kwargs), has_aux_value=True)(ray)

    return inner


def jacobian_transform(component):
    """Per-step (ray_out, jacobian) for scalar rays, via forward-mode duals."""

    def inner(ray):
        jac, out = jacobian_and_value(component)(ray)
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


def _model_fn(ray, components, propagator, composed: bool):
    """The per-ray function of the model: the elements step by step, or
    (composed=True, affine models) their composed affine-plus-pathlength form."""
    if not composed:
        return lambda r: _trace_model(r, components, propagator)
    from .compose import affine_form, apply_affine

    z = ray.z
    if isinstance(z, torch.Tensor):
        if z.numel() != 1:
            raise ValueError("composed=True needs one input z for all rays")
        z = z.item()
    form = affine_form(components, float(z), propagator)
    return lambda r: apply_affine(form, r)


def run_to_end(ray: Ray, components: Sequence, propagator: BasePropagator = FreeSpaceParaxial(), *,
               composed: bool = False) -> Ray:
    """Propagate `ray` through all components; one kernel launch for tensor rays.

    composed=True evaluates an affine model through its composed
    affine-plus-pathlength form (compose.affine_form): one short program
    whatever the model's length, rounding differently from the stepwise
    evaluation (about 1e-15 relative in float64).

    Tensor rays go through the plan cache: a call with the same model,
    propagator and ray structure (dtypes, shapes, strides, device) as an
    earlier one re-launches that call's compiled programs on the new arrays
    without tracing."""
    if _program.is_device_ray(ray):
        return _program.plan_cache.launch(
            ("run_to_end", tuple(components), propagator, composed, ray),
            lambda: _program.plan_fn(_model_fn(ray, components, propagator, composed), ray))
    return _model_fn(ray, components, propagator, composed)(ray)


def plan_run_to_image(ray: Ray, components: Sequence, acc=None, grid=None,
                      propagator: BasePropagator = FreeSpaceParaxial(), *, composed: bool = False) -> "_program.Plan":
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
        _model_fn(ray, components, propagator, composed), ray,
        bin_request_fn=lambda r: (spec, r.x, r.y),
    )
    plan.acc = acc
    return plan


def run_to_image(ray: Ray, components: Sequence, acc=None, grid=None,
                 propagator: BasePropagator = FreeSpaceParaxial(), *, composed: bool = False):
    """``run_to_end`` + ``grid.into_image`` fused into one launch.

    `grid` defaults to the last component (it must be a Grid, e.g. Detector).
    Returns (ray_out, acc).  Cached like ``run_to_end`` (the key adds the
    grid and the accumulator's dtype and shape).
    """
    from .grid import Grid, _prepare_acc, make_bin_spec

    grid = components[-1] if grid is None else grid
    if not isinstance(grid, Grid):
        raise ValueError("run_to_image needs a Grid (e.g. Detector) to bin into")
    device = _lib.require_device()
    acc = _prepare_acc(grid.shape, acc, device)
    if not _program.is_device_ray(ray):
        ray = ray.to_vector()

    def build():
        spec = make_bin_spec(grid.metres_to_pixels_mat, grid.shape, acc)
        return _program.plan_fn(_model_fn(ray, components, propagator, composed), ray,
                                bin_request_fn=lambda r: (spec, r.x, r.y))

    out = _program.plan_cache.launch(("run_to_image", tuple(components), grid, propagator, composed, ray, acc),
                                     build, acc=acc)
    return out, acc


def plan_run_to_end(ray: Ray, components: Sequence, propagator: BasePropagator = FreeSpaceParaxial(), *,
                    composed: bool = False) -> "_program.Plan":
    """Compile ``run_to_end`` for a vectorised ray once; ``plan.launch()`` runs it."""
    return _program.plan_fn(_model_fn(ray, components, propagator, composed), ray)


def calculate_derivatives(ray: Ray, model: Sequence, order: int):
    """Derivatives of run_to_end with respect to a scalar ray, orders 1 ..
    `order` (reference run.py:119-147, successive jax.jacfwd).  The dual
    numbers here are first order: order 1 gives [jacobian]; higher orders
    raise NotImplementedError."""
    if order > 1:
        raise NotImplementedError("calculate_derivatives: only first derivatives (order 1) are available")
    from .transfer import jacobian

    return [jacobian(lambda r: run_to_end(r, model))(ray)] if order == 1 else []


def solve_model(ray: Ray, model: Sequence, propagator: BasePropagator = FreeSpaceParaxial()):
    """Per-step 5x5 ABCD matrices along the model (reference run.py:150-179)."""
    from .transfer import custom_jacobian_matrix

    mats = []
    for _, jac in run_iter(ray, model, transform=jacobian_transform, propagator=propagator):
        mats.append(custom_jacobian_matrix(jac))
    return np.array(mats)


# ---------------------------------------------------------------------------
# Jacobians with respect to selected parameters (reference run.py:182-267)
# ---------------------------------------------------------------------------
def _grad_targets(input_ray, model, grad_vars):
    """[(path, where, keys)]: `where` is "ray" or the model index; `keys`
    lead from that object to one scalar parameter (or a ray field)."""
    from .tree_utils import PathBuilder, get_at, scalar_leaves

    out = []
    for var in grad_vars:
        found = []
        if var is input_ray:
            for f in dataclasses.fields(input_ray):
                found.append(((input_ray, f.name), "ray", (f.name,)))
        elif isinstance(var, PathBuilder):
            path = var._path()
            root, keys = path[0], tuple(path[1:])
            if root is input_ray:
                where = "ray"
            else:
                where = next((i for i, el in enumerate(model) if el is root), None)
            if where is not None:
                for leaf, _ in scalar_leaves(get_at(root, keys)):
                    found.append((path + leaf, where, keys + leaf))
        else:
            raise TypeError(f"run_with_grads: grad_vars holds {type(var).__name__}; use x.params.<field> paths")
        if not found:
            what = var._path() if isinstance(var, PathBuilder) else var
            raise RuntimeError(f"Cannot find {what} in parameters")
        out += found
    return out


def _seeded(value, is_ray_field: bool):
    if isinstance(value, torch.Tensor) and not is_ray_field:
        if value.numel() != 1:
            raise NotImplementedError("run_with_grads: array-valued model parameters (differentiate scalars)")
        value = value.item()
    return _program.FwdDual(value, 1.0)


def _split(out_ray, like):
    """(value ray, tangent ray) of an output ray whose fields may be FwdDual."""
    vals, tans = {}, {}
    for f in dataclasses.fields(out_ray):
        v = getattr(out_ray, f.name)
        if isinstance(v, _program.FwdDual):
            vals[f.name], tans[f.name] = v.v, v.t
        else:
            vals[f.name], tans[f.name] = v, 0.0
    return _program._rebuild(like, vals), _program._rebuild(like, tans)


def _shape_tangents(value_ray, tan_ray):
    """Tangents shaped like the values: uniform tangents of tensor fields
    become full tensors (the reference returns arrays)."""
    out = {}
    for f in dataclasses.fields(value_ray):
        v, t = getattr(value_ray, f.name), getattr(tan_ray, f.name)
        if isinstance(v, torch.Tensor) and not isinstance(t, torch.Tensor):
            out[f.name] = torch.full_like(v, float(t))
        else:
            out[f.name] = t
    return _program._rebuild(value_ray, out)


def run_with_grads(input_ray: Ray, model: Sequence, grad_vars: Sequence,
                   propagator: BasePropagator = FreeSpaceParaxial()):
    """Run the model and the Jacobians of the output ray with respect to the
    selected variables (reference run.py:182-267).

    grad_vars: paths built with ``.params`` -- ``model[i].params.focal_length``,
    ``input_ray.params.x``, a non-leaf such as ``descanner.params.descan_error``
    (every scalar under it), or the input ray itself (every field).

    Returns (value, grads): the output Ray, and a dict from each variable's
    path (root object, keys...) to a Ray of d(output field)/d(variable).
    Derivatives are forward-mode: for vectorised rays every variable costs one
    fused launch that computes the outputs and their derivatives together.
    A ray-field variable of a vectorised ray is differentiated per ray (the
    diagonal of the reference's [N, N] Jacobian); model parameters must be
    scalars.  Raises RuntimeError for a variable that is not in the model.
    """
    from .tree_utils import replace_at

    model = list(model)
    targets = _grad_targets(input_ray, model, grad_vars)
    value = run_to_end(input_ray, model, propagator)
    grads = {}
    for path, where, keys in targets:
        if where == "ray":
            ray_d = replace_at(input_ray, keys, lambda v: _seeded(v, True))
            model_d = model
        else:
            ray_d = input_ray
            model_d = list(model)
            model_d[where] = replace_at(model[where], keys, lambda v: _seeded(v, False))
        if _program.is_device_ray(input_ray):
            trace = _program.Trace(_lib.require_device())
            out = _trace_model(trace.symbolize(ray_d), model_d, propagator)
            val_ray, tan_ray = _split(out, input_ray)
            plan = _program.build_plan(trace, val_ray, extra_rays=[tan_ray])
            value_i = plan.launch()
            grads[path] = _shape_tangents(value_i, plan.extra_out[0])
        else:
            out = _trace_model(ray_d, model_d, propagator)
            _, tan_ray = _split(out, input_ray)
            grads[path] = tan_ray
    return value, grads
