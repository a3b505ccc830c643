"""Host compiler: component ``__call__`` bodies -> per-ray register programs.

A vectorised ray bundle is never processed by a chain of array operations.
Instead, the Python body of each component (``Lens.__call__``,
``FreeSpaceParaxial.propagate``, user-defined components ...) is executed
once on a *symbolic* ray whose tensor-valued fields are ``Sym`` objects.  Each
arithmetic operation on a ``Sym`` appends one instruction to a trace; the
trace is then dead-code eliminated, register-allocated and run by
``tg_ray_program_run`` (one kernel launch for the whole component list).

Scalar semantics follow the reference's JAX-with-x64 behaviour exactly
(reference run.py:47-116 executes the same Python bodies on jnp arrays):

* arithmetic between Python scalars stays in Python (double precision);
* a Python scalar meeting a per-ray value is rounded once to the ray dtype
  (weak typing) and the operation is carried out in that dtype;
* ``x ** 2`` is ``x * x`` (``lax.integer_pow``);
* every operation rounds once (no fused multiply-add).

Data-dependent Python control flow on a ``Sym`` raises ``TypeError``.
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import math
import numbers
import threading
import types
import weakref
from typing import Any, Callable

import numpy as np
import torch

from . import _lib

_OPC = _lib.OPCODES

# Fields of a Ray that are traced; other dataclass fields (e.g. the beam
# parameters of a GaussianRay) pass through untouched.
RAY_FIELDS = ("x", "y", "dx", "dy", "z", "pathlength", "_one")


def is_number(v) -> bool:
    return isinstance(v, numbers.Number) and not isinstance(v, complex)


def _is_array(v) -> bool:
    return isinstance(v, (torch.Tensor, np.ndarray))


def is_device_ray(ray) -> bool:
    """True if any traced field of `ray` holds a (non-scalar) array."""
    for name in RAY_FIELDS:
        v = getattr(ray, name, None)
        if _is_array(v) and v.ndim > 0:
            return True
    return False


def to_device_tensor(v, device) -> torch.Tensor:
    """Move an array-like onto `device`; integer/bool/half data become f64/f32."""
    if isinstance(v, np.ndarray):
        v = torch.from_numpy(np.ascontiguousarray(v))
    elif not isinstance(v, torch.Tensor):
        v = torch.as_tensor(v)
    if not v.is_floating_point():
        # jnp promotes an integer array meeting a weak float to the default
        # float type, which is f64 under jax_enable_x64.
        v = v.to(torch.float64)
    elif v.dtype not in (torch.float32, torch.float64):
        v = v.to(torch.float32)
    if v.device != device:
        v = v.to(device)
    return v


class Sym:
    """A per-ray value that lives in a register of the program being traced."""

    __slots__ = ("tr", "vid")
    __array_priority__ = 1000  # make numpy scalars defer to Sym operators

    def __init__(self, tr: "Trace", vid: int):
        self.tr = tr
        self.vid = vid

    # arithmetic (a FwdDual operand takes over: it carries a tangent) -------
    def __add__(self, o):
        return NotImplemented if isinstance(o, FwdDual) else self.tr.binary("add", self, o)

    def __radd__(self, o):
        return self.tr.binary("add", o, self)

    def __sub__(self, o):
        return NotImplemented if isinstance(o, FwdDual) else self.tr.binary("sub", self, o)

    def __rsub__(self, o):
        return self.tr.binary("sub", o, self)

    def __mul__(self, o):
        return NotImplemented if isinstance(o, FwdDual) else self.tr.binary("mul", self, o)

    def __rmul__(self, o):
        return self.tr.binary("mul", o, self)

    def __truediv__(self, o):
        return NotImplemented if isinstance(o, FwdDual) else self.tr.binary("div", self, o)

    def __rtruediv__(self, o):
        return self.tr.binary("div", o, self)

    def __neg__(self):
        return self.tr.unary("neg", self)

    def __pos__(self):
        return self

    def __pow__(self, p):
        if not is_number(p) or float(p) != int(p):
            raise NotImplementedError("ray programs support only integer powers of per-ray values")
        p = int(p)
        if p == 0:
            return 1.0
        if p < 0:
            return 1.0 / (self ** (-p))
        # lax.integer_pow: repeated squaring, x**2 == x*x exactly
        acc = None
        base = self
        while p:
            if p & 1:
                acc = base if acc is None else acc * base
            p >>= 1
            if p:
                base = base * base
        return acc

    def sign(self):
        return self.tr.unary("sign", self)

    # control flow on per-ray values is not expressible ---------------------
    def __bool__(self):
        raise TypeError("per-ray value used in Python control flow; ray programs are straight-line")

    def __float__(self):
        raise TypeError("per-ray value converted to a Python float inside a component")

    __index__ = __int__ = __float__

    def __repr__(self):
        return f"Sym(v{self.vid})"


def _zero(t) -> bool:
    """A structural zero tangent (emits no instructions)."""
    return is_number(t) and not isinstance(t, bool) and t == 0


def _tadd(a, b):
    if _zero(a):
        return b
    if _zero(b):
        return a
    return a + b


def _tmul(a, b):
    if _zero(a) or _zero(b):
        return 0.0
    return a * b


def _tneg(a):
    return 0.0 if _zero(a) else -a


class FwdDual:
    """(value, tangent) pair for forward-mode derivatives through component
    bodies.  Each part is a Python number or a per-ray ``Sym``; a tangent of
    0 is a structural zero and emits no instructions.  A parameter seeded as
    FwdDual(p, 1.0) yields d(output)/dp in the tangents (the role of
    jax.jacobian in reference run.py:182-267)."""

    __slots__ = ("v", "t")

    def __init__(self, v, t=0.0):
        self.v = v
        self.t = t

    @staticmethod
    def lift(o) -> "FwdDual":
        return o if isinstance(o, FwdDual) else FwdDual(o, 0.0)

    def __add__(self, o):
        o = FwdDual.lift(o)
        return FwdDual(self.v + o.v, _tadd(self.t, o.t))

    __radd__ = __add__

    def __sub__(self, o):
        o = FwdDual.lift(o)
        return FwdDual(self.v - o.v, _tadd(self.t, _tneg(o.t)))

    def __rsub__(self, o):
        return FwdDual.lift(o) - self

    def __mul__(self, o):
        o = FwdDual.lift(o)
        return FwdDual(self.v * o.v, _tadd(_tmul(self.t, o.v), _tmul(self.v, o.t)))

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = FwdDual.lift(o)
        q = self.v / o.v
        # d(a/b) = da/b - (a/b) db/b
        ta = 0.0 if _zero(self.t) else self.t / o.v
        tb = 0.0 if _zero(o.t) else (q * o.t) / o.v
        return FwdDual(q, _tadd(ta, _tneg(tb)))

    def __rtruediv__(self, o):
        return FwdDual.lift(o) / self

    def __neg__(self):
        return FwdDual(-self.v, _tneg(self.t))

    def __pos__(self):
        return self

    def __pow__(self, p):
        if not is_number(p) or float(p) != int(p):
            raise NotImplementedError("only integer powers of differentiated values")
        p = int(p)
        if p == 0:
            return FwdDual(1.0, 0.0)
        if p < 0:
            return 1.0 / (self ** (-p))
        acc, base = None, self
        while p:
            if p & 1:
                acc = base if acc is None else acc * base
            p >>= 1
            if p:
                base = base * base
        return acc

    # a differentiated value must not leave the algebra silently
    def __float__(self):
        raise TypeError("a differentiated parameter was converted to a Python float; use _program.cos/sin/sign")

    def __bool__(self):
        raise TypeError("a differentiated parameter was used in Python control flow")

    __index__ = __int__ = __float__


def cos(v):
    """cos of a host scalar parameter (float or FwdDual)."""
    if isinstance(v, FwdDual):
        if isinstance(v.v, Sym):
            raise TypeError("cos of a per-ray value is not a ray-program operation")
        return FwdDual(math.cos(v.v), _tmul(v.t, -math.sin(v.v)))
    return math.cos(v)


def sin(v):
    """sin of a host scalar parameter (float or FwdDual)."""
    if isinstance(v, FwdDual):
        if isinstance(v.v, Sym):
            raise TypeError("sin of a per-ray value is not a ray-program operation")
        return FwdDual(math.sin(v.v), _tmul(v.t, math.cos(v.v)))
    return math.sin(v)


def sign(v):
    """``jnp.sign`` for numbers and per-ray values."""
    if isinstance(v, FwdDual):
        return FwdDual(sign(v.v), 0.0)
    if isinstance(v, Sym):
        return v.sign()
    if _is_array(v):
        return torch.sign(v) if isinstance(v, torch.Tensor) else np.sign(v)
    if not is_number(v) and hasattr(v, "sign"):  # transfer.Dual (5x5 Jacobians)
        return v.sign()
    if v != v or v == 0:
        return v
    return 1.0 if v > 0 else -1.0


_PY = {
    "add": lambda a, b: a + b,
    "sub": lambda a, b: a - b,
    "mul": lambda a, b: a * b,
    "div": lambda a, b: a / b,
}


@dataclasses.dataclass
class _Instr:
    op: str
    dst: int
    a: int
    b: int
    imm: float


class Trace:
    """Records the instructions produced by symbolic execution."""

    def __init__(self, device: torch.device):
        self.device = device
        self.instrs: list[_Instr] = []
        self.n_values = 0
        self.slot_tensors: list[torch.Tensor] = []
        # the object each slot was made from (None for intermediate buffers
        # of split programs): what a cached plan re-binds on its next call
        self.slot_source: list = []
        self.slot_value: list[int] = []
        self._slot_by_id: dict[int, int] = {}
        self.deps: list[frozenset] = []   # value id -> slot indices it depends on
        self.vdtype: list[torch.dtype] = []  # value id -> dtype (per-operation promotion)
        self.value_slot: dict[int, int] = {}

    # value construction ----------------------------------------------------
    def _new_value(self, deps: frozenset, dtype: torch.dtype) -> Sym:
        vid = self.n_values
        self.n_values += 1
        self.deps.append(deps)
        self.vdtype.append(dtype)
        return Sym(self, vid)

    def input(self, arr) -> Sym:
        key = id(arr)
        if key in self._slot_by_id:
            s = self._slot_by_id[key]
            return Sym(self, self.slot_value[s])
        t = to_device_tensor(arr, self.device)
        s = len(self.slot_tensors)
        sym = self._new_value(frozenset((s,)), t.dtype)
        self.slot_tensors.append(t)
        self.slot_source.append(arr)
        self.slot_value.append(sym.vid)
        self._slot_by_id[key] = s
        self.value_slot[sym.vid] = s
        return sym

    def lift(self, v):
        """Sym stays Sym; arrays become input slots; scalars stay Python numbers."""
        if isinstance(v, Sym):
            if v.tr is not self:
                raise ValueError("per-ray values from two different traces were combined")
            return v
        if is_number(v):
            return v
        if _is_array(v):
            if v.ndim == 0:
                return v.item()
            return self.input(v)
        raise TypeError(f"unsupported operand of type {type(v).__name__} in a ray program")

    # instruction emission --------------------------------------------------
    def binary(self, op: str, a, b):
        a = self.lift(a)
        b = self.lift(b)
        sa, sb = isinstance(a, Sym), isinstance(b, Sym)
        if not sa and not sb:
            return _PY[op](a, b)
        if sa and sb:
            dt = torch.promote_types(self.vdtype[a.vid], self.vdtype[b.vid])
            out = self._new_value(self.deps[a.vid] | self.deps[b.vid], dt)
            self.instrs.append(_Instr(op, out.vid, a.vid, b.vid, 0.0))
            return out
        if sa:  # per-ray OP scalar
            out = self._new_value(self.deps[a.vid], self.vdtype[a.vid])
            code = {"add": "addi", "sub": "subi", "mul": "muli", "div": "divi"}[op]
            self.instrs.append(_Instr(code, out.vid, a.vid, a.vid, b))
            return out
        out = self._new_value(self.deps[b.vid], self.vdtype[b.vid])  # scalar OP per-ray
        code = {"add": "addi", "sub": "rsubi", "mul": "muli", "div": "rdivi"}[op]
        self.instrs.append(_Instr(code, out.vid, b.vid, b.vid, a))
        return out

    def unary(self, op: str, a: Sym) -> Sym:
        out = self._new_value(self.deps[a.vid], self.vdtype[a.vid])
        self.instrs.append(_Instr(op, out.vid, a.vid, a.vid, 0.0))
        return out

    # ray <-> symbols ---------------------------------------------------------
    def symbolize(self, ray):
        vals = {}
        for f in dataclasses.fields(ray):
            v = getattr(ray, f.name)
            if f.name not in RAY_FIELDS:
                vals[f.name] = v
            elif isinstance(v, FwdDual):
                vals[f.name] = FwdDual(self.lift(v.v), v.t if is_number(v.t) else self.lift(v.t))
            else:
                vals[f.name] = self.lift(v)
        return _rebuild(ray, vals)


def _rebuild(ray, vals: dict):
    """Construct an object of type(ray) from a full field dict."""
    init = {}
    late = {}
    for f in dataclasses.fields(ray):
        if f.init:
            init[f.name] = vals[f.name]
        else:
            late[f.name] = vals[f.name]
    out = type(ray)(**init)
    for k, v in late.items():
        object.__setattr__(out, k, v)
    return out


# --------------------------------------------------------------------------
# compilation
# --------------------------------------------------------------------------

@dataclasses.dataclass
class Compiled:
    """One kernel launch: a register program over one broadcast shape."""

    shape: tuple
    dtype: torch.dtype
    instrs: list            # (opcode, dst_reg, a_reg, b_reg, imm)
    slots: list             # (slot index, reg)
    outputs: list           # (value id, reg)
    keep_regs: dict         # value id -> reg at program end (kept-live values)
    bin_regs: tuple | None  # (x_reg or -1, y_reg or -1)


def _shape_of(trace: Trace, vid: int) -> tuple:
    shapes = [tuple(trace.slot_tensors[s].shape) for s in sorted(trace.deps[vid])]
    return tuple(torch.broadcast_shapes(*shapes)) if shapes else ()


def compile_group(trace: Trace, out_vids: list[int], keep_vids: list[int], shape: tuple) -> Compiled:
    """DCE + linear-scan register allocation for the values `out_vids` (stored)
    and `keep_vids` (kept live to the end, e.g. binning positions)."""
    live_out = set(out_vids) | set(keep_vids)
    kept, needed = _live_instrs(trace, live_out)
    slot_vids = [v for v in trace.slot_value if v in needed]
    if len(slot_vids) > _lib.TG_MAX_SLOT or len(set(out_vids)) > _lib.TG_MAX_SLOT:
        raise _TooWide(len(slot_vids), len(set(out_vids)))
    dtypes = {trace.vdtype[v] for v in needed}
    dtype = torch.float64 if torch.float64 in dtypes else torch.float32
    mixed = len(dtypes) > 1

    last_use: dict[int, int] = {}
    for i, ins in enumerate(kept):
        last_use[ins.a] = i
        last_use[ins.b] = i
    end = len(kept)
    for v in live_out:
        last_use[v] = end

    free = list(range(_lib.TG_MAX_REG - 1, -1, -1))
    reg: dict[int, int] = {}

    def alloc(v):
        if not free:
            raise _TooManyRegs(f"ray program needs more than {_lib.TG_MAX_REG} live per-ray values")
        reg[v] = free.pop()

    for v in slot_vids:
        alloc(v)
    slot_reg = {v: reg[v] for v in slot_vids}
    out_instrs = []
    for i, ins in enumerate(kept):
        ra, rb = reg[ins.a], reg[ins.b]
        # operands whose last use is this instruction release their register
        # before the destination is allocated (the kernel reads a and b first)
        for v in {ins.a, ins.b}:
            if last_use.get(v, -1) == i and v not in live_out and v in reg:
                free.append(reg.pop(v))
        if ins.dst in reg:  # cannot happen in SSA form
            raise AssertionError("value defined twice")
        alloc(ins.dst)
        code = _OPC[ins.op]
        if mixed and trace.vdtype[ins.dst] == torch.float32:
            code |= _lib.TG_OP_F32
        out_instrs.append((code, reg[ins.dst], ra, rb, float(ins.imm)))
        if last_use.get(ins.dst, -1) < i + 1 and ins.dst not in live_out:
            free.append(reg.pop(ins.dst))
    if len(out_instrs) > _lib.TG_MAX_INSTR:
        raise _TooLong(len(out_instrs))
    slots = [(trace.value_slot[v], slot_reg[v]) for v in slot_vids]
    outputs = [(v, reg[v]) for v in out_vids]
    keep_regs = {v: reg[v] for v in keep_vids}
    return Compiled(shape=shape, dtype=dtype, instrs=out_instrs, slots=slots, outputs=outputs,
                    keep_regs=keep_regs, bin_regs=None)


class _TooLong(Exception):
    pass


class _TooWide(Exception):
    """More input slots or outputs than one program holds."""


class _TooManyRegs(NotImplementedError):
    pass


def _live_instrs(trace: Trace, live_out: set):
    """Instructions needed for `live_out` (dead-code elimination), in order,
    and every value they touch.  Values bound to slots (inputs, or results
    materialised by an earlier program) are not recomputed."""
    needed = set(live_out)
    kept = []
    for ins in reversed(trace.instrs):
        if ins.dst in needed and ins.dst not in trace.value_slot:
            kept.append(ins)
            needed.add(ins.a)
            needed.add(ins.b)
    kept.reverse()
    return kept, needed


def _compile_segments(trace: Trace, out_vids: list, keep_vids: list, shape: tuple):
    """Compile one shape group, splitting it into several programs when it
    exceeds one program's instruction, register, slot or output limits.

    The group is cut after a prefix of its instructions.  The prefix program
    stores the final outputs it has computed and every value the rest still
    reads; those buffers become input slots of the rest, which is compiled
    the same way.  A cut is admissible when the rest's inputs (buffers and
    the original inputs it still reads) fit in one program's slots; the
    latest admissible cut whose prefix compiles is taken.  Returns
    ([(Program, None, keepalive)] of the prefixes, Compiled of the last
    segment, [(Compiled, outputs)] of the prefixes)."""
    limits = (_TooLong, _TooWide, _TooManyRegs)
    prefixes, prefix_comps = [], []
    out_vids = list(dict.fromkeys(out_vids))
    while len(out_vids) > _lib.TG_MAX_SLOT:
        # more outputs than one program stores: the first ones get programs
        # of their own, whose results later programs read instead of recomputing
        head, out_vids = out_vids[:_lib.TG_MAX_SLOT], out_vids[_lib.TG_MAX_SLOT:]
        p, comp, pc = _compile_segments(trace, head, [], shape)
        prog, outs, keepalive = prepare(trace, comp, None)
        prefixes += p + [(prog, None, keepalive + list(outs.values()))]
        prefix_comps += pc + [(comp, outs)]
        _bind_slots(trace, outs)
    while True:
        try:
            return prefixes, compile_group(trace, out_vids, keep_vids, shape), prefix_comps
        except limits:
            pass
        live_out = set(out_vids) | set(keep_vids)
        kept, _ = _live_instrs(trace, live_out)
        chosen = None
        for c in range(min(len(kept) - 1, _lib.TG_MAX_INSTR), 0, -1):
            defined = {ins.dst for ins in kept[:c]}
            reads_after = {v for ins in kept[c:] for v in (ins.a, ins.b)} | set(keep_vids)
            cross = defined & reads_after
            stores = sorted(cross | (defined & set(out_vids)))
            if not stores or len(stores) > _lib.TG_MAX_SLOT:
                continue
            still = reads_after | (live_out - defined)
            slot_in = {v for v in still if v in trace.value_slot}
            if len(slot_in | cross) > _lib.TG_MAX_SLOT:
                continue
            try:
                chosen = (stores, compile_group(trace, stores, [], shape))
            except limits:
                continue
            break
        if chosen is None:
            raise NotImplementedError("ray program cannot be split into programs that fit the kernel's limits")
        stores, comp = chosen
        prog, outs, keepalive = prepare(trace, comp, None)
        prefixes.append((prog, None, keepalive + list(outs.values())))
        prefix_comps.append((comp, outs))
        _bind_slots(trace, outs)
        # final outputs stored by the prefix need no later program
        out_vids = [v for v in out_vids if v not in outs]


def _bind_slots(trace: Trace, outs: dict):
    """Later programs read these values from their buffers."""
    for v, t in outs.items():
        s = len(trace.slot_tensors)
        trace.slot_tensors.append(t)
        trace.slot_source.append(None)
        trace.slot_value.append(v)
        trace.value_slot[v] = s


def _pad_shape(shape: tuple) -> list:
    if len(shape) > _lib.TG_MAX_DIM:
        lead = int(np.prod(shape[: len(shape) - _lib.TG_MAX_DIM + 1]))
        shape = (lead,) + tuple(shape[len(shape) - _lib.TG_MAX_DIM + 1:])
    return [1] * (_lib.TG_MAX_DIM - len(shape)) + list(shape)


def _slot_strides(t: torch.Tensor, shape: tuple) -> tuple[list, bool]:
    """Element strides of `t` broadcast to `shape` (padded to 4 dims)."""
    padded = _pad_shape(shape)
    nd = len(shape)
    if nd > _lib.TG_MAX_DIM:
        t = t.expand(shape).contiguous().reshape(padded)
        return list(t.stride()), True
    tshape = [1] * (nd - t.ndim) + list(t.shape)
    tstride = [0] * (nd - t.ndim) + list(t.stride())
    st = [0 if tshape[k] == 1 and shape[k] != 1 else tstride[k] for k in range(nd)]
    st = [0] * (_lib.TG_MAX_DIM - nd) + st
    contig = t.is_contiguous() and tuple(t.shape) == tuple(shape)
    return st, contig


def prepare(trace: Trace, comp: Compiled, bin_spec: _lib.BinSpec | None = None):
    """Allocate outputs and fill the C program struct for one compiled group.

    Returns (Program, outputs: value id -> tensor, keep-alive tensors)."""
    shape = tuple(comp.shape)
    n = int(np.prod(shape)) if shape else 1
    prog = _lib.Program()
    prog.dtype = _lib.TG_F64 if comp.dtype == torch.float64 else _lib.TG_F32
    prog.n = n
    for k, s in enumerate(_pad_shape(shape)):
        prog.shape[k] = s
    prog.n_instr = len(comp.instrs)
    for i, (op, d, a, b, imm) in enumerate(comp.instrs):
        ins = prog.instr[i]
        ins.op, ins.dst, ins.a, ins.b = op, d, a, b
        ins.imm = imm
        ins.immf = float(np.float32(imm))
    keep = []
    prog.n_slot = len(comp.slots)
    for k, (s, r) in enumerate(comp.slots):
        t = trace.slot_tensors[s]
        keep.append(t)
        st, contig = _slot_strides(t, shape)
        sl = prog.slot[k]
        sl.ptr = t.data_ptr()
        for j in range(_lib.TG_MAX_DIM):
            sl.stride[j] = st[j]
        sl.reg = r
        sl.contig = 1 if contig else 0
        sl.dtype = _lib.TG_F64 if t.dtype == torch.float64 else _lib.TG_F32
    outs = {}
    prog.n_out = len(comp.outputs)
    for k, (v, r) in enumerate(comp.outputs):
        odt = trace.vdtype[v]
        o = torch.empty(shape, dtype=odt, device=trace.device)
        outs[v] = o
        prog.out[k] = o.data_ptr()
        prog.out_reg[k] = r
        prog.out_dtype[k] = _lib.TG_F64 if odt == torch.float64 else _lib.TG_F32
    if bin_spec is not None and comp.bin_regs is not None:
        bin_spec.x_reg, bin_spec.y_reg = comp.bin_regs
    # what the struct's pointers were bound to (for PlanCache templates)
    prog.bound_slots = [s for s, _ in comp.slots]
    prog.bound_outputs = [(v, shape, trace.vdtype[v]) for v, _ in comp.outputs]
    return prog, outs, keep


class Plan:
    """A traced and compiled model application, ready to launch repeatedly.

    Launching re-runs the same kernels on the same input and output buffers
    (inputs may be refilled in place between launches).  ``launch()`` costs one
    C call per program, so a plan is also what a HIP graph capture records.
    """

    def __init__(self, trace: Trace, programs: list, out_ray, acc=None):
        self.trace = trace
        self.programs = programs  # list of (Program, BinSpec | None, keepalive)
        self.out_ray = out_ray
        self.acc = acc
        self.extra_out: list = []

    def launch(self, stream: int | None = None):
        lib = _lib.load()
        st = _lib.stream_ptr(self.trace.device) if stream is None else stream
        for prog, spec, _ in self.programs:
            status = lib.tg_ray_program_run(ctypes.byref(prog), ctypes.byref(spec) if spec is not None else None, st)
            _lib.check(status, "tg_ray_program_run")
        return self.out_ray


# --------------------------------------------------------------------------
# user-facing drivers
# --------------------------------------------------------------------------

def build_plan(trace: Trace, sym_ray, bin_request=None, extra_rays=()) -> Plan:
    """Compile the trace into launchable programs and the output ray.

    bin_request: (BinSpec, x, y) with x, y the symbolic final positions, or None.
    extra_rays: more symbolic rays of the same trace computed by the same
    launches (e.g. derivative rays); rebuilt into ``plan.extra_out``.
    """
    rays = [sym_ray] + list(extra_rays)
    fields = {(i, f.name): getattr(r, f.name) for i, r in enumerate(rays) for f in dataclasses.fields(r)}
    stored: dict = {}
    for name, v in fields.items():
        if isinstance(v, Sym):
            if v.tr is not trace:
                raise ValueError("output holds a per-ray value of a different trace")
            stored[name] = v.vid
    # values that are unchanged inputs are returned as the input arrays
    direct = {name: vid for name, vid in stored.items() if vid in trace.value_slot}
    computed = {name: vid for name, vid in stored.items() if name not in direct}
    groups: dict[tuple, list] = {}
    for name, vid in computed.items():
        groups.setdefault(_shape_of(trace, vid), []).append(vid)
    bin_shape = None
    keep: list[int] = []
    if bin_request is not None:
        _, bx, by = bin_request
        kv = [v.vid for v in (bx, by) if isinstance(v, Sym)]
        bin_shape = tuple(torch.broadcast_shapes(*[_shape_of(trace, v) for v in kv])) if kv else (1,)
        groups.setdefault(bin_shape, [])
        keep = kv
    results: dict[int, torch.Tensor] = {}
    programs = []
    for shape, vids in groups.items():
        uniq = list(dict.fromkeys(vids))
        is_bin = bin_request is not None and shape == bin_shape
        prefixes, comp, prefix_comps = _compile_segments(trace, uniq, keep if is_bin else [], shape)
        programs.extend(prefixes)
        for _, outs in prefix_comps:  # outputs stored by a prefix program
            results.update(outs)
        spec = None
        if is_bin:
            spec, bx, by = bin_request
            # registers of the binning positions (or uniform values)
            xr = comp.keep_regs[bx.vid] if isinstance(bx, Sym) else -1
            yr = comp.keep_regs[by.vid] if isinstance(by, Sym) else -1
            if not isinstance(bx, Sym):
                spec.x_uniform = _uniform_as(bx, comp.dtype)
            if not isinstance(by, Sym):
                spec.y_uniform = _uniform_as(by, comp.dtype)
            comp.bin_regs = (xr, yr)
        prog, outs, keepalive = prepare(trace, comp, spec)
        results.update(outs)
        programs.append((prog, spec, keepalive))
    out_vals: list[dict] = [{} for _ in rays]
    field_src = []  # the primary ray's fields: ("slot", s) | ("vid", v) | ("value", v)
    for (i, name), v in fields.items():
        if (i, name) in direct:
            s = trace.value_slot[direct[(i, name)]]
            out_vals[i][name] = trace.slot_tensors[s]
            src = ("slot", s)
        elif (i, name) in computed:
            out_vals[i][name] = results[computed[(i, name)]]
            src = ("vid", computed[(i, name)])
        else:
            out_vals[i][name] = v
            src = ("value", v)
        if i == 0:
            field_src.append((name, src))
    plan = Plan(trace, programs, _rebuild(sym_ray, out_vals[0]))
    plan.extra_out = [_rebuild(r, vals) for r, vals in zip(rays[1:], out_vals[1:])]
    plan.field_src = field_src
    return plan


def _materialize(trace: Trace, sym_ray, bin_request=None):
    return build_plan(trace, sym_ray, bin_request).launch()


def _uniform_as(v, dtype) -> float:
    """A weakly typed scalar position as the binning kernel sees it."""
    return float(np.float32(v)) if dtype == torch.float32 else float(v)


def plan_fn(fn: Callable, ray, bin_request_fn: Callable | None = None) -> Plan:
    """Trace `fn` (a Ray -> Ray function of component bodies) on a vectorised
    ray and compile it.  If `bin_request_fn(sym_out)` is given it returns
    (BinSpec, x, y) to bin the final positions in the same launch."""
    device = _lib.require_device()
    trace = Trace(device)
    sym_in = trace.symbolize(ray)
    sym_out = fn(sym_in)
    req = bin_request_fn(sym_out) if bin_request_fn is not None else None
    return build_plan(trace, sym_out, req)


def apply_fn(fn: Callable, ray, bin_request_fn: Callable | None = None):
    """Trace, compile and launch `fn` on a vectorised ray; returns the output ray."""
    return plan_fn(fn, ray, bin_request_fn).launch()


# --------------------------------------------------------------------------
# plan cache: repeated drop-in calls skip tracing and compilation
# --------------------------------------------------------------------------

class Uncacheable(Exception):
    """A call whose program may depend on something the cache key cannot see."""


_ATOM_TYPES = frozenset((type(None), str, bool, torch.device, torch.dtype))
_DC_INFO: dict = {}        # class -> (field names, frozen) or False
_FROZEN_KEYS: dict = {}    # id(obj) -> (weakref, key) of frozen, array-free dataclasses
_FROZEN_KEYS_MAX = 4096
_MUTABLE = -1              # `seen` entry counting mutable containers walked


def _dc_info(t):
    info = _DC_INFO.get(t)
    if info is None:
        info = (tuple(f.name for f in dataclasses.fields(t)), bool(t.__dataclass_params__.frozen)) \
            if dataclasses.is_dataclass(t) else False
        _DC_INFO[t] = info
    return info


def _array_key(obj, leaves: list, seen: dict):
    first = seen.get(id(obj))
    if first is not None:
        return ("alias", first)
    seen[id(obj)] = len(leaves)
    leaves.append(obj)
    if isinstance(obj, torch.Tensor):
        k = (obj.dtype, obj.shape, obj.stride(), obj.device, obj.requires_grad)
        single = obj.numel() == 1
    else:
        k = (obj.dtype.str, obj.shape, obj.strides)
        single = obj.size == 1
    if single:
        v = obj.item()
        k += (type(v), v.hex() if type(v) is float else v)
    return k


def _structure_key(obj, leaves: list, seen: dict):
    """Hashable key of everything in `obj` a traced program can depend on.

    Python scalars fold into programs as immediates, so they are keyed by type
    and exact value.  Arrays become input slots: they are keyed by dtype,
    shape, strides and device, and collected in `leaves` (in walk order) so a
    cached plan can bind the next call's arrays.  One-element arrays are also
    keyed by value (host math may fold them: ``float(p)``, ``.item()``).  An
    array met again is keyed as an alias of its first occurrence, because the
    tracer gives each distinct object one slot.  Dataclasses, tuples, lists
    and dicts are walked, and so are the instance attributes of classes that
    set ``_plan_key_by_attributes`` (propagators); functions without closures
    are keyed by identity; anything else is Uncacheable.  The key of a frozen
    dataclass that holds no arrays and no mutable containers cannot change,
    so it is remembered per object.
    """
    t = type(obj)
    if t is float:
        return (float, obj.hex())
    if t is torch.Tensor or t is np.ndarray:
        return _array_key(obj, leaves, seen)
    if t is int:
        return (int, obj)
    if t in _ATOM_TYPES:
        return obj
    info = _DC_INFO.get(t)
    if info is None:
        info = _dc_info(t)
    if info:
        names, frozen = info
        if not frozen:
            return (t,) + tuple([_structure_key(getattr(obj, n), leaves, seen) for n in names])
        memo = _FROZEN_KEYS.get(id(obj))
        if memo is not None and memo[0]() is obj:
            return memo[1]
        n_leaves, n_mutable = len(leaves), seen.get(_MUTABLE, 0)
        key = (t,) + tuple([_structure_key(getattr(obj, n), leaves, seen) for n in names])
        if len(leaves) == n_leaves and seen.get(_MUTABLE, 0) == n_mutable:
            if len(_FROZEN_KEYS) >= _FROZEN_KEYS_MAX:
                _FROZEN_KEYS.clear()
            try:
                _FROZEN_KEYS[id(obj)] = (weakref.ref(obj), key)
            except TypeError:  # no weak references to this class
                pass
        return key
    if isinstance(obj, tuple):
        return (t,) + tuple([_structure_key(v, leaves, seen) for v in obj])
    if isinstance(obj, (torch.Tensor, np.ndarray)):
        return _array_key(obj, leaves, seen)
    if isinstance(obj, type):
        return obj
    if isinstance(obj, numbers.Number):
        if isinstance(obj, (complex, np.complexfloating)):
            return (t, float(obj.real).hex(), float(obj.imag).hex())
        if isinstance(obj, numbers.Integral):
            return (t, int(obj))
        return (t, float(obj).hex())
    if isinstance(obj, list):
        seen[_MUTABLE] = seen.get(_MUTABLE, 0) + 1
        return (t,) + tuple([_structure_key(v, leaves, seen) for v in obj])
    if isinstance(obj, dict):
        seen[_MUTABLE] = seen.get(_MUTABLE, 0) + 1
        return (dict,) + tuple((k, _structure_key(obj[k], leaves, seen)) for k in sorted(obj))
    if isinstance(obj, types.FunctionType) and obj.__closure__ is None:
        return obj
    if getattr(t, "_plan_key_by_attributes", False):
        seen[_MUTABLE] = seen.get(_MUTABLE, 0) + 1
        return (t, _structure_key(getattr(obj, "__dict__", {}), leaves, seen))
    raise Uncacheable(t.__name__)


class _PlanTemplate:
    """The compiled programs of one cache key with their pointers unbound."""

    __slots__ = ("device", "programs", "fields", "ray_type", "convert")

    def __init__(self, plan: Plan, leaves: list):
        trace = plan.trace
        where = {}
        for i, leaf in enumerate(leaves):
            where.setdefault(id(leaf), i)
        slot_src = []
        # leaves the tracer converted (host arrays, other dtypes) are converted
        # again on every call; device tensors it used as they were are bound as-is
        self.convert = set()
        for s, src in enumerate(trace.slot_source):
            if src is None:
                slot_src.append(("vid", trace.slot_value[s]))
            elif id(src) in where:
                slot_src.append(("leaf", where[id(src)]))
                if trace.slot_tensors[s] is not src:
                    self.convert.add(where[id(src)])
            else:
                raise Uncacheable("an input array is not reachable from the call's arguments")
        self.device = trace.device
        self.programs = []
        for prog, spec, _ in plan.programs:
            groups: dict = {}
            for k, (vid, shape, dt) in enumerate(prog.bound_outputs):
                groups.setdefault((tuple(shape), dt), []).append((k, vid))
            self.programs.append((_lib.Program.from_buffer_copy(prog),
                                  None if spec is None else _lib.BinSpec.from_buffer_copy(spec),
                                  [slot_src[s] for s in prog.bound_slots], list(groups.items())))
        self.fields = []
        for name, (kind, x) in plan.field_src:
            if kind == "slot":
                src = slot_src[x]
                if src[0] != "leaf":
                    raise Uncacheable("an output field is an intermediate buffer")
                self.fields.append((name, "input", src[1]))
            elif kind == "vid":
                self.fields.append((name, "vid", x))
            elif isinstance(x, (torch.Tensor, np.ndarray)):
                if id(x) not in where:
                    raise Uncacheable("an output array is not one of the call's arguments")
                self.fields.append((name, "leaf", where[id(x)]))
            elif x is None or isinstance(x, (str, bool, numbers.Number)):
                self.fields.append((name, "const", x))
            else:
                raise Uncacheable(type(x).__name__)
        self.ray_type = type(plan.out_ray)

    def bind(self, leaves: list, acc: torch.Tensor | None):
        """Copies of the programs with this call's arrays and fresh outputs
        bound: ([(Program, BinSpec | None)], output ray, (inputs, outputs)).
        The outputs of one program that share a shape and dtype are views of
        one allocation."""
        dev = self.device
        inputs: dict[int, torch.Tensor] = {}
        outs: dict[int, torch.Tensor] = {}
        convert = self.convert

        def input_tensor(i):
            t = inputs.get(i)
            if t is None:
                t = inputs[i] = to_device_tensor(leaves[i], dev) if i in convert else leaves[i]
            return t

        bound = []
        for prog_t, spec_t, binds, groups in self.programs:
            prog = _lib.Program.from_buffer_copy(prog_t)
            slots = prog.slot
            for k, (kind, x) in enumerate(binds):
                slots[k].ptr = (input_tensor(x) if kind == "leaf" else outs[x]).data_ptr()
            for (shape, dt), members in groups:
                if len(members) == 1:
                    views = (torch.empty(shape, dtype=dt, device=dev),)
                else:
                    views = torch.empty((len(members),) + shape, dtype=dt, device=dev).unbind(0)
                for (k, vid), o in zip(members, views):
                    outs[vid] = o
                    prog.out[k] = o.data_ptr()
            spec = None
            if spec_t is not None:
                spec = _lib.BinSpec.from_buffer_copy(spec_t)
                spec.acc = acc.data_ptr()
            bound.append((prog, spec))
        vals = {}
        for name, kind, x in self.fields:
            if kind == "vid":
                vals[name] = outs[x]
            elif kind == "input":
                vals[name] = input_tensor(x)
            elif kind == "leaf":
                vals[name] = leaves[x]
            else:
                vals[name] = x
        return bound, _rebuild_type(self.ray_type, vals), (inputs, outs)

    def launch(self, leaves: list, acc: torch.Tensor | None, stream: int):
        """Bind this call's arrays and fresh outputs, launch, return the output ray."""
        run = _lib.load().tg_ray_program_run
        bound, out_ray, _ = self.bind(leaves, acc)
        for prog, spec in bound:
            st = run(ctypes.byref(prog), None if spec is None else ctypes.byref(spec), stream)
            if st != 0:
                _lib.check(st, "tg_ray_program_run")
        return out_ray


_INIT_SPLIT: dict = {}


def _rebuild_type(cls, vals: dict):
    split = _INIT_SPLIT.get(cls)
    if split is None:
        fs = dataclasses.fields(cls)
        split = _INIT_SPLIT[cls] = (tuple(f.name for f in fs if f.init), tuple(f.name for f in fs if not f.init))
    out = cls(**{n: vals[n] for n in split[0]})
    for n in split[1]:
        object.__setattr__(out, n, vals[n])
    return out


class PlanCache:
    """LRU map from a call's structure key to its compiled programs.

    ``launch(key_objects, build, acc)`` keys `key_objects` (the entry point's
    name, the model, the propagator, the ray, ...).  On a miss it runs
    ``build() -> Plan``, launches it and keeps a template; on a hit it binds
    the call's arrays into copies of the template's programs and launches
    them, without tracing.  Calls whose structure cannot be keyed, or whose
    plan reads arrays the key did not see, always take the uncached path.
    """

    def __init__(self, max_entries: int = 128):
        self.max_entries = max_entries
        self._map: "collections.OrderedDict[tuple, _PlanTemplate | None]" = collections.OrderedDict()
        self._lock = threading.Lock()
        self.hits = self.misses = self.uncached = 0

    def clear(self):
        with self._lock:
            self._map.clear()
            self.hits = self.misses = self.uncached = 0

    def __len__(self):
        return len(self._map)

    def launch(self, key_objects: tuple, build: Callable[[], Plan], acc: torch.Tensor | None = None,
               stream: int | None = None):
        leaves: list = []
        try:
            key = _structure_key(key_objects, leaves, {})
        except Uncacheable:
            self.uncached += 1
            plan = build()
            return plan.launch(stream)
        with self._lock:
            tpl = self._map.get(key, False)
            if tpl is not False:
                self._map.move_to_end(key)
        if tpl is None:  # known to be uncacheable
            self.uncached += 1
            return build().launch(stream)
        if tpl is not False:
            self.hits += 1
            st = _lib.stream_ptr(tpl.device) if stream is None else stream
            return tpl.launch(leaves, acc, st)
        self.misses += 1
        plan = build()
        try:
            tpl = _PlanTemplate(plan, leaves)
        except Uncacheable:
            tpl = None
        with self._lock:
            self._map[key] = tpl
            while len(self._map) > self.max_entries:
                self._map.popitem(last=False)
        return plan.launch(stream)


plan_cache = PlanCache()
