"""ctypes binding of libtemgym_hip.so (the C ABI declared in include/temgym_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m temgymcore_amd.build``).  Loading it is mandatory for every
vectorised (tensor) computation: there is no CPU fallback, and a missing or
stale library raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtemgym_hip.so")
ABI_VERSION = 7

TG_MAX_INSTR = 112
TG_MAX_REG = 64
TG_MAX_SLOT = 32
TG_MAX_DIM = 4

TG_F32 = 0
TG_F64 = 1
TG_OP_F32 = 0x80

OPCODES = {
    "add": 0, "sub": 1, "mul": 2, "div": 3,
    "addi": 4, "subi": 5, "rsubi": 6, "muli": 7, "divi": 8, "rdivi": 9,
    "neg": 10, "mov": 11, "set": 12, "sign": 13,
}


class Instr(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_uint8), ("dst", ctypes.c_uint8),
        ("a", ctypes.c_uint8), ("b", ctypes.c_uint8),
        ("immf", ctypes.c_float), ("imm", ctypes.c_double),
    ]


class Slot(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("stride", ctypes.c_int64 * TG_MAX_DIM),
        ("reg", ctypes.c_int32),
        ("contig", ctypes.c_int16),
        ("dtype", ctypes.c_int16),
    ]


class Program(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int32), ("n_instr", ctypes.c_int32),
        ("n_slot", ctypes.c_int32), ("n_out", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("shape", ctypes.c_int64 * TG_MAX_DIM),
        ("instr", Instr * TG_MAX_INSTR),
        ("slot", Slot * TG_MAX_SLOT),
        ("out", ctypes.c_void_p * TG_MAX_SLOT),
        ("out_reg", ctypes.c_int32 * TG_MAX_SLOT),
        ("out_dtype", ctypes.c_int32 * TG_MAX_SLOT),
    ]


class BinSpec(ctypes.Structure):
    _fields_ = [
        ("enabled", ctypes.c_int32), ("acc_dtype", ctypes.c_int32),
        ("x_reg", ctypes.c_int32), ("y_reg", ctypes.c_int32),
        ("x_uniform", ctypes.c_double), ("y_uniform", ctypes.c_double),
        ("m", ctypes.c_double * 6),
        ("height", ctypes.c_int64), ("width", ctypes.c_int64),
        ("acc", ctypes.c_void_p),
        ("w_reg", ctypes.c_int32), ("grid_hint", ctypes.c_int32),
    ]


# Entry points: name -> (restype, argtypes).  Every declared symbol of the
# header is listed here; tests check that the built library exports them all.
_SIGNATURES: dict[str, tuple] = {
    "tg_last_error": (ctypes.c_char_p, []),
    "tg_abi_version": (ctypes.c_int32, []),
    "tg_device_cu_count": (ctypes.c_int32, []),
    "tg_tune_set": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
    "tg_tune_get": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "tg_tune_clear": (ctypes.c_int, []),
    "tg_ray_program_run": (ctypes.c_int, [ctypes.POINTER(Program), ctypes.POINTER(BinSpec), ctypes.c_void_p]),
    "tg_hbm_read_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p]),
}

_lock = threading.Lock()
_lib = None


def register_signatures(extra: dict[str, tuple]) -> None:
    """Add entry points defined by other kernel families (beamlet, STEM...)."""
    _SIGNATURES.update(extra)
    if _lib is not None:
        _bind(_lib, extra)


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load (once) and return the ctypes handle of libtemgym_hip.so."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"temgymcore_amd: HIP library not built ({LIB_PATH} missing); "
                "run `python -m temgymcore_amd.build` (or __graft_entry__.build())"
            )
        # torch must own the HIP runtime first so that the library binds to
        # the same libamdhip64 instance (same SONAME) that torch loaded.
        torch.cuda.is_available  # noqa: B018  (import side effect only)
        lib = ctypes.CDLL(LIB_PATH)
        _bind(lib, _SIGNATURES)
        ver = lib.tg_abi_version()
        if ver != ABI_VERSION:
            raise RuntimeError(f"temgymcore_amd: library ABI {ver} != expected {ABI_VERSION}; rebuild")
        _lib = lib
        return lib


def check(status: int, what: str) -> None:
    """Map a C-ABI status code to a Python exception."""
    if status == 0:
        return
    msg = load().tg_last_error().decode(errors="replace")
    if status == -1:
        raise ValueError(f"{what}: {msg}")
    if status == -3:
        raise NotImplementedError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def require_device() -> torch.device:
    """The HIP device vectorised work runs on (torch's current device)."""
    if not torch.cuda.is_available():
        raise RuntimeError(
            "temgymcore_amd: vectorised rays run on the MI355X HIP kernels and need a GPU; "
            "no HIP device is visible"
        )
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device: torch.device | None = None) -> int:
    """Raw hipStream_t of torch's current stream on `device`."""
    return torch.cuda.current_stream(device).cuda_stream
