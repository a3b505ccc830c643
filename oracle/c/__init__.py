"""OpenMP C port used as the timed CPU baseline (test / baseline infrastructure).

``build()`` compiles cpu_port.c with gcc into oracle/_build/libcpu_port.so;
``load()`` returns a ctypes handle with typed signatures.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(HERE), "_build")
LIB = os.path.join(OUT_DIR, "libcpu_port.so")


def build(verbose: bool = False) -> str:
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        raise RuntimeError("no C compiler for the CPU port")
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [cc, "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", os.path.join(HERE, "cpu_port.c"),
           "-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        lib.cpu_threads.restype = ctypes.c_int
        lib.plain_run.argtypes = [ctypes.c_int64, P, P, P, P, ctypes.c_double, P, ctypes.c_int64, ctypes.c_int64,
                                  P, P, P, P, P, P]
        lib.plain_run.restype = None
        lib.stem_fit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, P, P, P, ctypes.c_int64, ctypes.c_int64, P]
        lib.stem_fit.restype = None
        _lib = lib
    return _lib
