"""ctypes wrapper of the C oracle (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "liboracle_ros2.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build_oracle
            build_oracle.build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_ros2_steps.restype = C.c_int
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def ros2_steps(y, k, atm_records, dt, net, nthreads=1):
    """One Ros2 step for every column; y (B, nz, ni), k (B, S, nz),
    atm_records (B, L) packed records, dt (B,). Returns (sol, delta)."""
    lib = _load()
    y = _c(y, np.float64)
    B, nz, ni = y.shape
    k = _c(k, np.float64)
    atm = _c(atm_records, np.float64)
    dtv = _c(np.broadcast_to(np.asarray(dt, dtype=np.float64), (B,)), np.float64)
    r_idx = _c(net.reactant_idx, np.int64)
    p_idx = _c(net.product_idx, np.int64)
    r_st = _c(net.reactant_stoich, np.float64)
    p_st = _c(net.product_stoich, np.float64)
    is3 = _c(net.is_three_body, np.uint8)
    sol = np.empty_like(y)
    delta = np.empty_like(y)
    P = C.c_void_p
    rc = lib.oracle_ros2_steps(
        C.c_int(B), C.c_int(nz), C.c_int(ni), C.c_int(net.nr), C.c_int(r_idx.shape[1]),
        P(r_idx.ctypes.data), P(p_idx.ctypes.data), P(r_st.ctypes.data), P(p_st.ctypes.data),
        P(is3.ctypes.data), P(y.ctypes.data), P(k.ctypes.data), P(atm.ctypes.data),
        C.c_int64(atm.shape[1]), P(dtv.ctypes.data), P(sol.ctypes.data), P(delta.ctypes.data),
        C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError("oracle_ros2_steps failed")
    return sol, delta
