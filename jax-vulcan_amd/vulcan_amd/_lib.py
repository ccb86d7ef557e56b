"""ctypes binding of ``libvulcan_hip.so`` (the C ABI in include/vulcan_ros2.h).

The library is required: importing a compute wrapper without it raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvulcan_hip.so"

VK_OK = 0
_ERRNAMES = {1: "bad argument", 2: "HIP error", 3: "out of memory", 4: "unsupported size"}

p = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
sz = C.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "vk_last_error": (C.c_char_p, []),
    "vk_version": (i32, []),
    "vk_max_species": (i32, []),
    "vk_network_create": (i32, [i32, i32, i32, p, p, p, p, p, C.POINTER(p)]),
    "vk_network_destroy": (i32, [p]),
    "vk_batch_create": (i32, [i32, p, p, p, p, p, p, p, p, p, i32, C.POINTER(p)]),
    "vk_batch_uniform": (i32, [i32, i32, p, C.POINTER(p)]),
    "vk_batch_destroy": (i32, [p]),
    "vk_block_doubles": (i64, [i32]),
    "vk_batch_block_doubles": (i64, [p]),
    "vk_batch_set_latency_mode": (i32, [p, i32]),
    "vk_bt_cr_workspace_bytes": (sz, [p]),
    "vk_bt_cr_factor_f64": (i32, [p, p, p, p, i32, p, sz, p, p]),
    "vk_bt_cr_solve_f64": (i32, [p, p, p, p, p, p, sz, p]),
    "vk_atm_record_layout": (i64, [i32, i32, p]),
    "vk_chem_rhs_f64": (i32, [p, p, p, p, p, p, p]),
    "vk_chem_jac_f64": (i32, [p, p, p, p, p, p, p]),
    "vk_bt_factor_f64": (i32, [p, p, p, p, i32, p, p, p]),
    "vk_bt_solve_f64": (i32, [p, p, p, p, p, p, p, p]),
    "vk_bt_inverse_natural_f64": (i32, [p, p, p, p]),
    "vk_bt_dense_f64": (i32, [p, p, p, p, p, p, p, p, p, p]),
    "vk_diff_grav_f64": (i32, [p, p, p, p, p, p]),
    "vk_ros2_workspace_bytes": (sz, [p]),
    "vk_ros2_step_f64": (i32, [p, p, p, p, p, p, p, p, p, sz, p, p]),
    "vk_ros2_step_noalloc_f64": (i32, [p, p, p, p, p, p, p, p, p, p]),
    "vk_residual_f64": (i32, [p, p, p, p, p, p]),
    "vk_jacobian_blocks_f64": (i32, [p, p, p, p, p, p, p, p]),
    "vk_ss_init_f64": (i32, [p, p, p, p, p]),
    "vk_ss_attempt_f64": (i32, [p, p, p, p, p, p, sz, p]),
    "vk_photo_f64": (i32, [p, i32, p, p, p, p, p]),
    "vk_ss_vjp_f64": (i32, [p, p, p, p, p, p, p, p]),
    "vk_step_error_f64": (i32, [p, p, p, p, i64, p]),
    "vk_tune_set": (i32, [i32, i32, i32]),
    "vk_batch_slots": (i32, [i32, i32, i32, i32, i32, p]),
    "vk_batch_set_column": (i32, [p, i32, i32, p, p]),
    "vk_ss_init_range_f64": (i32, [p, i32, i32, p, p, p, p]),
    "vk_tune_get": (i32, [i32, i32]),
    "vk_timing_enable": (i32, [i32]),
    "vk_timing_reset": (i32, []),
    "vk_timing_read": (i32, [p, p, i32]),
    "vk_rates_create": (i32, [i32, i32, i32, i32, p, p, p, p, p, p, p, p, p, p, C.POINTER(p)]),
    "vk_rates_destroy": (i32, [p]),
    "vk_rates_eval_f64": (i32, [p, i32, i32, p, p, i32, p, p, p, p]),
    "vk_rates_vjp_f64": (i32, [p, i32, p, p, i32, p, p, p, p]),
    "vk_batch_set_rates": (i32, [p, p, i32]),
    "vk_refresh_record_layout": (i64, [i32, i32, p]),
    "vk_atm_refresh_f64": (i32, [p, i32, p, p, p, p, p, p, p]),
    "vk_rates_build_batch_f64": (i32, [p, i32, p, p, p]),
}

_lib = None


class VulcanError(RuntimeError):
    """Raised when a C entry point returns a non-zero status."""


def exported_symbols():
    return list(_SIGS)


class _Bound:
    """Only the entry points with a declared signature are reachable, so a
    64-bit device pointer can never be passed through ctypes' default
    (32-bit int) conversion."""

    def __init__(self, lib):
        self._cdll = lib
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def __getattr__(self, name):
        raise AttributeError(f"{name} has no ctypes signature in vulcan_amd._lib._SIGS")


def load():
    """Load the library once; raise if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            f"g.build()'` (hipcc, gfx950).  There is no CPU fallback.")
    # PyTorch-ROCm carries its own HIP and HSA runtimes beside the system
    # ones this library links.  The device tensors the wrappers pass are
    # torch's, so torch's runtime opens the GPU first: a first allocation
    # made by this library in a process where torch has not yet touched the
    # GPU failed (hipMalloc in vk_network_create).
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    _lib = _Bound(C.CDLL(str(LIB_PATH)))
    return _lib


def check(rc: int, what: str = ""):
    if rc != VK_OK:
        msg = load().vk_last_error().decode(errors="replace")
        raise VulcanError(f"{what}: {_ERRNAMES.get(rc, rc)}: {msg}")


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_ptr(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
