"""Fixed-dt Ros2 integration on the GPU (drop-in for integrate.py).

``jax_integrate_fixed_dt(y0, k_arr, dt, n_steps, atm, net) -> (y_final,
deltas)`` keeps the reference's signature (integrate.py:20-36): n_steps
steps at a frozen k_arr, and per step the scalar error
max(delta_arr / max(|sol|, 1e-30)).  A leading batch axis on ``y0`` (and
optionally ``k_arr``) integrates a batch of columns; ``deltas`` is then
(B, n_steps).  Every step is one vk_ros2_step_f64 plus one
vk_step_error_f64 launch on the current stream; the state ping-pongs
between two device buffers.
"""
from __future__ import annotations

import torch

from . import _lib
from ._types import AtmStatic, NetworkArrays
from .core import F64, device, uniform_batch
from .jax_step import _atm_buffer

__all__ = ["jax_integrate_fixed_dt"]


def jax_integrate_fixed_dt(y0, k_arr, dt: float, n_steps: int, atm: AtmStatic,
                           net: NetworkArrays):
    """Take n_steps fixed-dt Ros2 steps; returns (y_final, deltas)."""
    n_steps = int(n_steps)
    if n_steps < 0:
        raise ValueError("n_steps must be >= 0")
    dev = device()
    y = torch.as_tensor(y0, dtype=F64, device=dev)
    squeeze = y.dim() == 2
    if squeeze:
        y = y[None]
    B, nz, ni = y.shape
    if ni != net.ni:
        raise ValueError(f"y0 has {ni} species, the network {net.ni}")
    k = torch.as_tensor(k_arr, dtype=F64, device=dev)
    if k.dim() == 2:
        k = k[None].expand(B, *k.shape)
    if k.shape != (B, net.nr + 1, nz):
        raise ValueError(f"k_arr must be ({net.nr + 1}, {nz}) per column, got {tuple(k.shape)}")
    k = k.contiguous()
    atm_buf = _atm_buffer(atm, B, dev)
    b = uniform_batch(B, nz, net)
    ws = b.workspace()
    lib = _lib.load()
    dts = torch.full((B,), float(dt), dtype=F64, device=dev)
    status = torch.zeros(B, dtype=torch.int32, device=dev)
    bufs = [y.contiguous().clone(), torch.empty_like(y)]
    delta = torch.empty_like(y)
    deltas = torch.empty((B, max(n_steps, 1)), dtype=F64, device=dev)
    st = _lib.stream_ptr()
    for s in range(n_steps):
        cur, nxt = bufs[s % 2], bufs[(s + 1) % 2]
        _lib.check(lib.vk_ros2_step_f64(b.handle, cur.data_ptr(), k.data_ptr(), atm_buf.data_ptr(),
                                        dts.data_ptr(), None, nxt.data_ptr(), delta.data_ptr(),
                                        ws.data_ptr(), ws.numel(), status.data_ptr(), st),
                   "vk_ros2_step_f64")
        _lib.check(lib.vk_step_error_f64(b.handle, nxt.data_ptr(), delta.data_ptr(),
                                         deltas.data_ptr() + 8 * s, deltas.shape[1], st),
                   "vk_step_error_f64")
    y_final = bufs[n_steps % 2]
    deltas = deltas[:, :n_steps]
    if squeeze:
        return y_final[0], deltas[0]
    return y_final, deltas
