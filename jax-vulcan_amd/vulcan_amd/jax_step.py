"""The Ros2 column step on the GPU (drop-in for jax_step.jax_ros2_step).

``jax_ros2_step(y, k_arr, dt, atm, net, fix_mask=None) -> (sol, delta_arr)``
keeps the reference's argument order and shapes (jax_step.py:249-315).  A
leading batch axis on ``y``/``k_arr`` runs a batch of columns in one set of
launches, like ``jax.vmap(jax_ros2_step, in_axes=(0, 0, None, None, None))``;
``atm`` may then be one AtmStatic shared by all columns or a sequence of
them, and ``dt`` a scalar or one value per column.
"""
from __future__ import annotations

from typing import Sequence

import torch

from . import _lib
from ._types import AtmStatic, NetworkArrays
from .core import F64, device, pack_atm_record, uniform_batch

__all__ = ["AtmStatic", "jax_ros2_step", "GAMMA"]

GAMMA = 1.0 + 1.0 / 2.0 ** 0.5


def _atm_buffer(atm, B: int, dev) -> torch.Tensor:
    if isinstance(atm, torch.Tensor):
        return atm
    if isinstance(atm, AtmStatic):
        rec = pack_atm_record(atm).to(dev)
        return rec.repeat(B)
    if isinstance(atm, Sequence):
        if len(atm) != B:
            raise ValueError(f"{len(atm)} atmospheres for a batch of {B}")
        return torch.cat([pack_atm_record(a) for a in atm]).to(dev)
    raise TypeError("atm must be an AtmStatic, a sequence of them or a packed record tensor")


def jax_ros2_step(y, k_arr, dt, atm, net: NetworkArrays, fix_mask=None):
    """One second-order Rosenbrock step; returns (sol, delta_arr)."""
    dev = device()
    y = torch.as_tensor(y, dtype=F64, device=dev)
    squeeze = y.dim() == 2
    if squeeze:
        y = y[None]
    y = y.contiguous()
    B, nz, ni = y.shape
    if ni != net.ni:
        raise ValueError(f"y has {ni} species, the network {net.ni}")
    k = torch.as_tensor(k_arr, dtype=F64, device=dev)
    if k.dim() == 2:
        k = k[None].expand(B, *k.shape)
    if k.shape != (B, net.nr + 1, nz):
        raise ValueError(f"k_arr must be ({net.nr + 1}, {nz}) per column, got {tuple(k.shape)}")
    k = k.contiguous()
    dt_t = torch.as_tensor(dt, dtype=F64, device=dev).reshape(-1)
    if dt_t.numel() == 1:
        dt_t = dt_t.expand(B)
    dt_t = dt_t.contiguous()
    atm_buf = _atm_buffer(atm, B, dev)
    fix = None
    if fix_mask is not None:
        fix = torch.as_tensor(fix_mask, device=dev).to(torch.uint8)
        if fix.dim() == 2:
            fix = fix[None].expand(B, nz, ni)
        fix = fix.contiguous()
    b = uniform_batch(B, nz, net)
    ws = b.workspace()
    sol = torch.empty_like(y)
    delta = torch.empty_like(y)
    status = torch.zeros(B, dtype=torch.int32, device=dev)
    lib = _lib.load()
    _lib.check(lib.vk_ros2_step_f64(b.handle, y.data_ptr(), k.data_ptr(), atm_buf.data_ptr(),
                                    dt_t.data_ptr(), _lib.ptr(fix), sol.data_ptr(),
                                    delta.data_ptr(), ws.data_ptr(), ws.numel(),
                                    status.data_ptr(), _lib.stream_ptr()),
               "vk_ros2_step_f64")
    if squeeze:
        return sol[0], delta[0]
    return sol, delta
