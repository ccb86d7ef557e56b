"""Chemistry right-hand side and analytical Jacobian on the GPU.

Same names, argument order and shapes as chem.py of the reference:
``chem_rhs(y, M, k, net)`` (chem.py:128-131) and
``chem_jac_analytical(y, M, k, net)`` (chem.py:211-214).  A leading batch
axis on ``y`` plays the role of ``jax.vmap``; ``M`` and ``k`` may be shared
or batched.  Inputs are torch tensors (any device) or NumPy arrays; the
results are float64 tensors on the GPU.
"""
from __future__ import annotations

import torch

from . import _lib
from ._types import NetworkArrays
from .core import F64, device, device_network, uniform_batch

__all__ = ["NetworkArrays", "to_jax", "chem_rhs", "chem_jac_analytical"]


def to_jax(net: NetworkArrays):
    """Upload a network once; the device handle is cached per object."""
    return device_network(net)


def _prep(y, M, k):
    dev = device()
    y = torch.as_tensor(y, dtype=F64, device=dev)
    squeeze = y.dim() == 2
    if squeeze:
        y = y.unsqueeze(0)
    B, nz, ni = y.shape
    M = torch.as_tensor(M, dtype=F64, device=dev)
    if M.dim() == 1:
        M = M.unsqueeze(0).expand(B, nz)
    k = torch.as_tensor(k, dtype=F64, device=dev)
    if k.dim() == 2:
        k = k.unsqueeze(0).expand(B, *k.shape)
    return y.contiguous(), M.contiguous(), k.contiguous(), squeeze


def chem_rhs(y, M, k, net) -> torch.Tensor:
    """prod - loss per species and layer (chem.py:83-131); shape of y."""
    y, M, k, squeeze = _prep(y, M, k)
    B, nz, ni = y.shape
    b = uniform_batch(B, nz, net)
    out = torch.empty_like(y)
    _lib.check(_lib.load().vk_chem_rhs_f64(b.handle, y.data_ptr(), M.data_ptr(), None,
                                           k.data_ptr(), out.data_ptr(), _lib.stream_ptr()),
               "vk_chem_rhs_f64")
    return out[0] if squeeze else out


def chem_jac_analytical(y, M, k, net) -> torch.Tensor:
    """d(chem_rhs)/dy per layer (chem.py:144-214); shape (..., nz, ni, ni)."""
    y, M, k, squeeze = _prep(y, M, k)
    B, nz, ni = y.shape
    b = uniform_batch(B, nz, net)
    out = torch.empty((B, nz, ni, ni), dtype=F64, device=y.device)
    _lib.check(_lib.load().vk_chem_jac_f64(b.handle, y.data_ptr(), M.data_ptr(), None,
                                           k.data_ptr(), out.data_ptr(), _lib.stream_ptr()),
               "vk_chem_jac_f64")
    return out[0] if squeeze else out
