"""Block-tridiagonal Thomas solver with diagonal off-diagonal blocks (GPU).

Mirrors solver.py of the reference (solver.py:23-176): factor once, solve
many right-hand sides; ``block_thomas`` is the dense-off-diagonal form.  The factors hold the stack of block inverses
``inv(A'_j)`` (tile layout, see include/vulcan_ros2.h) instead of LU
factors and pivots; ``diag_inv_natural`` returns them as (nz, ni, ni).
A leading batch axis on every argument batches independent systems.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from ._types import NetworkArrays
from .core import F64, device, device_network, uniform_batch

_EMPTY_NETS: dict = {}


def _solver_network(ni: int):
    """A network with no reactions: the solver only needs ni from it."""
    net = _EMPTY_NETS.get(ni)
    if net is None:
        net = NetworkArrays(ni=ni, nr=0, reactant_idx=np.full((1, 1), ni, np.int64),
                            product_idx=np.full((1, 1), ni, np.int64),
                            reactant_stoich=np.zeros((1, 1)), product_stoich=np.zeros((1, 1)),
                            is_three_body=np.zeros(1, bool))
        _EMPTY_NETS[ni] = net
    return device_network(net)


class BlockThomasDiagFactors(NamedTuple):
    """Factored system: stack of block inverses plus the off-diagonals."""
    diag_inv: torch.Tensor   # flat tile-layout stack (B * nz * block_doubles)
    sup_d: torch.Tensor      # (B, nz-1, ni)
    sub_d: torch.Tensor      # (B, nz-1, ni)
    shape: tuple             # (B, nz, ni, squeezed)
    status: torch.Tensor     # (B,) int32 status bits (1 = zero pivot met)


def _as(x):
    return torch.as_tensor(x, dtype=F64, device=device())


def factor_block_thomas_diag_offdiag(diag, sup_d, sub_d, transpose: bool = False):
    """A'_j = D_j - (c_j b_{j-1}^T) o inv(A'_{j-1}); solver.py:33-64.

    ``transpose=True`` factors the system with blocks diag[j]^T without
    materialising the transpose.
    """
    diag, sup_d, sub_d = _as(diag), _as(sup_d), _as(sub_d)
    squeeze = diag.dim() == 3
    if squeeze:
        diag, sup_d, sub_d = diag[None], sup_d[None], sub_d[None]
    B, nz, ni, _ = diag.shape
    if nz < 2:
        raise ValueError("block Thomas needs nz >= 2")
    if sup_d.shape != (B, nz - 1, ni) or sub_d.shape != (B, nz - 1, ni):
        raise ValueError(f"sup_d/sub_d must be ({B}, {nz - 1}, {ni})")
    diag, sup_d, sub_d = diag.contiguous(), sup_d.contiguous(), sub_d.contiguous()
    b = uniform_batch(B, nz, _solver_network(ni))
    # sup/sub live at off_s = c*nz*ni with nz-1 rows used
    sup_p = torch.zeros((B, nz, ni), dtype=F64, device=diag.device)
    sub_p = torch.zeros((B, nz, ni), dtype=F64, device=diag.device)
    sup_p[:, :-1] = sup_d
    sub_p[:, :-1] = sub_d
    inv = torch.empty(b.blk_len, dtype=F64, device=diag.device)
    status = torch.zeros(B, dtype=torch.int32, device=diag.device)
    _lib.check(_lib.load().vk_bt_factor_f64(b.handle, diag.data_ptr(), sup_p.data_ptr(),
                                            sub_p.data_ptr(), int(bool(transpose)),
                                            inv.data_ptr(), status.data_ptr(), _lib.stream_ptr()),
               "vk_bt_factor_f64")
    return BlockThomasDiagFactors(inv, sup_p, sub_p, (B, nz, ni, squeeze), status)


def solve_block_thomas_diag_offdiag(factors: BlockThomasDiagFactors, rhs):
    """Forward sweep then back substitution; solver.py:67-112."""
    B, nz, ni, squeeze = factors.shape
    rhs = _as(rhs)
    if rhs.dim() == 2:
        rhs = rhs[None]
    rhs = rhs.expand(B, nz, ni).contiguous()
    b = uniform_batch(B, nz, _solver_network(ni))
    x = torch.empty_like(rhs)
    scratch = torch.empty_like(rhs)
    _lib.check(_lib.load().vk_bt_solve_f64(b.handle, factors.diag_inv.data_ptr(),
                                           factors.sup_d.data_ptr(), factors.sub_d.data_ptr(),
                                           rhs.data_ptr(), x.data_ptr(), scratch.data_ptr(),
                                           _lib.stream_ptr()),
               "vk_bt_solve_f64")
    return x[0] if squeeze else x


def block_thomas_diag_offdiag(diag, sup_d, sub_d, rhs):
    """Factor + solve; solver.py:115-126."""
    return solve_block_thomas_diag_offdiag(factor_block_thomas_diag_offdiag(diag, sup_d, sub_d), rhs)


def block_thomas(diag, sup, sub, rhs):
    """Dense block-tridiagonal Thomas solve (solver.py:129-176).

    Shapes: diag (nz, ni, ni), sup/sub (nz-1, ni, ni), rhs (nz, ni) -> (nz, ni);
    ``sup[j]`` couples layer j to j+1 and ``sub[j]`` layer j+1 to j.  A
    leading batch axis on every argument batches independent systems.  One
    launch factors and solves (explicit inverses of the Schur complements,
    Gauss-Jordan with partial pivoting); a zero pivot leaves the result
    non-finite, as the reference's LU does.
    """
    diag, sup, sub, rhs = _as(diag), _as(sup), _as(sub), _as(rhs)
    squeeze = diag.dim() == 3
    if squeeze:
        diag, sup, sub, rhs = diag[None], sup[None], sub[None], rhs[None]
    B, nz, ni, _ = diag.shape
    if nz < 2:
        raise ValueError("block Thomas needs nz >= 2")
    if sup.shape != (B, nz - 1, ni, ni) or sub.shape != (B, nz - 1, ni, ni):
        raise ValueError(f"sup/sub must be ({B}, {nz - 1}, {ni}, {ni})")
    if rhs.shape != (B, nz, ni):
        raise ValueError(f"rhs must be ({B}, {nz}, {ni})")
    b = uniform_batch(B, nz, _solver_network(ni))
    sup_p = torch.zeros((B, nz, ni, ni), dtype=F64, device=diag.device)
    sub_p = torch.zeros((B, nz, ni, ni), dtype=F64, device=diag.device)
    sup_p[:, :-1] = sup
    sub_p[:, :-1] = sub
    diag, rhs = diag.contiguous(), rhs.contiguous()
    x = torch.empty_like(rhs)
    scratch = torch.empty_like(rhs)
    inv = torch.empty(b.blk_len, dtype=F64, device=diag.device)
    status = torch.zeros(B, dtype=torch.int32, device=diag.device)
    _lib.check(_lib.load().vk_bt_dense_f64(b.handle, diag.data_ptr(), sup_p.data_ptr(),
                                           sub_p.data_ptr(), rhs.data_ptr(), x.data_ptr(),
                                           inv.data_ptr(), scratch.data_ptr(), status.data_ptr(),
                                           _lib.stream_ptr()), "vk_bt_dense_f64")
    return x[0] if squeeze else x


def diag_inv_natural(factors: BlockThomasDiagFactors) -> torch.Tensor:
    """The inverse stack as (B, nz, ni, ni) natural blocks."""
    B, nz, ni, squeeze = factors.shape
    b = uniform_batch(B, nz, _solver_network(ni))
    nat = torch.zeros((B, nz, ni, ni), dtype=F64, device=factors.diag_inv.device)
    _lib.check(_lib.load().vk_bt_inverse_natural_f64(b.handle, factors.diag_inv.data_ptr(),
                                                     nat.data_ptr(), _lib.stream_ptr()),
               "vk_bt_inverse_natural_f64")
    return nat[0] if squeeze else nat
