"""The Ros2 column step on the GPU (drop-in for jax_step.jax_ros2_step).

``jax_ros2_step(y, k_arr, dt, atm, net, fix_mask=None) -> (sol, delta_arr)``
keeps the reference's argument order and shapes (jax_step.py:249-315).  A
leading batch axis on ``y``/``k_arr`` runs a batch of columns in one set of
launches, like ``jax.vmap(jax_ros2_step, in_axes=(0, 0, None, None, None))``;
``atm`` may then be one AtmStatic shared by all columns or a sequence of
them, and ``dt`` a scalar or one value per column.
"""
from __future__ import annotations

from typing import NamedTuple, Sequence

import numpy as np
import torch

from . import _lib
from ._types import AtmStatic, NetworkArrays
from .core import F64, device, pack_atm_record, uniform_batch

__all__ = ["AtmStatic", "DiffGrav", "compute_diff_grav", "make_atm_static", "jax_ros2_step",
           "GAMMA"]

GAMMA = 1.0 + 1.0 / 2.0 ** 0.5


class DiffGrav(NamedTuple):
    """The y-independent parts of the molecular-diffusion coefficients
    (jax_step.py:52-64), float64 tensors on the GPU."""
    A_grav_int: torch.Tensor     # (nz-2, ni)
    B_grav_int: torch.Tensor     # (nz-2, ni)
    C_grav_int: torch.Tensor     # (nz-2, ni)
    bdry0_grav: torch.Tensor     # (2, ni): A and B of layer 0
    bdry_top_grav: torch.Tensor  # (2, ni): A and C of layer nz-1
    dz_ave: torch.Tensor         # (nz-2,)


def compute_diff_grav(atm: AtmStatic) -> DiffGrav:
    """jax_step.compute_diff_grav (jax_step.py:67-147) on the GPU
    (vk_diff_grav_f64): gravity / thermal diffusion, or the upwind vm form,
    plus settling, per the atmosphere's flags."""
    dev = device()
    nz, ni = atm.nz, atm.ni
    if nz < 3:
        raise ValueError("compute_diff_grav needs nz >= 3")
    b = uniform_batch(1, nz, _grav_network(ni))
    rec = pack_atm_record(atm).to(dev)
    A = torch.zeros((nz, ni), dtype=F64, device=dev)
    Bm = torch.zeros_like(A)
    Cm = torch.zeros_like(A)
    _lib.check(_lib.load().vk_diff_grav_f64(b.handle, rec.data_ptr(), A.data_ptr(), Bm.data_ptr(),
                                            Cm.data_ptr(), _lib.stream_ptr()), "vk_diff_grav_f64")
    dzi = torch.as_tensor(np.asarray(atm.dzi, dtype=np.float64), device=dev)
    return DiffGrav(A_grav_int=A[1:-1], B_grav_int=Bm[1:-1], C_grav_int=Cm[1:-1],
                    bdry0_grav=torch.stack([A[0], Bm[0]]),
                    bdry_top_grav=torch.stack([A[-1], Cm[-1]]),
                    dz_ave=0.5 * (dzi[:-1] + dzi[1:]))


def _grav_network(ni: int):
    """A network with no reactions: the coefficient kernel only needs ni."""
    from .solver import _solver_network
    return _solver_network(ni)


class _ReferenceDefaults:
    """The run-configuration flags make_atm_static reads, at the defaults it
    assumes when vulcan_cfg does not set them (jax_step.py:320-323)."""
    use_vm_mol = False
    use_moldiff = True
    use_settling = False
    use_topflux = False
    use_botflux = False


def make_atm_static(atm, ni: int, nz: int, cfg=None) -> AtmStatic:
    """An AtmStatic from a legacy atmosphere container (jax_step.py:318-351).

    ``atm`` carries Kzz, Dzz, dzi, vz, Hpi, Ti, Tco, g, ms, alpha, M, vm, vs,
    top_flux, bot_flux, bot_vdep and gas_indx (indices of the gas species).
    ``cfg`` is the run configuration (any object with use_vm_mol,
    use_moldiff, use_settling, use_topflux, use_botflux; missing ones take
    the reference's defaults); without it the module ``vulcan_cfg`` is read
    when importable, as the reference does.  Molecular-diffusion leaves the
    configuration switches off become zeros.
    """
    if cfg is None:
        try:
            import vulcan_cfg as cfg  # the reference's run configuration module
        except ImportError:
            cfg = _ReferenceDefaults
    flag = lambda name: bool(getattr(cfg, name, getattr(_ReferenceDefaults, name)))
    moldiff = flag("use_moldiff")
    use_vm = flag("use_vm_mol") and moldiff
    use_set = flag("use_settling") and moldiff
    f64 = lambda v: np.asarray(v, dtype=np.float64)
    gas = np.zeros(ni, dtype=bool)
    gas[np.asarray(atm.gas_indx, dtype=np.int64)] = True
    return AtmStatic(
        Kzz=f64(atm.Kzz),
        Dzz=f64(atm.Dzz) if moldiff else np.zeros((nz - 1, ni)),
        dzi=f64(atm.dzi), vz=f64(atm.vz), Hpi=f64(atm.Hpi), Ti=f64(atm.Ti), Tco=f64(atm.Tco),
        g=f64(atm.g), ms=f64(atm.ms), alpha=f64(atm.alpha), M=f64(atm.M),
        vm=f64(atm.vm) if use_vm else np.zeros((nz, ni)),
        vs=f64(atm.vs) if use_set else np.zeros((nz - 1, ni)),
        top_flux=f64(atm.top_flux), bot_flux=f64(atm.bot_flux), bot_vdep=f64(atm.bot_vdep),
        gas_indx_mask=gas, use_vm_mol=use_vm, use_settling=use_set,
        use_topflux=flag("use_topflux"), use_botflux=flag("use_botflux"))


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
