"""Implicit-function-theorem gradients of a converged column, on the GPU.

Drop-in for steady_state_grad.py of the reference (steady_state_grad.py:
1-330), with torch autograd in place of ``jax.custom_vjp``.  The converged
state y* satisfies f(y*, theta) = 0, so for a loss L(y*)

    dL/dtheta = -((df/dy)^{-T} v) . (df/dtheta),   v = dL/dy*.

Backward = one transposed block-Thomas factor + solve (``vk_bt_factor_f64``
with ``transpose=1`` and the off-diagonals swapped, steady_state_grad.py:
314-317) and one parameter VJP of f (``vk_ss_vjp_f64``: the chemistry part
for k and M, dual-number diffusion rows for every other atmosphere leaf).
Nothing is taped through the forward run, so memory is O(1) in step count;
gradient accuracy is bounded by the forward residual ||f(y*, theta)||.

Inputs are torch tensors (any device) or NumPy arrays; results are float64
tensors on the GPU.  ``y_star`` is (nz, ni); the leaves follow AtmStatic.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from ._types import ATM_FLAGS, ATM_RECORD_FIELDS, AtmStatic, NetworkArrays
from .core import F64, atm_record_layout, device, uniform_batch

__all__ = [
    "SteadyStateInputs", "build_steady_state_inputs", "steady_state_residual",
    "steady_state_residual_inputs", "validate_steady_state_solution",
    "differentiable_steady_state_inputs", "checked_differentiable_steady_state",
    "steady_state_value_and_grad", "differentiable_steady_state", "newton_steady_state",
    "adjoint_state", "pseudo_transient_steady_state",
]

ST_SINGULAR = 1


class SteadyStateInputs(NamedTuple):
    """Differentiable inputs of the residual (steady_state_grad.py:36-59)."""
    k_arr: object
    Kzz: object
    Dzz: object
    dzi: object
    vz: object
    Hpi: object
    Ti: object
    Tco: object
    g: object
    ms: object
    alpha: object
    M: object
    vm: object
    vs: object
    top_flux: object
    bot_flux: object
    bot_vdep: object
    gas_indx_mask: object
    use_vm_mol: object
    use_settling: object
    use_topflux: object
    use_botflux: object


_FLOAT_LEAVES = ("k_arr",) + tuple(f for f in ATM_RECORD_FIELDS if f != "gas_indx_mask")
_INT_LEAVES = ("gas_indx_mask",) + ATM_FLAGS


def _t(x, dtype=F64):
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=dtype)
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device())


def build_steady_state_inputs(k_arr, atm: AtmStatic) -> SteadyStateInputs:
    """Pack an AtmStatic plus k_arr (steady_state_grad.py:62-87)."""
    vals = {"k_arr": _t(k_arr)}
    for f in ATM_RECORD_FIELDS:
        vals[f] = _t(getattr(atm, f), torch.bool if f == "gas_indx_mask" else F64)
    for f in ATM_FLAGS:
        vals[f] = torch.as_tensor(bool(np.asarray(getattr(atm, f))))
    return SteadyStateInputs(**vals)


def _atm_from_inputs(inputs: SteadyStateInputs) -> AtmStatic:
    return AtmStatic(**{f: getattr(inputs, f) for f in ATM_RECORD_FIELDS + ATM_FLAGS})


def _record(inputs: SteadyStateInputs, nz: int, ni: int) -> torch.Tensor:
    """The packed atmosphere record (vk_atm_record_layout order) on the device."""
    parts = [_t(getattr(inputs, f)).detach().reshape(-1) for f in ATM_RECORD_FIELDS]
    parts.append(torch.tensor([float(bool(getattr(inputs, f))) for f in ATM_FLAGS], dtype=F64,
                              device=device()))
    rec = torch.cat(parts)
    total, _ = atm_record_layout(nz, ni)
    if rec.numel() != total:
        raise ValueError(f"atmosphere leaves hold {rec.numel()} values, expected {total} "
                         f"for nz={nz}, ni={ni}")
    return rec


def _split_record(rec: torch.Tensor, inputs: SteadyStateInputs, nz: int, ni: int) -> dict:
    _, off = atm_record_layout(nz, ni)
    out = {}
    for j, f in enumerate(ATM_RECORD_FIELDS):
        shape = tuple(np.shape(getattr(inputs, f)))
        n = int(np.prod(shape)) if shape else 1
        out[f] = rec[off[j]:off[j] + n].reshape(shape)
    return out


def _prep(y, k_arr, net):
    y = _t(y).contiguous()
    if y.dim() != 2:
        raise ValueError(f"y must be (nz, ni), got {tuple(y.shape)}")
    nz, ni = y.shape
    if ni != net.ni:
        raise ValueError(f"y has {ni} species, the network {net.ni}")
    k = _t(k_arr).contiguous()
    if tuple(k.shape) != (net.nr + 1, nz):
        raise ValueError(f"k_arr must be ({net.nr + 1}, {nz}), got {tuple(k.shape)}")
    return y, k, nz, ni


def _residual(y, k, rec, net, nz) -> torch.Tensor:
    b = uniform_batch(1, nz, net)
    out = torch.empty_like(y)
    _lib.check(_lib.load().vk_residual_f64(b.handle, y.data_ptr(), k.data_ptr(), rec.data_ptr(),
                                           out.data_ptr(), _lib.stream_ptr()),
               "vk_residual_f64")
    return out


def steady_state_residual_inputs(y, inputs: SteadyStateInputs, net: NetworkArrays,
                                 grav=None) -> torch.Tensor:
    """f(y, inputs) = chem_rhs(y) + diffusion(y) (steady_state_grad.py:116-130).

    ``grav`` is accepted for signature parity; the device recomputes the
    y-independent molecular-diffusion terms in the same launch.
    """
    y, k, nz, ni = _prep(y, inputs.k_arr, net)
    return _residual(y, k, _record(inputs, nz, ni), net, nz)


def steady_state_residual(y, k_arr, atm: AtmStatic, net: NetworkArrays, grav=None):
    """f(y, k_arr) = chem_rhs(y) + diffusion(y) (steady_state_grad.py:133-150)."""
    return steady_state_residual_inputs(y, build_steady_state_inputs(k_arr, atm), net, grav)


def _jacobian_blocks(y, k, rec, net, nz, ni):
    b = uniform_batch(1, nz, net)
    diag = torch.empty((nz, ni, ni), dtype=F64, device=y.device)
    sup = torch.zeros((nz, ni), dtype=F64, device=y.device)
    sub = torch.zeros((nz, ni), dtype=F64, device=y.device)
    _lib.check(_lib.load().vk_jacobian_blocks_f64(b.handle, y.data_ptr(), k.data_ptr(),
                                                  rec.data_ptr(), diag.data_ptr(), sup.data_ptr(),
                                                  sub.data_ptr(), _lib.stream_ptr()),
               "vk_jacobian_blocks_f64")
    return diag, sup, sub


def _build_jacobian_blocks(y, k_arr, atm: AtmStatic, net: NetworkArrays):
    """(diag, sup_d, sub_d) of J_y = df/dy (steady_state_grad.py:153-175)."""
    inputs = build_steady_state_inputs(k_arr, atm)
    y, k, nz, ni = _prep(y, k_arr, net)
    diag, sup, sub = _jacobian_blocks(y, k, _record(inputs, nz, ni), net, nz, ni)
    return diag, sup[:-1].clone(), sub[:-1].clone()


_LATENCY = {"on": False}


def set_latency_mode(on: bool) -> None:
    """Factor and solve the steady-state systems (Newton polish, adjoint) by
    block cyclic reduction (vk_bt_cr_factor_f64): about log2(nz) block
    inversions deep instead of nz, for one or a few columns.  Results agree
    with the default block Thomas to the system's conditioning."""
    _LATENCY["on"] = bool(on)


def _factor_solve(diag, sup, sub, rhs, net, nz, transpose: bool):
    """Factor (optionally the transposed system) and solve one right-hand side."""
    b = uniform_batch(1, nz, net)
    lib = _lib.load()
    if _LATENCY["on"]:
        status = torch.zeros(1, dtype=torch.int32, device=diag.device)
        a, c = (sub, sup) if transpose else (sup, sub)
        nbytes = int(lib.vk_bt_cr_workspace_bytes(b.handle))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=diag.device)
        _lib.check(lib.vk_bt_cr_factor_f64(b.handle, diag.data_ptr(), a.data_ptr(), c.data_ptr(),
                                           int(transpose), ws.data_ptr(), nbytes,
                                           status.data_ptr(), _lib.stream_ptr()),
                   "vk_bt_cr_factor_f64")
        x = torch.empty_like(rhs)
        _lib.check(lib.vk_bt_cr_solve_f64(b.handle, a.data_ptr(), c.data_ptr(),
                                          rhs.contiguous().data_ptr(), x.data_ptr(),
                                          ws.data_ptr(), nbytes, _lib.stream_ptr()),
                   "vk_bt_cr_solve_f64")
        return x, status
    inv = torch.empty(b.blk_len, dtype=F64, device=diag.device)
    status = torch.zeros(1, dtype=torch.int32, device=diag.device)
    a, c = (sub, sup) if transpose else (sup, sub)
    _lib.check(lib.vk_bt_factor_f64(b.handle, diag.data_ptr(), a.data_ptr(), c.data_ptr(),
                                    int(transpose), inv.data_ptr(), status.data_ptr(),
                                    _lib.stream_ptr()), "vk_bt_factor_f64")
    x = torch.empty_like(rhs)
    scratch = torch.empty_like(rhs)
    _lib.check(lib.vk_bt_solve_f64(b.handle, inv.data_ptr(), a.data_ptr(), c.data_ptr(),
                                   rhs.contiguous().data_ptr(), x.data_ptr(), scratch.data_ptr(),
                                   _lib.stream_ptr()), "vk_bt_solve_f64")
    return x, status


def adjoint_state(y_star, inputs: SteadyStateInputs, net: NetworkArrays, v):
    """lambda = J_y^{-T} v by the transposed block Thomas (steady_state_grad.py:314-317).

    Returns (lambda, singular) where ``singular`` is True when the factor met
    a zero pivot (lambda is then not finite, as in the reference).
    """
    y, k, nz, ni = _prep(y_star, inputs.k_arr, net)
    rec = _record(inputs, nz, ni)
    diag, sup, sub = _jacobian_blocks(y, k, rec, net, nz, ni)
    lam, status = _factor_solve(diag, sup, sub, _t(v).reshape(nz, ni), net, nz, True)
    return lam, bool(int(status.item()) & ST_SINGULAR)


def _param_vjp(y, k, rec, lam, net, nz, want_k: bool, want_atm: bool):
    b = uniform_batch(1, nz, net)
    cot_k = torch.empty_like(k) if want_k else None
    cot_rec = torch.empty_like(rec) if want_atm else None
    _lib.check(_lib.load().vk_ss_vjp_f64(b.handle, y.data_ptr(), k.data_ptr(), rec.data_ptr(),
                                         lam.data_ptr(), _lib.ptr(cot_k), _lib.ptr(cot_rec),
                                         _lib.stream_ptr()), "vk_ss_vjp_f64")
    return cot_k, cot_rec


def validate_steady_state_solution(y_star, inputs: SteadyStateInputs, net: NetworkArrays,
                                   residual_rtol: float = 1e-6,
                                   residual_atol: float = 0.0) -> float:
    """Residual gate before implicit AD (steady_state_grad.py:178-196)."""
    resid = steady_state_residual_inputs(y_star, inputs, net)
    resid_inf = float(resid.abs().max().item())
    state_scale = float(_t(y_star).abs().max().item())
    tol = max(residual_atol, residual_rtol * max(state_scale, 1.0))
    if not resid_inf <= tol:
        raise ValueError(
            "Steady-state residual is too large for reliable implicit differentiation: "
            f"||f||_inf={resid_inf:.3e} exceeds tolerance {tol:.3e}. "
            "Tighten the forward convergence criterion or provide a more converged state.")
    return resid_inf


class _ImplicitSteadyState(torch.autograd.Function):
    """Identity on y* forward; implicit-function-theorem VJP backward."""

    @staticmethod
    def forward(ctx, y_star, ctx_data, *leaves):
        net, inputs = ctx_data
        ctx.net = net
        ctx.inputs = inputs
        ctx.save_for_backward(y_star.detach())
        return y_star.detach().clone()

    @staticmethod
    def backward(ctx, v):
        (y_star,) = ctx.saved_tensors
        net, inputs = ctx.net, ctx.inputs
        need = ctx.needs_input_grad[2:]
        want_k = need[0]
        want_atm = any(need[1:])
        grads = [None] * len(_FLOAT_LEAVES)
        if want_k or want_atm:
            y, k, nz, ni = _prep(y_star, inputs.k_arr, net)
            rec = _record(inputs, nz, ni)
            diag, sup, sub = _jacobian_blocks(y, k, rec, net, nz, ni)
            lam, _ = _factor_solve(diag, sup, sub, v.detach().to(F64).reshape(nz, ni), net, nz,
                                   True)
            cot_k, cot_rec = _param_vjp(y, k, rec, lam, net, nz, want_k, want_atm)
            if want_k:
                grads[0] = cot_k
            if want_atm:
                parts = _split_record(cot_rec, inputs, nz, ni)
                for j, f in enumerate(_FLOAT_LEAVES[1:], start=1):
                    if need[j]:
                        grads[j] = parts[f]
        # the y* cotangent is zero: y* comes from an external solver
        return (torch.zeros_like(y_star), None, *grads)


def differentiable_steady_state_inputs(inputs: SteadyStateInputs, y_star,
                                       net: NetworkArrays) -> torch.Tensor:
    """y* as a differentiable function of every float leaf of ``inputs``
    (steady_state_grad.py:199-233).  Leaves that are tensors with
    ``requires_grad`` receive gradients."""
    leaves = [getattr(inputs, f) for f in _FLOAT_LEAVES]
    leaves = [x if isinstance(x, torch.Tensor) else _t(x) for x in leaves]
    plain = SteadyStateInputs(**{f: (x.detach() if isinstance(x, torch.Tensor) else x)
                                 for f, x in zip(SteadyStateInputs._fields, inputs)})
    return _ImplicitSteadyState.apply(_t(y_star), (net, plain), *leaves)


def checked_differentiable_steady_state(inputs: SteadyStateInputs, y_star, net: NetworkArrays,
                                        residual_rtol: float = 1e-6,
                                        residual_atol: float = 0.0) -> torch.Tensor:
    """Residual gate, then the implicit wrapper (steady_state_grad.py:236-251)."""
    validate_steady_state_solution(y_star, inputs, net, residual_rtol=residual_rtol,
                                   residual_atol=residual_atol)
    return differentiable_steady_state_inputs(inputs, y_star, net)


def steady_state_value_and_grad(loss_fn, inputs: SteadyStateInputs, y_star, net: NetworkArrays,
                                residual_rtol: float = 1e-6, residual_atol: float = 0.0):
    """(loss, gradient wrt every leaf of ``inputs``) (steady_state_grad.py:254-281).

    The gradient is a SteadyStateInputs; the gas mask and mode flags, which
    are not differentiable, get zero tensors of their own dtype (the
    reference's float0 leaves).
    """
    validate_steady_state_solution(y_star, inputs, net, residual_rtol=residual_rtol,
                                   residual_atol=residual_atol)
    leaves = {f: _t(getattr(inputs, f)).detach().clone().requires_grad_(True)
              for f in _FLOAT_LEAVES}
    others = {f: getattr(inputs, f) for f in _INT_LEAVES}
    req = SteadyStateInputs(**leaves, **others)
    with torch.enable_grad():
        y = differentiable_steady_state_inputs(req, y_star, net)
        value = loss_fn(y)
        grads = torch.autograd.grad(value, [leaves[f] for f in _FLOAT_LEAVES], allow_unused=True)
    out = {f: (g if g is not None else torch.zeros_like(leaves[f]))
           for f, g in zip(_FLOAT_LEAVES, grads)}
    for f in _INT_LEAVES:
        x = getattr(inputs, f)
        out[f] = torch.zeros_like(x) if isinstance(x, torch.Tensor) else np.zeros_like(x)
    return value.detach(), SteadyStateInputs(**out)


def differentiable_steady_state(k_arr, y_star, atm: AtmStatic, net: NetworkArrays):
    """y* as a differentiable function of k_arr (steady_state_grad.py:284-330).

    Pass a ``k_arr`` tensor with ``requires_grad=True``; ``torch.autograd``
    then returns dL/dk through the implicit-function theorem.
    """
    inputs = build_steady_state_inputs(k_arr, atm)
    if isinstance(k_arr, torch.Tensor):
        inputs = inputs._replace(k_arr=k_arr)
    return differentiable_steady_state_inputs(inputs, y_star, net)


def newton_steady_state(y0, k_arr, atm: AtmStatic, net: NetworkArrays, tol: float = 1e-9,
                        max_iter: int = 100):
    """Newton on f(y) = 0 with the J_y blocks and the block Thomas on the GPU.

    The polish step of the reference's gradient tests and examples
    (test_steady_state_grad.py:135-154).  Returns (y, iterations, ||f||_inf).
    """
    inputs = build_steady_state_inputs(k_arr, atm)
    y, k, nz, ni = _prep(y0, k_arr, net)
    rec = _record(inputs, nz, ni)
    y = y.clone()
    nrm = float("inf")
    for it in range(max_iter):
        res = _residual(y, k, rec, net, nz)
        nrm = float(res.abs().max().item())
        if nrm < tol:
            return y, it, nrm
        diag, sup, sub = _jacobian_blocks(y, k, rec, net, nz, ni)
        dy, _ = _factor_solve(diag, sup, sub, -res, net, nz, False)
        y = y + dy
    raise RuntimeError(f"Newton did not converge: ||res|| = {nrm:.3e}")


def pseudo_transient_steady_state(y0, k_arr, atm: AtmStatic, net: NetworkArrays,
                                  dt0: float = 1e-6, grow: float = 4.0, shrink: float = 8.0,
                                  dt_newton: float = 1e30, polish: int = 8,
                                  max_iter: int = 1000, floor: float = 0.0):
    """Steady state of f(y) = 0 by pseudo-transient continuation on the GPU.

    Each iteration solves (I/dt - J_y) dy = f(y) with the J_y blocks and the
    block-Thomas factor (a linearly implicit Euler step); a step that keeps
    every density positive is taken and dt grows by ``grow``, otherwise dt
    shrinks by ``shrink``.  Once dt reaches ``dt_newton`` the iteration is
    Newton's, and ``polish`` more Newton steps finish it.  This is the
    forward solve of the differentiation workload; the reference reaches the
    same state with its adaptive Ros2 loop (outer_loop.py:492-1112), whose
    stopping rule leaves a residual the implicit gradient then carries.

    ``floor`` > 0 takes a step that would push a density below ``floor``
    times its current value with that density held there instead, and
    leaves dt unchanged after such a step (a positivity-preserving damping
    for networks whose trace densities span many decades; the fixed point
    is the same, since near it no density is held).

    Returns (y, iterations, ||f||_inf).
    """
    inputs = build_steady_state_inputs(k_arr, atm)
    y, k, nz, ni = _prep(y0, k_arr, net)
    rec = _record(inputs, nz, ni)
    y = y.clone()
    dt = float(dt0)
    left = int(polish)
    eye = torch.arange(ni, device=y.device)
    for it in range(max_iter):
        res = _residual(y, k, rec, net, nz)
        diag, sup, sub = _jacobian_blocks(y, k, rec, net, nz, ni)
        newton = dt >= dt_newton
        if not newton:
            diag[:, eye, eye] -= 1.0 / dt
        dy, status = _factor_solve(diag, sup, sub, -res, net, nz, False)
        yn = y + dy
        held = False
        if floor > 0.0:
            low = yn < floor * y
            held = bool(low.any().item())
            yn = torch.where(low, floor * y, yn)
            ok = bool(torch.isfinite(yn).all().item())
        else:
            ok = bool(torch.all(yn > 0).item())
        ok = ok and not (int(status.item()) & ST_SINGULAR)
        if ok:
            y = yn
            if newton:
                left -= 1
                if left <= 0:
                    nrm = float(_residual(y, k, rec, net, nz).abs().max().item())
                    return y, it + 1, nrm
            if not held:
                dt = min(dt * grow, dt_newton)
        else:
            dt /= shrink
    raise RuntimeError(f"pseudo-transient continuation did not converge in {max_iter} steps")
