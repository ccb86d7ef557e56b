"""GPU parity: block-Thomas, chemistry and the Ros2 step vs the CPU oracle.

Every call goes through the C ABI of libvulcan_hip.so.  Tolerances:
* block-Thomas vs a dense np.linalg.solve: the reference's own bars
  (1e-12 small / 1e-8 VULCAN-shaped, test_block_thomas_diag.py:65,101);
* chemistry rhs / Jacobian: identical summation order as the oracle, so
  1e-15 relative per element (differences only from pow vs multiply);
* Ros2 step: 1e-10 relative per step, measured per
  layer against the layer's largest number density.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available, relerr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _np(t):
    return t.detach().cpu().numpy()


def _dense(diag, sup, sub):
    from oracle.ros2_np import assemble_dense
    return assemble_dense(diag, sup, sub)


def test_block_thomas_small_vs_dense():
    from vulcan_amd import solver
    rng = np.random.default_rng(7)  # test_block_thomas_diag.py:41-49
    nz, ni = 8, 5
    diag = rng.standard_normal((nz, ni, ni)) + 10.0 * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni))
    sub = rng.standard_normal((nz - 1, ni))
    rhs = rng.standard_normal((nz, ni))
    x = _np(solver.block_thomas_diag_offdiag(diag, sup, sub, rhs))
    ref = np.linalg.solve(_dense(diag, sup, sub), rhs.reshape(-1)).reshape(nz, ni)
    assert relerr(x, ref, 1e-12) < 1e-12


def test_block_thomas_vulcan_shape_vs_oracle_and_dense():
    from oracle import ros2_np as O
    from vulcan_amd import solver
    rng = np.random.default_rng(7)
    nz, ni = 40, 93  # test_block_thomas_diag.py:71-77 magnitudes
    diag = rng.standard_normal((nz, ni, ni)) + 1e10 * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni)) * 1e-3
    sub = rng.standard_normal((nz - 1, ni)) * 1e-3
    rhs = rng.standard_normal((nz, ni))
    x = _np(solver.block_thomas_diag_offdiag(diag, sup, sub, rhs))
    ref = np.linalg.solve(_dense(diag, sup, sub), rhs.reshape(-1)).reshape(nz, ni)
    assert relerr(x, ref, 1e-12) < 1e-8
    xo = O.block_thomas_diag_offdiag(diag, sup, sub, rhs)
    assert relerr(x, xo, 1e-12) < 1e-8


def test_block_thomas_pivoting_needed():
    """Blocks whose diagonal is tiny: partial pivoting is exercised."""
    from vulcan_amd import solver
    rng = np.random.default_rng(11)
    nz, ni = 12, 37
    diag = rng.standard_normal((nz, ni, ni))
    diag[:, np.arange(ni), np.arange(ni)] *= 1e-6
    sup = rng.standard_normal((nz - 1, ni)) * 0.1
    sub = rng.standard_normal((nz - 1, ni)) * 0.1
    rhs = rng.standard_normal((nz, ni))
    x = _np(solver.block_thomas_diag_offdiag(diag, sup, sub, rhs))
    A = _dense(diag, sup, sub)
    ref = np.linalg.solve(A, rhs.reshape(-1)).reshape(nz, ni)
    cond = np.linalg.cond(A)
    assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < 1e-15 * cond * 10


def test_block_thomas_transposed_vs_dense():
    from vulcan_amd import solver
    rng = np.random.default_rng(3)
    nz, ni = 10, 20
    diag = rng.standard_normal((nz, ni, ni)) + 8.0 * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni))
    sub = rng.standard_normal((nz - 1, ni))
    v = rng.standard_normal((nz, ni))
    fac = solver.factor_block_thomas_diag_offdiag(diag, sub, sup, transpose=True)
    lam = _np(solver.solve_block_thomas_diag_offdiag(fac, v))
    A = _dense(diag, sup, sub)
    ref = np.linalg.solve(A.T, v.reshape(-1)).reshape(nz, ni)
    assert relerr(lam, ref, 1e-12) < 1e-12


def test_block_inverse_stack_matches_numpy():
    from vulcan_amd import solver
    rng = np.random.default_rng(5)
    nz, ni = 6, 50
    diag = rng.standard_normal((nz, ni, ni)) + 5.0 * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni))
    sub = rng.standard_normal((nz - 1, ni))
    fac = solver.factor_block_thomas_diag_offdiag(diag, sup, sub)
    inv = _np(solver.diag_inv_natural(fac))
    A = diag[0].copy()
    for j in range(nz):
        if j > 0:
            A = diag[j] - (sub[j - 1][:, None] * sup[j - 1][None, :]) * np.linalg.inv(A)
        np.testing.assert_allclose(inv[j], np.linalg.inv(A), rtol=1e-11, atol=1e-14)


def test_block_thomas_batch_is_bitwise_single():
    from vulcan_amd import solver
    rng = np.random.default_rng(9)
    B, nz, ni = 3, 9, 69
    diag = rng.standard_normal((B, nz, ni, ni)) + 20.0 * np.eye(ni)
    sup = rng.standard_normal((B, nz - 1, ni))
    sub = rng.standard_normal((B, nz - 1, ni))
    rhs = rng.standard_normal((B, nz, ni))
    xb = _np(solver.block_thomas_diag_offdiag(diag, sup, sub, rhs))
    for c in range(B):
        xs = _np(solver.block_thomas_diag_offdiag(diag[c], sup[c], sub[c], rhs[c]))
        assert np.array_equal(xs, xb[c])


def test_chem_rhs_and_jac_match_oracle(cfg1):
    from oracle import ros2_np as O
    from vulcan_amd import chem
    c = cfg1
    f = _np(chem.chem_rhs(c.y, c.atm.M, c.k, c.net))
    fo = O.chem_rhs(c.y, c.atm.M, c.k, c.net)
    scale = np.abs(fo).max(axis=1, keepdims=True)
    assert np.max(np.abs(f - fo) / scale) < 1e-15
    J = _np(chem.chem_jac_analytical(c.y, c.atm.M, c.k, c.net))
    Jo = O.chem_jac(c.y, c.atm.M, c.k, c.net)
    assert relerr(J, Jo, 1e-300) < 1e-14


def _step_err(sol, ref):
    """max over layers of max_i |sol - ref| / max_i |ref| (per-layer scale)."""
    scale = np.max(np.abs(ref), axis=1, keepdims=True)
    return float(np.max(np.abs(sol - ref) / scale))


def test_ros2_step_cfg1_matches_oracle(cfg1):
    from oracle import ros2_np as O
    from vulcan_amd import jax_step
    c = cfg1
    sol, delta = jax_step.jax_ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    so, do = O.ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    e_sol = _step_err(_np(sol), so)
    e_del = _step_err(_np(delta), do)
    print(f"ros2 cfg1: sol {e_sol:.3e} delta {e_del:.3e}")
    assert e_sol < 1e-10
    assert np.all(np.isfinite(_np(sol)))


def test_ros2_step_fix_mask(cfg1):
    from oracle import ros2_np as O
    from vulcan_amd import jax_step
    c = cfg1
    rng = np.random.default_rng(4)
    fix = rng.random((c.nz, c.ni)) < 0.05
    sol, _ = jax_step.jax_ros2_step(c.y, c.k, c.dt, c.atm, c.net, fix_mask=fix)
    so, _ = O.ros2_step(c.y, c.k, c.dt, c.atm, c.net, fix_mask=fix)
    assert _step_err(_np(sol), so) < 1e-10


def test_ros2_step_batch_equals_single(cfg1):
    from vulcan_amd import jax_step
    import torch
    c = cfg1
    y2 = np.stack([c.y, c.y * 1.01])
    k2 = np.stack([c.k, c.k])
    sol_b, _ = jax_step.jax_ros2_step(y2, k2, c.dt, c.atm, c.net)
    sol_0, _ = jax_step.jax_ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    assert torch.equal(sol_b[0], sol_0)
