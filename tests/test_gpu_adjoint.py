"""GPU implicit-function-theorem gradients vs the oracle and finite differences.

Known-answer problem: the open chain of test_steady_state_grad.py:44-132
(0 -> A, A -> B, B -> 0, nz = 4, weak diffusion).  Its gradient is checked
against central differences with eps = 1e-6 at 5e-4 of the gradient scale
(the reference's tolerance), against the oracle's adjoint at 1e-10, and the
structured-input API against the k-only wrapper at 1e-12
(test_steady_state_grad.py:206-213).  Every parameter cotangent of
vk_ss_vjp_f64 is checked against central differences of lam . diffusion
and lam . chem_rhs (the oracle's operators, each differenced alone) on small
random columns with the diffusion modes switched on and off, and lambda
against a dense LAPACK solve of J_y^T.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _np(t):
    return t.detach().cpu().numpy()


def _open_chain():
    from test_oracle_pins import _open_chain as oc
    return oc()


def test_open_chain_newton_matches_oracle():
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    y0, k, atm, net = _open_chain()
    yg, it, res = G.newton_steady_state(y0, k, atm, net)
    yo, _, _ = O.newton_steady_state(y0, k, atm, net)
    assert res < 1e-9
    np.testing.assert_allclose(_np(yg), yo, rtol=1e-10)
    # well-mixed closed form [A] = k1/k3, [B] = k1/k5 up to the diffusion shift
    np.testing.assert_allclose(_np(yg)[:, 0], 2.0, rtol=5e-2)
    np.testing.assert_allclose(_np(yg)[:, 1], 1.0 / 0.3, rtol=5e-2)


def test_open_chain_gradient_vs_fd_oracle_and_structured_api():
    import torch
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    y0, k, atm, net = _open_chain()
    ys, _, _ = G.newton_steady_state(y0, k, atm, net)
    ys = _np(ys)

    def loss(y):
        return (y[:, 1] ** 2).sum()

    kt = torch.tensor(k, dtype=torch.float64, device="cuda", requires_grad=True)
    yd = G.differentiable_steady_state(kt, ys, atm, net)
    (g,) = torch.autograd.grad(loss(yd), kt)
    g = _np(g)
    assert np.all(np.isfinite(g))
    v = np.zeros_like(ys)
    v[:, 1] = 2.0 * ys[:, 1]
    go, lam_o = O.adjoint_grad_k(ys, k, atm, net, v)
    scale = np.abs(go).max()
    assert np.max(np.abs(g - go)) / scale < 1e-10
    # central differences, Newton re-solved on the device on each side
    fd = np.zeros_like(k)
    eps = 1e-6
    for r in (1, 3, 5):
        for z in range(k.shape[1]):
            kp, km = k.copy(), k.copy()
            kp[r, z] += eps
            km[r, z] -= eps
            yp, _, _ = G.newton_steady_state(ys, kp, atm, net, tol=1e-12)
            ym, _, _ = G.newton_steady_state(ys, km, atm, net, tol=1e-12)
            fd[r, z] = float((loss(yp) - loss(ym)).item()) / (2 * eps)
    err = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    print(f"open chain: |g - FD| / max|FD| = {err:.3e}")
    assert err < 5e-4
    # structured-input API == wrapper (test_steady_state_grad.py:206-213)
    inputs = G.build_steady_state_inputs(k, atm)
    val, gi = G.steady_state_value_and_grad(loss, inputs, ys, net, residual_rtol=1e-6)
    assert abs(float(val) - float((ys[:, 1] ** 2).sum())) < 1e-12 * float(val)
    assert np.max(np.abs(_np(gi.k_arr) - g)) / np.abs(g).max() <= 1e-12
    assert torch.isfinite(gi.Dzz).all() and torch.isfinite(gi.Kzz).all()
    # the residual gate
    with pytest.raises(ValueError):
        G.validate_steady_state_solution(y0 * 3.0, inputs, net)


def _mode_column(seed, flags):
    """Small random column with nonzero vz, vm, vs, alpha and boundary fluxes."""
    from vulcan_amd import synth
    c = synth.make_column(seed, nz=7, ni=5, n_rev=8, dt=1e-6, p_bot=1e5, p_top=1e1)
    rng = np.random.default_rng(seed + 100)
    nz, ni = c.nz, c.ni
    a = c.atm
    sgn = lambda shape: rng.choice([-1.0, 1.0], size=shape)
    atm = a._replace(
        Dzz=a.Kzz[:, None] * rng.uniform(0.05, 0.5, (nz - 1, ni)),
        vz=sgn(nz - 1) * rng.uniform(1.0, 10.0, nz - 1),
        vm=sgn((nz, ni)) * rng.uniform(0.5, 5.0, (nz, ni)),
        vs=sgn((nz - 1, ni)) * rng.uniform(0.5, 5.0, (nz - 1, ni)),
        alpha=rng.uniform(-0.4, 0.4, ni),
        ms=rng.uniform(2.0, 60.0, ni),
        top_flux=rng.uniform(-1e8, 1e8, ni), bot_flux=rng.uniform(1e8, 1e10, ni),
        bot_vdep=rng.uniform(0.01, 1.0, ni),
        gas_indx_mask=np.array([True] * (ni - 1) + [False]),
        use_vm_mol=flags[0], use_settling=flags[1], use_topflux=flags[2], use_botflux=flags[3])
    # two three-body rows so the M cotangent is exercised
    tb = np.zeros(c.net.nr + 1, bool)
    tb[[1, 4]] = True
    net = c.net._replace(is_three_body=tb)
    return c, atm, net


@pytest.mark.parametrize("flags", [(False, False, False, False), (True, True, True, True),
                                   (False, True, True, False)])
def test_parameter_cotangents_vs_fd(flags):
    import torch
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    from vulcan_amd._types import ATM_RECORD_FIELDS
    c, atm, net = _mode_column(11, flags)
    y, k = c.y, c.k
    rng = np.random.default_rng(5)
    lam = rng.standard_normal(y.shape) / np.maximum(np.abs(y), 1.0)
    inputs = G.build_steady_state_inputs(k, atm)
    yt, kt, nz, ni = G._prep(y, k, net)
    rec = G._record(inputs, nz, ni)
    cot_k, cot_rec = G._param_vjp(yt, kt, rec, torch.as_tensor(lam, device="cuda"), net, nz,
                                  True, True)
    parts = G._split_record(cot_rec, inputs, nz, ni)

    def phi_diff(a):
        # the diffusion operator alone: the chemistry term is many orders of
        # magnitude larger and would swamp the differences of the sum
        A, B, C, Ai, Bi, Ci, _ = O.build_diff_coeffs(y, a, O.compute_diff_grav(a))
        return float(np.sum(lam * O.apply_diffusion(y, A, B, C, Ai, Bi, Ci, a)))

    def phi_chem(a):
        return float(np.sum(lam * O.chem_rhs(y, a.M, k, net)))

    # k: d(lam . f)/dk is the chemistry VJP of the oracle (linear in k)
    ck = -O.chem_vjp_k(y, atm.M, k, net, lam)
    assert np.max(np.abs(_np(cot_k) - ck)) <= 1e-12 * np.abs(ck).max()
    # The boundary-flux leaves enter one row each, additively: closed forms
    # (jax_step.py:240-245).  A difference quotient cannot resolve them, since
    # their term is ~1e-8 of the row it is added to.
    use_top, use_bot = float(bool(atm.use_topflux)), float(bool(atm.use_botflux))
    closed = {"top_flux": -use_top * lam[-1] / atm.dzi[-1],
              "bot_flux": -use_bot * lam[0] / atm.dzi[0],
              "bot_vdep": use_bot * lam[0] * y[0] / atm.dzi[0]}
    # Leaves the operator is linear in take a large step (no truncation
    # error, less rounding); the piecewise-linear velocities half their own
    # size (no sign change); dzi, Hpi and Ti a relative 1e-4.
    linear = {"Kzz", "Dzz", "g", "Tco", "ms", "alpha", "M"}
    piecewise = {"vz", "vm", "vs"}
    worst = {}
    for f in ATM_RECORD_FIELDS:
        if f == "gas_indx_mask":
            assert np.all(_np(parts[f]) == 0.0)
            continue
        base = np.asarray(getattr(atm, f), dtype=np.float64)
        got = _np(parts[f]).reshape(base.shape)
        if f in closed:
            ref = closed[f]
            scale = np.abs(ref).max()
            worst[f] = 0.0 if scale == 0.0 and not got.any() else \
                float(np.max(np.abs(got - ref)) / max(scale, 1e-300))
            continue
        fd = np.zeros_like(base)
        for idx in np.ndindex(base.shape):
            if f in linear:
                h = 0.5 * max(abs(base[idx]), 0.1)
            elif f in piecewise:
                h = 0.5 * abs(base[idx])
            else:
                h = 1e-4 * abs(base[idx])
            ap, am = base.copy(), base.copy()
            ap[idx] += h
            am[idx] -= h
            phi = phi_chem if f == "M" else phi_diff
            fd[idx] = -(phi(atm._replace(**{f: ap})) - phi(atm._replace(**{f: am}))) / (2 * h)
        scale = max(np.abs(fd).max(), np.abs(got).max())
        if scale == 0.0:
            worst[f] = 0.0
            continue
        worst[f] = float(np.max(np.abs(got - fd)) / scale)
    print("cotangent errors:", {f: f"{e:.1e}" for f, e in worst.items()})
    bad = {f: e for f, e in worst.items() if e > 1e-6}
    assert not bad, bad


def test_adjoint_lambda_vs_dense_solve():
    """lambda = J_y^{-T} v from the transposed device factor vs LAPACK on J_y^T,
    at the open chain's steady state (a well-posed J_y: the random 2 -> 2
    networks conserve molecules and leave J_y singular up to the boundary
    terms)."""
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    y0, k, atm, net = _open_chain()
    ys, _, _ = O.newton_steady_state(y0, k, atm, net)
    v = np.random.default_rng(2).standard_normal(ys.shape)
    lam, singular = G.adjoint_state(ys, G.build_steady_state_inputs(k, atm), net, v)
    assert not singular
    lam = _np(lam)
    diag, sup, sub = O.jacobian_blocks(ys, k, atm, net)
    A = O.assemble_dense(diag, sup, sub)
    ref = np.linalg.solve(A.T, v.reshape(-1))
    r = A.T @ lam.reshape(-1) - v.reshape(-1)
    berr = np.abs(r).max() / (np.abs(A).max() * np.abs(lam).max() + np.abs(v).max())
    cond = np.linalg.cond(A)
    err = np.abs(lam.reshape(-1) - ref).max() / np.abs(ref).max()
    print(f"lambda: backward error {berr:.2e}, cond {cond:.2e}, rel diff {err:.2e}")
    assert berr < 1e-14
    assert err < max(1e-12, 1e-15 * cond)
