"""GPU: the differentiation workload at its own size.

Problem (``synth.config4``): one open column, nz = 150, ni = 93 (92 traces
and a background gas), 1200 reactions (first-order sinks of every trace,
zeroth-order sources of 20, 500 reversible three-body isomerisations, 587
reversible 2->2 reactions with the workload-1 Arrhenius distributions,
detailed-balance K_eq), eddy diffusion with the workload-1 Kzz, bottom
fluxes of 10 traces (log-uniform 1e5..1e12 cm-2 s-1) and deposition.

Forward: the steady state y* on the GPU by pseudo-transient continuation
(block cyclic reduction), Newton-polished.

Loss L = sum over layers of log ymix of 5 random traces.  Parameters: one
log multiplier per reaction (1200), log T(z) (150; entering k(T), Tco, Ti,
Hpi, g and the hydrostatic dzi at fixed number density) and the 10 log
bottom fluxes.  Gradient on the GPU: one transposed block-Thomas factor and
solve for lambda = J_y^{-T} dL/dy*, one parameter VJP of f
(steady_state_grad.py:306-327), then the chain rule through the leaf
generator: d k / d log T = k (b + E/T) exactly, the atmosphere leaves'
tangents by central differences in log T.

Checks:
* all 1200 rate gradients against the oracle's restatement of the
  reference custom_vjp (lambda by the oracle's LU block Thomas, the rate
  VJP of chem_rhs);
* 20 random parameters (8 rate multipliers among those whose gradient is at
  least 1e-2 of the largest, 8 temperatures, 4 fluxes) against the same
  restatement (df/dtheta at fixed y*: its rate part through the oracle's
  rate VJP, its atmosphere part by central differences of the oracle's
  residual), and against central differences of L itself
  (h = 1e-6 in log space, Newton re-solved on the GPU on each side).  The
  finite differences resolve a gradient only to the noise of the Newton
  fixed point, which the test measures (L over successive Newton
  iterates) and states as the absolute part of that bar.

The reference's J_y treats the diffusion coefficients' dependence on ysum
as constant (jax_step.py:150-220); the traces here stay below 1e-12 of the
background, and T moves no number density, so that approximation is below
the finite differences' resolution.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

H = 1e-6
LEAVES_T = ("Tco", "Ti", "Hpi", "g", "dzi", "M")


def _loss_np(y, sp):
    ymix = y / y.sum(axis=1, keepdims=True)
    return float(np.log(ymix[:, sp]).sum())


def _dloss_np(y, sp):
    s = y.sum(axis=1, keepdims=True)
    g = np.broadcast_to(-len(sp) / s, y.shape).copy()
    g[:, sp] += 1.0 / y[:, sp]
    return g


def _theta(pb, kind, i, d):
    logm, T, logbf = np.zeros(pb.nq), pb.T0.copy(), pb.logbf0.copy()
    if kind == "m":
        logm[i] += d
    elif kind == "T":
        T[i] *= np.exp(d)
    else:
        logbf[i] += d
    return pb.leaves(T, logm, logbf)


def _per_reaction(k, cot_k):
    """sum over the two rows of each reaction and over layers of k * cot_k (nq,)."""
    return (k[1::2] * cot_k[1::2] + k[2::2] * cot_k[2::2]).sum(axis=1)


def _rate_part_T(pb, k, cot_k, z):
    """The k(T) part of dL/dlog T_z: d k / d log T_z = k (b + E/T) on layer z."""
    s = pb.dlogk_dlogT(pb.T0)[:, z]
    return float(np.sum((k[1::2, z] * cot_k[1::2, z] + k[2::2, z] * cot_k[2::2, z]) * s))


@pytest.fixture(scope="module")
def solved():
    from vulcan_amd import steady_state_grad as G
    from vulcan_amd import synth
    pb = synth.config4()
    k0, atm0 = pb.leaves(pb.T0)
    print(f"config 4: nz={pb.nz} ni={pb.ni} reactions={pb.nq} rate rows={pb.net.nr + 1}",
          flush=True)
    G.set_latency_mode(True)
    try:
        t = time.perf_counter()
        ys, it, res = G.pseudo_transient_steady_state(pb.initial_state(), k0, atm0, pb.net)
        t = time.perf_counter() - t
    finally:
        G.set_latency_mode(False)
    ys = ys.cpu().numpy()
    print(f"steady state: {it} iterations, {t:.2f} s, ||f||_inf = {res:.3e} "
          f"(max y {ys.max():.3e})", flush=True)
    return pb, k0, atm0, ys


def _newton(pb, k, atm, y0, iters):
    """Newton iterates on the GPU (residual, J_y blocks, block Thomas)."""
    import torch
    from vulcan_amd import steady_state_grad as G
    inputs = G.build_steady_state_inputs(k, atm)
    y, kt, nz, ni = G._prep(y0, k, pb.net)
    rec = G._record(inputs, nz, ni)
    out = []
    for _ in range(iters):
        res = G._residual(y, kt, rec, pb.net, nz)
        diag, sup, sub = G._jacobian_blocks(y, kt, rec, pb.net, nz, ni)
        dy, _ = G._factor_solve(diag, sup, sub, -res, pb.net, nz, False)
        y = y + dy
        out.append(y)
    torch.cuda.synchronize()
    return [v.cpu().numpy() for v in out]


def test_config4_steady_state_matches_oracle_newton(solved):
    """y* is a fixed point of the oracle's Newton map too: one oracle Newton
    step from the GPU's y* moves no cell by more than the fixed point's own
    noise."""
    from oracle import ros2_np as O
    pb, k0, atm0, ys = solved
    assert np.all(np.isfinite(ys)) and np.all(ys > 0)
    f = O.total_rhs(ys, k0, atm0, pb.net)
    scale = np.abs(O.chem_term_scale(ys, atm0.M, k0, pb.net)).max()
    d, sp, sb = O.jacobian_blocks(ys, k0, atm0, pb.net)
    dy = O.block_thomas_diag_offdiag(d, sp, sb, -f)
    rel = np.abs(dy) / ys
    ymix = ys / ys.sum(axis=1, keepdims=True)
    print(f"oracle residual at y*: {np.abs(f).max():.3e} (largest gross term {scale:.3e}); "
          f"oracle Newton step from y*: worst cell {rel.max():.2e}; trace total "
          f"{(1.0 - ymix[:, pb.bg]).max():.2e} of ysum")
    assert rel.max() < 1e-8
    assert (1.0 - ymix[:, pb.bg]).max() < 1e-12


def test_config4_gradients_match_oracle_and_finite_differences(solved):
    import torch
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    pb, k0, atm0, ys = solved
    sp = pb.loss_sp
    spt = torch.as_tensor(sp, device="cuda")

    def loss_t(y):
        ymix = y / y.sum(dim=1, keepdim=True)
        return ymix[:, spt].log().sum()

    # --- GPU gradient of every parameter
    inputs = G.build_steady_state_inputs(k0, atm0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    val, gi = G.steady_state_value_and_grad(loss_t, inputs, ys, pb.net, residual_rtol=1e-9)
    gk = gi.k_arr.detach().cpu().numpy()
    t = time.perf_counter() - t
    assert abs(float(val) - _loss_np(ys, sp)) < 1e-12 * abs(_loss_np(ys, sp))
    lg = {f: np.asarray(getattr(gi, f).detach().cpu()) for f in LEAVES_T + ("bot_flux",)}
    g_m = _per_reaction(k0, gk)
    g_bf = np.exp(pb.logbf0) * lg["bot_flux"][pb.flux_sp]
    g_T = np.zeros(pb.nz)
    hT = 1e-7
    for z in range(pb.nz):
        (_, ap), (_, am) = _theta(pb, "T", z, hT), _theta(pb, "T", z, -hT)
        tot = 0.0
        for f in LEAVES_T:
            tot += np.sum(lg[f] * (np.asarray(getattr(ap, f)) - np.asarray(getattr(am, f))))
        g_T[z] = tot / (2 * hT) + _rate_part_T(pb, k0, gk, z)
    grad = {"m": g_m, "T": g_T, "b": g_bf}
    print(f"GPU gradient of L = {float(val):.6e} wrt {g_m.size} + {g_T.size} + {g_bf.size} "
          f"parameters: adjoint {t * 1e3:.1f} ms; max |dL/dlog m| {np.abs(g_m).max():.3e}, "
          f"|dL/dlog T| {np.abs(g_T).max():.3e}, |dL/dlog F| {np.abs(g_bf).max():.3e}",
          flush=True)

    # --- the oracle's restatement of the custom_vjp
    lam = O.adjoint_lambda(ys, k0, atm0, pb.net, _dloss_np(ys, sp))
    gk_o = -O.chem_vjp_k(ys, atm0.M, k0, pb.net, lam)
    g_m_o = _per_reaction(k0, gk_o)
    e_m = np.abs(g_m - g_m_o).max() / np.abs(g_m_o).max()
    print(f"all {g_m.size} rate gradients vs the oracle custom_vjp: {e_m:.2e} of the largest")
    assert e_m < 1e-9

    def g_oracle(kind, i):
        # -lambda . df/dtheta at fixed y*: the rate part through the oracle's
        # rate VJP, the atmosphere part by central differences of its residual
        (_, ap), (_, am) = _theta(pb, kind, i, H), _theta(pb, kind, i, -H)
        df = (O.total_rhs(ys, k0, ap, pb.net) - O.total_rhs(ys, k0, am, pb.net)) / (2 * H)
        g = -float(np.sum(lam * df))
        return g + (_rate_part_T(pb, k0, gk_o, i) if kind == "T" else 0.0)

    # --- 20 random parameters
    rng = np.random.default_rng(4)
    big = np.flatnonzero(np.abs(g_m) >= 1e-2 * np.abs(g_m).max())
    picks = ([("m", int(q)) for q in rng.choice(big, 8, replace=False)]
             + [("T", int(z)) for z in rng.choice(pb.nz, 8, replace=False)]
             + [("b", int(i)) for i in rng.choice(pb.flux_sp.size, 4, replace=False)])
    # the finite differences' resolution: L over successive Newton iterates at theta_0
    its = _newton(pb, k0, atm0, ys, 4)
    Ls = [_loss_np(v, sp) for v in its]
    noise = max(abs(a - b) for a, b in zip(Ls[1:], Ls[:-1]))
    atol = max(20.0 * noise / (2 * H), 1e-5)
    print(f"Newton fixed-point noise of L: {noise:.2e} -> finite differences resolve "
          f"{atol:.1e} absolute", flush=True)
    worst_o = worst_fd = 0.0
    for kind, i in picks:
        g = float(grad[kind][i])
        go = float(g_m_o[i]) if kind == "m" else g_oracle(kind, i)
        Lp = _loss_np(_newton(pb, *_theta(pb, kind, i, H), ys, 4)[-1], sp)
        Lm = _loss_np(_newton(pb, *_theta(pb, kind, i, -H), ys, 4)[-1], sp)
        gf = (Lp - Lm) / (2 * H)
        eo = abs(g - go) / max(abs(go), 1e-300)
        ef = abs(g - gf) / max(abs(gf), 1e-300)
        print(f"  {kind}{i:5d}: GPU {g:+.9e}  custom_vjp {go:+.9e} ({eo:.1e})  "
              f"FD {gf:+.9e} ({ef:.1e})", flush=True)
        assert abs(g - go) <= 1e-7 * abs(go) + 1e-6
        assert abs(g - gf) <= 1e-3 * abs(gf) + atol
        worst_o = max(worst_o, eo if abs(go) > 1e-3 else 0.0)
        worst_fd = max(worst_fd, ef if abs(gf) > 100 * atol else 0.0)
    print(f"20 parameters: worst relative error vs the custom_vjp {worst_o:.1e}, vs central "
          f"differences {worst_fd:.1e} (on gradients above 100x the resolution)")
