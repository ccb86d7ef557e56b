"""Gradient of a steady-state loss w.r.t. physical parameters (the
differentiation workload) through the GPU implicit-function-theorem path.

Problem: an open column (nz = 12, 8 species) on the weak-diffusion
atmosphere of the reference's open-chain test: zeroth-order sources, a ring
of reversible first-order reactions (two of them three-body), first-order
sinks, Arrhenius rates A (T/T_ref)^b exp(-E/T), eddy and molecular
diffusion, and bottom fluxes of 4 species (log-uniform 1e5..1e12) with
deposition.  Chemistry dominates the weak diffusion, so Newton with the
reference's approximate Jacobian converges; it runs on the GPU.

Loss L = sum_z sum_{5 species} log ymix.  Parameters: one log multiplier per
reaction (shared by its forward and reverse rows), T(z) (entering k(T),
M = T_ref / T, Tco, Ti and Hpi), and log bottom flux.
dL/dtheta = sum over leaves of (GPU cotangent) . (d leaf / d theta), the
leaf tangents from the host-side input generator.  Checked against central
differences with h = 1e-6 in log space on 20 parameters (Newton re-solved on
each side) and, for the rate part, against the oracle's restatement of the
reference custom_vjp (steady_state_grad.py:306-327).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

NZ, NT, NI = 12, 8, 9                              # 8 trace species + a background gas
BG = 8
RING = [(i, (i + 1) % NT) for i in range(NT)]       # reversible i <-> i+1
SOURCES = [0, 3, 6, BG]                            # 0 -> s (zeroth order)
SINKS = [2, 5, 7, BG]                              # s -> 0 (first order)
NQ = len(RING) + len(SOURCES) + len(SINKS)         # reactions; rows 2q+1 (fwd), 2q+2 (rev)
SPECIES = [0, 2, 3, 5, 7]
T_REF = 1500.0
N0 = 1e20                                          # background density scale
# The reference's J_y treats the diffusion coefficients' dependence on the
# column density ysum as constant (approximate by design, jax_step.py:150-220),
# so its implicit gradient differs from central differences wherever that
# dependence matters -- e.g. T moves the background density and so ysum.
# Central differences are therefore taken without diffusion (Kzz = 0: every
# layer is independent and J_y is exact); with diffusion on, the GPU gradient
# is compared with the reference's definition, the implicit gradient built
# from J_y, restated by the oracle.  A background gas produced at k_d * M and
# lost at k_d keeps ysum ~ M ~ 1e20 >> the traces.
KZZ_ON = 1e-2
K_BG = 1.0


def _network():
    from vulcan_amd._types import NetworkArrays
    nr = 2 * NQ
    r_idx = np.full((nr + 1, 1), NI, np.int64)
    p_idx = np.full((nr + 1, 1), NI, np.int64)
    r_st = np.zeros((nr + 1, 1))
    p_st = np.zeros((nr + 1, 1))
    is3 = np.zeros(nr + 1, bool)

    def put(row, reac, prod):
        if reac is not None:
            r_idx[row, 0], r_st[row, 0] = reac, 1.0
        if prod is not None:
            p_idx[row, 0], p_st[row, 0] = prod, 1.0

    q = 0
    for i, j in RING:
        put(2 * q + 1, i, j)
        put(2 * q + 2, j, i)
        q += 1
    for s in SOURCES:  # forward 0 -> s; the reverse row (s -> 0) has rate 0
        put(2 * q + 1, None, s)
        put(2 * q + 2, s, None)
        q += 1
    for s in SINKS:    # forward s -> 0; the reverse row (0 -> s) has rate 0
        put(2 * q + 1, s, None)
        put(2 * q + 2, None, s)
        q += 1
    is3[[1, 2, 7, 8]] = True  # ring pairs 0 and 3 react with a third body (rate k M y)
    return NetworkArrays(NI, nr, r_idx, p_idx, r_st, p_st, is3)


class Problem:
    """Open column on the atmosphere of the reference's open-chain test
    (test_steady_state_grad.py:96-124) with T(z) entering the rates
    (Arrhenius), the third-body / background density M = N0 T_ref / T, Tco,
    Ti and the scale height Hpi = Ti / T_ref."""

    def __init__(self, seed=21, kzz=0.0):
        rng = np.random.default_rng(seed)
        self.kzz = float(kzz)
        self.net = _network()
        z = np.linspace(0.0, 1.0, NZ)
        self.T0 = 1000.0 + 400.0 * z + 80.0 * np.sin(3.0 * z + rng.uniform(0, 1))
        self.A = 10.0 ** rng.uniform(-0.5, 0.5, NQ)
        self.A[[0, 3]] /= N0  # three-body pairs: k M with M ~ N0
        self.b = rng.uniform(-1.0, 1.0, NQ)
        self.E = rng.uniform(0.0, 1500.0, NQ)
        self.logK = rng.uniform(-0.5, 0.5, NQ)
        self.flux_sp = np.array([1, 2, 4, 6])
        self.logbf0 = np.log(10.0 ** rng.uniform(5.0, 12.0, 4))
        self.vdep = np.full(NI, 0.1)
        self.vdep[BG] = 0.0
        self.logm0 = np.zeros(NQ)

    def leaves(self, T, logm, logbf):
        """(k, atm) for temperatures T, rate multipliers exp(logm), fluxes exp(logbf)."""
        from vulcan_amd._types import AtmStatic
        M = N0 * T_REF / T
        k = np.zeros((2 * NQ + 1, NZ))
        nring = len(RING)
        for q in range(NQ):
            kf = np.exp(logm[q]) * self.A[q] * (T / T_REF) ** self.b[q] * np.exp(-self.E[q] / T)
            k[2 * q + 1] = kf
            k[2 * q + 2] = kf / 10.0 ** self.logK[q] if q < nring else 0.0
        q_src = nring + SOURCES.index(BG)
        q_snk = nring + len(SOURCES) + SINKS.index(BG)
        k[2 * q_src + 1] = np.exp(logm[q_src]) * K_BG * M  # background made at k_d M ...
        k[2 * q_snk + 1] = np.exp(logm[q_snk]) * K_BG       # ... and lost at k_d
        Ti = 0.5 * (T[1:] + T[:-1])
        bf = np.zeros(NI)
        bf[self.flux_sp] = np.exp(logbf)
        atm = AtmStatic(
            Kzz=np.full(NZ - 1, self.kzz), Dzz=np.full((NZ - 1, NI), 0.1 * self.kzz), dzi=np.ones(NZ - 1),
            vz=np.zeros(NZ - 1), Hpi=Ti / T_REF, Ti=Ti, Tco=T, g=np.ones(NZ), ms=np.ones(NI),
            alpha=np.zeros(NI), M=M, vm=np.zeros((NZ, NI)), vs=np.zeros((NZ - 1, NI)),
            top_flux=np.zeros(NI), bot_flux=bf, bot_vdep=self.vdep,
            gas_indx_mask=np.ones(NI, bool), use_botflux=True)
        return k, atm


def _loss_t(y):
    ymix = y / y.sum(dim=1, keepdim=True)
    return ymix[:, SPECIES].log().sum()


def _loss_np(y):
    ymix = y / y.sum(axis=1, keepdims=True)
    return float(np.log(ymix[:, SPECIES]).sum())


def _solve(pb, k, atm, y0, iters=12):
    """A fixed number of Newton steps on the GPU (residual, J_y blocks, block
    Thomas): a residual tolerance would stop before a 1e-6 perturbation of a
    trace species registers."""
    import torch
    from vulcan_amd import steady_state_grad as G
    inputs = G.build_steady_state_inputs(k, atm)
    y, kt, nz, ni = G._prep(np.asarray(y0, dtype=np.float64), k, pb.net)
    rec = G._record(inputs, nz, ni)
    y = y.clone()
    for _ in range(iters):
        res = G._residual(y, kt, rec, pb.net, nz)
        diag, sup, sub = G._jacobian_blocks(y, kt, rec, pb.net, nz, ni)
        dy, _ = G._factor_solve(diag, sup, sub, -res, pb.net, nz, False)
        y = y + dy
    torch.cuda.synchronize()
    return y.detach().cpu().numpy()


def _gpu_gradients(pb):
    """y*, and dL/d(log m), dL/d(log T), dL/d(log bf) from the GPU adjoint."""
    from vulcan_amd import steady_state_grad as G
    k0, atm0 = pb.leaves(pb.T0, pb.logm0, pb.logbf0)
    y_init = np.full((NZ, NI), 1.0)
    y_init[:, BG] = N0
    ys = _solve(pb, k0, atm0, y_init, 30)
    assert np.all(ys > 0)
    inputs = G.build_steady_state_inputs(k0, atm0)
    val, gi = G.steady_state_value_and_grad(_loss_t, inputs, ys, pb.net, residual_rtol=1e-9)
    assert abs(float(val) - _loss_np(ys)) < 1e-12 * abs(_loss_np(ys))
    gk = gi.k_arr.detach().cpu().numpy()
    lg = {f: np.asarray(getattr(gi, f).detach().cpu()) for f in ("M", "Tco", "Ti", "Hpi", "bot_flux")}
    g_logm = np.array([np.sum(k0[2 * q + 1] * gk[2 * q + 1] + k0[2 * q + 2] * gk[2 * q + 2])
                       for q in range(NQ)])
    g_logbf = np.exp(pb.logbf0) * lg["bot_flux"][pb.flux_sp]
    g_logT = np.zeros(NZ)
    for z in range(NZ):  # chain rule through the leaf generator
        hT = 1e-6
        Tp, Tm = pb.T0.copy(), pb.T0.copy()
        Tp[z] *= np.exp(hT)
        Tm[z] *= np.exp(-hT)
        (kp, ap), (km, am) = pb.leaves(Tp, pb.logm0, pb.logbf0), pb.leaves(Tm, pb.logm0, pb.logbf0)
        tot = np.sum(gk * (kp - km))
        for f in ("M", "Tco", "Ti", "Hpi"):
            tot += np.sum(lg[f] * (np.asarray(getattr(ap, f)) - np.asarray(getattr(am, f))))
        g_logT[z] = tot / (2 * hT)
    return ys, k0, atm0, gk, {"m": g_logm, "T": g_logT, "bf": g_logbf}


def _perturbed(pb, kind, i, d):
    logm, T, logbf = pb.logm0.copy(), pb.T0.copy(), pb.logbf0.copy()
    if kind == "m":
        logm[i] += d
    elif kind == "T":
        T[i] *= np.exp(d)
    else:
        logbf[i] += d
    return pb.leaves(T, logm, logbf)


def _params():
    return ([("m", q) for q in range(NQ)] + [("T", z) for z in range(NZ)] +
            [("bf", i) for i in range(4)])


def test_parameter_gradients_vs_central_differences():
    """No diffusion: exact J_y, so the implicit gradient is the true one."""
    import torch
    from oracle import ros2_np as O
    pb = Problem(kzz=0.0)
    ys, k0, atm0, gk, grads = _gpu_gradients(pb)
    # the rate part against the oracle's restatement of the reference custom_vjp
    yt = torch.tensor(ys, dtype=torch.float64, requires_grad=True)
    v = torch.autograd.grad(_loss_t(yt), yt)[0].numpy()
    go, _ = O.adjoint_grad_k(ys, k0, atm0, pb.net, v)
    assert np.max(np.abs(gk - go)) <= 1e-8 * np.abs(go).max()
    rng = np.random.default_rng(0)
    params = _params()
    pick = [params[i] for i in rng.choice(len(params), 20, replace=False)]
    h = 1e-6
    scale = max(np.abs(g).max() for g in grads.values())
    worst = 0.0
    for kind, i in pick:
        vals = [_loss_np(_solve(pb, *_perturbed(pb, kind, i, sgn * h), ys)) for sgn in (1.0, -1.0)]
        fd = (vals[0] - vals[1]) / (2 * h)
        g = grads[kind][i]
        err = abs(g - fd) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        print(f"{kind}[{i}]: adjoint {g:+.6e}  FD {fd:+.6e}  rel {err:.2e}")
    print(f"worst relative error over 20 parameters: {worst:.2e}")
    assert worst < 1e-4


def test_parameter_gradients_with_diffusion_vs_reference_definition():
    """Diffusion on: every parameter against the oracle's implicit gradient,
    -lambda^T df/dtheta with lambda = J_y^{-T} dL/dy* (steady_state_grad.py:
    306-327), df/dtheta by central differences of the oracle residual."""
    import torch
    from oracle import ros2_np as O
    pb = Problem(kzz=KZZ_ON)
    ys, k0, atm0, gk, grads = _gpu_gradients(pb)
    yt = torch.tensor(ys, dtype=torch.float64, requires_grad=True)
    v = torch.autograd.grad(_loss_t(yt), yt)[0].numpy()
    lam = O.adjoint_lambda(ys, k0, atm0, pb.net, v)
    h = 1e-6
    worst = 0.0
    scale = max(np.abs(g).max() for g in grads.values())
    for kind, i in _params():
        fp = O.total_rhs(ys, *_perturbed(pb, kind, i, h), pb.net)
        fm = O.total_rhs(ys, *_perturbed(pb, kind, i, -h), pb.net)
        ref = -float(np.sum(lam * (fp - fm))) / (2 * h)
        g = grads[kind][i]
        err = abs(g - ref) / max(abs(ref), 1e-3 * scale)
        worst = max(worst, err)
    print(f"worst relative difference to the reference definition over {len(_params())} "
          f"parameters: {worst:.2e}")
    assert worst < 1e-6
