"""GPU: the steady-state gradient with respect to T(z) when the rates are
built on the device from T: dL/dT = (dL/dk) dk/dT + (dL/dM) dM/dT + the
atmosphere leaves, with dk/dT and dk/dM from the rate build's
vector-Jacobian product (vk_rates_vjp_f64) instead of differences of a
host-side rate generator.

Problem: an open column (nz = 12, 8 trace species and a background gas)
without diffusion, so every layer is independent and the implicit gradient
is the true one.  A ring of reversible first-order reactions whose reverse
rates come from NASA-9 equilibrium constants (the reference's
build_rate_array path, rates.py:185-208), two of them three-body; zeroth-
order sources and first-order sinks past stop_rev_indx (no reverse); the
background gas made at K_BG M (a three-body source) and lost at K_BG;
M = N0 T_ref / T.  Loss L = sum_z sum_{5 species} log ymix.  Checked
against central differences in log T (h = 1e-6) of the whole chain: device
rate build, Newton solve on the device, loss.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

NZ, NT, NI = 12, 8, 9
BG = 8
RING = [(i, (i + 1) % NT) for i in range(NT)]
SOURCES = [0, 3, 6, BG]
SINKS = [2, 5, 7, BG]
NQ = len(RING) + len(SOURCES) + len(SINKS)
SPECIES = [0, 2, 3, 5, 7]
T_REF = 1500.0
N0 = 1e20
K_BG = 1.0


def _tables():
    """Reaction tables: ring pairs first (rows 1..16, reversible), then the
    sources and sinks (forward rows from 17 on, reverses held at zero)."""
    nr = 2 * NQ
    r_idx = np.full((nr + 1, 1), NI, np.int64)
    p_idx = np.full((nr + 1, 1), NI, np.int64)
    r_st = np.zeros((nr + 1, 1))
    p_st = np.zeros((nr + 1, 1))

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
    for s in SOURCES:
        put(2 * q + 1, None, s)
        put(2 * q + 2, s, None)
        q += 1
    for s in SINKS:
        put(2 * q + 1, s, None)
        put(2 * q + 2, None, s)
        q += 1
    return nr, r_idx, p_idx, r_st, p_st


class Problem:
    def __init__(self, seed=31):
        from vulcan_amd._types import NetworkArrays, RateNetwork
        rng = np.random.default_rng(seed)
        nr, r_idx, p_idx, r_st, p_st = _tables()
        q_bg_src = len(RING) + SOURCES.index(BG)
        is3 = np.zeros(nr + 1, bool)
        is3[[1, 2, 7, 8]] = True               # ring pairs 0 and 3: rate k M y
        is3[2 * q_bg_src + 1] = True           # background source: k M
        self.net = NetworkArrays(NI, nr, r_idx, p_idx, r_st, p_st, is3)
        z = np.linspace(0.0, 1.0, NZ)
        self.T0 = 1000.0 + 400.0 * z + 80.0 * np.sin(3.0 * z + rng.uniform(0, 1))
        A = 10.0 ** rng.uniform(-0.5, 0.5, NQ)
        A[[0, 3]] /= N0
        b = rng.uniform(-1.0, 1.0, NQ)
        E = rng.uniform(0.0, 1500.0, NQ)
        a = A * T_REF ** (-b)                  # A (T / T_ref)^b = a T^b
        a[q_bg_src], b[q_bg_src], E[q_bg_src] = K_BG, 0.0, 0.0
        q_bg_snk = len(RING) + len(SOURCES) + SINKS.index(BG)
        a[q_bg_snk], b[q_bg_snk], E[q_bg_snk] = K_BG, 0.0, 0.0
        S = nr + 1
        col = lambda v: np.concatenate([[0.0], np.repeat(v, 2) * np.tile([1.0, 0.0], NQ)])
        fwd = np.zeros(S, bool)
        fwd[1::2] = True
        f = lambda: np.zeros(S, bool)
        self.rnet = RateNetwork(
            species=tuple(f"S{j}" for j in range(NI)), ni=NI, nr=nr, reactant_idx=r_idx,
            product_idx=p_idx, reactant_stoich=r_st, product_stoich=p_st, a=col(a), n=col(b),
            E=col(E), a_inf=np.zeros(S), n_inf=np.zeros(S), E_inf=np.zeros(S), is_forward=fwd,
            has_kinf=f(), is_special=f(), is_photo=f(), is_ion=f(), is_conden=f(),
            is_radiative=f(), stop_rev_indx=2 * len(RING) + 1, Rf={})
        # g/RT = 2.5 (1 - ln T) + a8 / T - a9: K_eq of the ring within e^{+-2}
        c = np.zeros((NI, 2, 10))
        c[:, :, 2] = 2.5
        c[:, :, 8] = rng.uniform(-500.0, 500.0, NI)[:, None]
        c[:, :, 9] = rng.uniform(-0.5, 0.5, NI)[:, None]
        self.nasa9 = c
        self.flux_sp = np.array([1, 2, 4, 6])
        self.bf = 10.0 ** rng.uniform(5.0, 12.0, 4)
        self.vdep = np.full(NI, 0.1)
        self.vdep[BG] = 0.0

    def M_of(self, T):
        return N0 * T_REF / T

    def atm(self, T):
        from vulcan_amd._types import AtmStatic
        Ti = 0.5 * (T[1:] + T[:-1])
        bf = np.zeros(NI)
        bf[self.flux_sp] = self.bf
        return AtmStatic(
            Kzz=np.zeros(NZ - 1), Dzz=np.zeros((NZ - 1, NI)), dzi=np.ones(NZ - 1),
            vz=np.zeros(NZ - 1), Hpi=Ti / T_REF, Ti=Ti, Tco=T, g=np.ones(NZ), ms=np.ones(NI),
            alpha=np.zeros(NI), M=self.M_of(T), vm=np.zeros((NZ, NI)), vs=np.zeros((NZ - 1, NI)),
            top_flux=np.zeros(NI), bot_flux=bf, bot_vdep=self.vdep,
            gas_indx_mask=np.ones(NI, bool), use_botflux=True)

    def k(self, T):
        from types import SimpleNamespace
        from vulcan_amd import rates
        cfg = SimpleNamespace(use_lowT_limit_rates=False, remove_list=None)
        return rates.build_rate_array(cfg, self.rnet, SimpleNamespace(Tco=T, M=self.M_of(T)),
                                      self.nasa9)


def _loss_t(y):
    ymix = y / y.sum(dim=1, keepdim=True)
    return ymix[:, SPECIES].log().sum()


def _loss_np(y):
    ymix = y / y.sum(axis=1, keepdims=True)
    return float(np.log(ymix[:, SPECIES]).sum())


def _solve(pb, k, atm, y0, iters):
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


def test_T_gradient_through_device_rates_vs_central_differences():
    from vulcan_amd import rates
    from vulcan_amd import steady_state_grad as G
    pb = Problem()
    T0 = pb.T0
    k0 = pb.k(T0)
    atm0 = pb.atm(T0)
    y_init = np.full((NZ, NI), 1.0)
    y_init[:, BG] = N0
    ys = _solve(pb, k0, atm0, y_init, 30)
    assert np.all(ys > 0)
    inputs = G.build_steady_state_inputs(k0, atm0)
    val, gi = G.steady_state_value_and_grad(_loss_t, inputs, ys, pb.net, residual_rtol=1e-9)
    # chain rule: k(T, M(T)) through the device VJP, the atmosphere leaves in closed form
    d = rates.device_rates(pb.rnet, pb.nasa9, None)
    M0 = pb.M_of(T0)
    Tb, Mb = d.vjp(T0, M0, gi.k_arr)
    Tb, Mb = Tb.cpu().numpy(), Mb.cpu().numpy()
    leaf = {f: np.asarray(getattr(gi, f).detach().cpu()) for f in ("M", "Tco", "Ti", "Hpi")}
    dM = -M0 / T0
    g_T = Tb + (Mb + leaf["M"]) * dM + leaf["Tco"]
    dTi = 0.5 * (leaf["Ti"] + leaf["Hpi"] / T_REF)
    g_T[1:] += dTi
    g_T[:-1] += dTi
    g_logT = g_T * T0
    assert np.all(np.isfinite(g_logT)) and np.abs(Tb).max() > 0
    h = 1e-6
    worst = 0.0
    scale = np.abs(g_logT).max()
    for z in range(NZ):
        vals = []
        for sgn in (1.0, -1.0):
            T = T0.copy()
            T[z] *= np.exp(sgn * h)
            vals.append(_loss_np(_solve(pb, pb.k(T), pb.atm(T), ys, 12)))
        fd = (vals[0] - vals[1]) / (2 * h)
        err = abs(g_logT[z] - fd) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        print(f"T[{z}]: adjoint {g_logT[z]:+.6e}  FD {fd:+.6e}  rel {err:.2e}  "
              f"(through k: {Tb[z] * T0[z]:+.3e})")
    print(f"worst relative error over {NZ} layers: {worst:.2e}")
    assert worst < 1e-4
