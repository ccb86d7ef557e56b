"""GPU: the steady-state gradient on the HD 189733b network, open at the bottom.

Column: the reference's HD 189733b network (thermo/NCHO_photo_network.txt,
69 species, 878 rate rows, photolysis rows at k = 0) on the nz = 40
HD 189733b-like column of ``test_gpu_real_network._column`` (1000 bar to
1 mbar, 1700 -> 1000 K, Kzz = 1e8), rates built on the device from the
reference's NASA-9 fits.  Open bottom: every species deposits at 1 cm/s,
and H2, He, H2O, CO, CH4, N2 and NH3 get bottom fluxes equal to their
deposition from the starting state (steady_state_grad.py:153-175 puts the
deposition on J_y's bottom block).

Forward: pseudo-transient continuation on the GPU (cyclic-reduction
factor, densities held at no less than 1/100 of their value per step) to a
state that passes the reference's residual gate, ||f||_inf <= 1e-6 max y*
(steady_state_grad.py:178-196), checked with the oracle's residual too.

Adjoint: lambda = J_y^{-T} dL/dy* for L = sum over layers of log ymix of
CH4, NH3, HCN, CO2 and C2H2 (steady_state_grad.py:314-317).  This J_y is
nearly singular: cond(Y^-1 J_y Y) = 2.8e22 (chemistry at 1700 K and
1000 bar relaxes in ~1e-10 s, the column's element content through the
bottom in ~1e10 s).  So:

* the GPU's lambda is checked by its backward error, ||J^T lambda - v|| /
  (||J^T|| ||lambda|| + ||v||) <= 1e-12 (measured 3.4e-15; the oracle's LU
  solves reach 1e-23 and 4e-22);
* dL/dlog k per rate row is not determined in FP64 by any of them: the
  oracle's LU block Thomas and one dense LAPACK solve of the same J_y give
  rows that differ by 91 % in the median (of the 80 rows above 1e-2 of the
  largest), and the GPU's differ from the dense solve's by as much, no
  more: the reference's gradient on a column like this is ill-posed, not
  merely approximate.

The reference's J_y also holds the diffusion coefficients' dependence on
ysum fixed: a finite-difference Jacobian of f at y* differs from it by 1.5x
(H2), 8x (He) and 0.5x (H2O) of those rows' largest entries (the median row
agrees to 5e-8), so a Newton iteration on it stalls near 3e-3 on the major
species.  The synthetic open column (test_gpu_config4.py) keeps its traces
below 1e-12 of the background and checks the same machinery against the
oracle's custom_vjp and finite differences.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

LOSS_SP = ("CH4", "NH3", "HCN", "CO2", "C2H2")
FLUX_SP = ("H2", "He", "H2O", "CO", "CH4", "N2", "NH3")


def _open_column():
    from test_gpu_real_network import _column, _k
    from vulcan_amd import network as N
    net, c9, _, y0, atm = _column(nz=40, kzz=1e8, seed=3)
    k = _k(net, c9, atm)
    vdep = np.full(net.ni, 1.0)
    fsp = np.array([net.species_idx[s] for s in FLUX_SP])
    bf = np.zeros(net.ni)
    bf[fsp] = vdep[fsp] * y0[0, fsp]
    atm = atm._replace(bot_vdep=vdep, bot_flux=bf, use_botflux=True)
    return net, N.to_arrays(net), k, atm, y0, fsp


@pytest.fixture(scope="module")
def solved():
    import time
    from vulcan_amd import steady_state_grad as G
    net, arr, k, atm, y0, fsp = _open_column()
    print(f"HD 189733b open column: nz=40 ni={net.ni} rate rows={arr.nr}", flush=True)
    G.set_latency_mode(True)
    try:
        t = time.perf_counter()
        ys, it, res = G.pseudo_transient_steady_state(y0, k, atm, arr, floor=0.01,
                                                      polish=8, max_iter=3000)
        t = time.perf_counter() - t
    finally:
        G.set_latency_mode(False)
    ys = ys.cpu().numpy()
    print(f"steady state: {it} iterations, {t:.2f} s, ||f||_inf = {res:.3e}, "
          f"max y {ys.max():.3e} (gate {1e-6 * ys.max():.3e})", flush=True)
    return net, arr, k, atm, ys, res, fsp


def test_open_hd189_steady_state_passes_the_reference_gate(solved):
    from oracle import ros2_np as O
    net, arr, k, atm, ys, res, fsp = solved
    assert np.all(np.isfinite(ys)) and np.all(ys > 0)
    f = O.total_rhs(ys, k, atm, arr)
    print(f"oracle ||f||_inf at y*: {np.abs(f).max():.3e}; GPU {res:.3e}")
    assert np.abs(f).max() <= 1e-6 * ys.max()
    assert res <= 1e-6 * ys.max()


def _dense(d, sp, sb):
    nz, ni = d.shape[0], d.shape[1]
    J = np.zeros((nz * ni, nz * ni))
    e = np.arange(ni)
    for j in range(nz):
        J[j * ni:(j + 1) * ni, j * ni:(j + 1) * ni] = d[j]
        if j < nz - 1:
            J[j * ni + e, (j + 1) * ni + e] = sp[j]
            J[(j + 1) * ni + e, j * ni + e] = sb[j]
    return J


def test_open_hd189_adjoint_and_gradient(solved):
    import torch
    from oracle import ros2_np as O
    from vulcan_amd import steady_state_grad as G
    net, arr, k, atm, ys, res, fsp = solved
    nz, ni = ys.shape
    lsp = np.array([net.species_idx[s] for s in LOSS_SP])
    lspt = torch.as_tensor(lsp, device="cuda")

    def loss_t(y):
        ymix = y / y.sum(dim=1, keepdim=True)
        return ymix[:, lspt].log().sum()

    s = ys.sum(axis=1, keepdims=True)
    v = np.broadcast_to(-len(lsp) / s, ys.shape).copy()
    v[:, lsp] += 1.0 / ys[:, lsp]
    inputs = G.build_steady_state_inputs(k, atm)
    lam_g, singular = G.adjoint_state(ys, inputs, arr, torch.as_tensor(v, device="cuda"))
    lam_g = lam_g.cpu().numpy()
    assert not singular and np.all(np.isfinite(lam_g))
    d, sp, sb = O.jacobian_blocks(ys, k, atm, arr)
    J = _dense(d, sp, sb)
    lam_bt = O.adjoint_lambda(ys, k, atm, arr, v)
    lam_d = np.linalg.solve(J.T, v.reshape(-1)).reshape(nz, ni)
    nJ = np.abs(J).sum(axis=0).max()          # ||J^T||_inf
    eta = {}
    for name, lam in (("GPU", lam_g), ("oracle LU block Thomas", lam_bt), ("dense LAPACK", lam_d)):
        r = J.T @ lam.reshape(-1) - v.reshape(-1)
        eta[name] = np.abs(r).max() / (nJ * np.abs(lam).max() + np.abs(v).max())
    Ds = ys.reshape(-1)
    cond = np.linalg.cond(J * Ds[None, :] / Ds[:, None])

    val, gi = G.steady_state_value_and_grad(loss_t, inputs, ys, arr)
    gk = gi.k_arr.detach().cpu().numpy()
    row_g = (gk * k).sum(axis=1)
    rows = {}
    for name, lam in (("bt", lam_bt), ("dense", lam_d)):
        rows[name] = (-O.chem_vjp_k(ys, atm.M, k, arr, lam) * k).sum(axis=1)
    big = np.abs(rows["dense"]) >= 1e-2 * np.abs(rows["dense"]).max()
    spread = np.abs(rows["bt"] - rows["dense"])
    off = np.abs(row_g - rows["dense"])
    print(f"cond(Y^-1 J_y Y) = {cond:.2e}; backward error of lambda: "
          + ", ".join(f"{n} {e:.1e}" for n, e in eta.items()))
    med_spread = np.median(spread[big] / np.abs(rows["dense"][big]))
    med_off = np.median(off[big] / np.abs(rows["dense"][big]))
    print(f"dL/dlog k rows above 1e-2 of the largest: {big.sum()}; the two oracle solves differ "
          f"by median {med_spread:.2f}, GPU vs dense median {med_off:.2f}, worst GPU / spread "
          f"{np.max(off[big] / np.maximum(spread[big], 1e-300)):.2f}")
    assert eta["GPU"] <= 1e-12
    assert np.all(np.isfinite(row_g))
    assert med_off <= 2.0 * med_spread + 1e-6
