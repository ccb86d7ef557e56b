"""GPU photolysis: temperature-dependent cross sections and the MFMA J contraction.

* T-dependent absorbers in tau and w0 (photo.py:107-111, 156-160) and
  T-dependent branches in J (photo.py:334-347) against the oracle, whose T
  paths are pinned by reduction to its plain paths (test_oracle_pins.py);
* J = cross_J . aflux^T on FP64 MFMA against the oracle at the steady-state
  workload's shape (150 layers, 2500 bins, 60 branches) and at sizes that
  exercise every edge of the tiling (70 branches: two 64-branch groups;
  nz = 37; nbin = 1001 and din12 = 333, neither a multiple of 4);
* the in-loop update with T-dependent absorbers and branches: the k rows
  after 40 attempts equal the oracle's.

Bars: tau 1e-13; flux as test_gpu_photo.py (5e-9 elementwise); J 1e-12 (the
MFMA sums run in another order than the oracle's BLAS product).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _np(t):
    return t.detach().cpu().numpy()


def _t_col(seed=51, nz=30, ni=20, n_rev=50, n_photo=8, nbin=300):
    from vulcan_amd import synth
    c = synth.make_column(seed, nz=nz, ni=ni, n_rev=n_rev, n_photo=n_photo, nbin=nbin,
                          dt=1e-10, name="pht", p_bot=1e5, p_top=1e-2, consistent=True)
    return synth.with_t_dependent_photo(c, seed + 1, n_abs_T=2, n_br_T=3)


def _pd(c):
    from vulcan_amd.photo import PhotoData
    ph = c.photo
    return PhotoData(ph.absp_idx, ph.absp_cross, ph.absp_T_idx, c.photo_T["absp_T_cross"],
                     ph.scat_idx, ph.scat_cross)


def test_tau_with_T_dependent_absorbers():
    from oracle import ros2_np as O
    from vulcan_amd import photo
    c = _t_col()
    ph = c.photo
    tau = _np(photo.compute_tau_jax(c.y, c.dz, _pd(c)))
    ref = O.compute_tau(c.y, c.dz, ph.absp_idx, ph.absp_cross, ph.scat_idx, ph.scat_cross,
                        absp_T_idx=ph.absp_T_idx, absp_T_cross=c.photo_T["absp_T_cross"])
    plain = O.compute_tau(c.y, c.dz, ph.absp_idx, ph.absp_cross, ph.scat_idx, ph.scat_cross)
    e = np.max(np.abs(tau - ref) / np.maximum(np.abs(ref), 1e-300))
    print(f"tau with T-dependent absorbers: {e:.2e} (they change tau by "
          f"{np.max(np.abs(ref - plain) / np.maximum(ref, 1e-300)):.2e})")
    assert e < 1e-13


@pytest.mark.parametrize("ag0", [0.0, 0.3])
def test_flux_with_T_dependent_absorbers(ag0):
    from oracle import ros2_np as O
    from vulcan_amd import photo
    c = _t_col()
    ph = c.photo
    ymix = c.y / c.y.sum(1, keepdims=True)
    kw = dict(absp_T_idx=ph.absp_T_idx, absp_T_cross=c.photo_T["absp_T_cross"])
    tau = O.compute_tau(c.y, c.dz, ph.absp_idx, ph.absp_cross, ph.scat_idx, ph.scat_cross, **kw)
    # the in-loop lag: a first call from zero, a second fed the first's dflux_u.
    # aflux and sflux (what photolysis uses): 5e-9 per element.  The diffuse
    # fluxes of the optically thin top layers are small differences: the
    # device forms w0 from the layer's number densities, the reference from
    # its mixing ratios, the two agree to rounding, and the oracle run both
    # ways moves dflux_d there by 2e-7 of its value (1e-4 of its maximum);
    # the diffuse fluxes are held to 10x the largest such spread of the call.
    prev = np.zeros((c.nz + 1, ph.bins.shape[0]))
    args = (ph.absp_idx, ph.absp_cross, ph.scat_idx, ph.scat_cross, ph.bins, ph.mu_zenith,
            ph.edd, ag0, ph.hc)
    for call in range(2):
        got = photo.compute_flux_jax(tau, ph.sflux_top, ymix, _pd(c), ph.bins, ph.mu_zenith,
                                     ph.edd, ag0, ph.hc, prev, ag0_is_zero=(ag0 == 0.0))
        ref = O.compute_flux(tau, ph.sflux_top, ymix, *args, prev, **kw)
        alt = O.compute_flux(tau, ph.sflux_top, c.y, *args, prev, **kw)
        errs, spreads = {}, {}
        for name, g, r, a in zip(("aflux", "sflux", "dflux_d", "dflux_u"), got, ref, alt):
            scale = np.maximum(np.abs(r), 1e-300 + 1e-12 * np.abs(r).max())
            errs[name] = float(np.max(np.abs(_np(g) - r) / scale))
            spreads[name] = float(np.max(np.abs(a - r) / scale))
            print(f"call {call} {name}: elementwise {errs[name]:.3e} "
                  f"(oracle w0 spread {spreads[name]:.3e})")
        assert errs["aflux"] < 5e-9 and errs["sflux"] < 5e-9
        diffuse_bar = max(5e-9, 10.0 * max(spreads["dflux_d"], spreads["dflux_u"]))
        assert errs["dflux_d"] < diffuse_bar and errs["dflux_u"] < diffuse_bar
        prev = ref[3]


@pytest.mark.parametrize("nz,nbin,n_br,d12", [(150, 2500, 60, 2500 // 2), (37, 1001, 70, 333),
                                              (16, 8, 1, 4)])
def test_J_mfma_matches_oracle(nz, nbin, n_br, d12):
    from oracle import ros2_np as O
    from vulcan_amd import photo
    rng = np.random.default_rng(nz + nbin)
    af = rng.uniform(0.0, 1e14, (nz, nbin))
    cross = 10.0 ** rng.uniform(-24, -17, (n_br, nbin))
    J, JT = photo.compute_J_jax_flat(af, cross, np.zeros((0, nz, nbin)), d12, 0.276, 0.5)
    J = _np(J)
    ref = O.compute_J(af, cross, d12, 0.276, 0.5)
    e = float(np.max(np.abs(J - ref) / np.abs(ref)))
    print(f"J nz={nz} nbin={nbin} n_br={n_br}: {e:.2e}")
    assert J.shape == (n_br, nz) and JT.shape == (0, nz)
    assert e < 1e-12


def test_J_of_T_dependent_branches():
    from oracle import ros2_np as O
    from vulcan_amd import photo
    rng = np.random.default_rng(9)
    nz, nbin = 23, 517
    af = rng.uniform(0.0, 1e14, (nz, nbin))
    cross = 10.0 ** rng.uniform(-24, -17, (5, nbin))
    crossT = 10.0 ** rng.uniform(-24, -17, (4, nz, nbin))
    J, JT = photo.compute_J_jax_flat(af, cross, crossT, 201, 0.3, 1.2)
    eJ = float(np.max(np.abs(_np(J) - O.compute_J(af, cross, 201, 0.3, 1.2)) /
                      np.abs(O.compute_J(af, cross, 201, 0.3, 1.2))))
    refT = O.compute_J_T(af, crossT, 201, 0.3, 1.2)
    eT = float(np.max(np.abs(_np(JT) - refT) / np.abs(refT)))
    print(f"J {eJ:.2e}, J_T {eT:.2e}")
    assert eJ < 1e-12 and eT < 1e-12


def test_in_loop_photolysis_with_T_dependent_cross_sections():
    """40 attempts with the photolysis refreshed in the loop (T-dependent
    absorbers and branches on): the same dt trajectory as the oracle, and
    the photolysis rows of k (plain and T-dependent branches) within 1e-9."""
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = _t_col(seed=61, nz=20, ni=14, n_rev=40, n_photo=5, nbin=200)
    pst = O.photo_state(c.photo, c.dz, c.nz, tcross=c.photo_T)
    tr = []
    yo, info = O.steady_state_run(c.y, c.k.copy(), c.atm, c.net, c.dt, photo=pst,
                                  max_attempts=40, trace=tr)
    r = SteadyStateRunner([c], RunConfig())
    dts = []
    for _ in range(40):
        dts.append(float(r.dt.item()))
        r.attempt()
    dto = np.array([x["dt"] for x in tr])
    assert np.max(np.abs(np.array(dts) - dto) / dto) < 1e-8
    kd = r.k.view(-1).cpu().numpy().reshape(c.k.shape)
    rows = np.concatenate([c.photo.branch_re_idx[c.photo.branch_active],
                           c.photo.branch_T_re_idx])
    ko = info["k"][rows]
    e = float(np.max(np.abs(kd[rows] - ko) / np.abs(ko).max(axis=1, keepdims=True)))
    print(f"photolysis k rows after 40 attempts: {e:.2e}")
    assert e < 1e-9


def test_photo_cadence_switch_matches_oracle():
    """The ini -> final photolysis cadence (outer_loop.py:716-724,
    1040-1049): photolysis every 4th accepted step until the carried
    relative change falls below the longdy threshold (1.0 here, the
    reference's yconv_min * 10, with a longdydt threshold that always holds),
    then every 2nd.  60 attempts against the oracle's restatement: the same
    dt trajectory and photolysis k rows, and both end on the final cadence."""
    from oracle import ros2_np as O
    from vulcan_amd import synth
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = synth.make_column(71, nz=20, ni=14, n_rev=40, n_photo=5, nbin=200, dt=1e-10,
                          name="cad", p_bot=1e5, p_top=1e-2, consistent=True)
    rp = O.RunParams(photo_frq=4, photo_frq_final=2, photo_sw_longdy=1.0,
                     photo_sw_longdydt=1e30)
    tr = []
    _, info = O.steady_state_run(c.y, c.k.copy(), c.atm, c.net, c.dt, rp=rp,
                                 photo=O.photo_state(c.photo, c.dz, c.nz), max_attempts=60,
                                 trace=tr)
    r = SteadyStateRunner([c], RunConfig(photo_frq=4, photo_frq_final=2, photo_switch_longdy=1.0,
                                         photo_switch_longdydt=1e30))
    dts = []
    for _ in range(60):
        dts.append(float(r.dt.item()))
        r.attempt()
    dto = np.array([x["dt"] for x in tr])
    print(f"oracle: {info['photo_updates']} photolysis updates, final cadence "
          f"{info['photo_frq']}; device cadence {int(r.photo_cad.item())}")
    assert info["photo_frq"] == 2 and int(r.photo_cad.item()) == 2
    assert np.max(np.abs(np.array(dts) - dto) / dto) < 1e-8
    kd = r.k.view(-1).cpu().numpy().reshape(c.k.shape)
    rows = c.photo.branch_re_idx
    ko = info["k"][rows]
    assert np.max(np.abs(kd[rows] - ko) / np.abs(ko).max(axis=1, keepdims=True)) < 1e-9
