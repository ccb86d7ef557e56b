"""GPU: the atmosphere refresh (atm_refresh.py:36-122) against the oracle,
standalone and inside the attempt loop.

* update_mu_dz_jax / update_phi_esc_jax on an HD 189733b-sized column
  (nz = 60, 69 species, reference layer 12, finite planet radius, two
  escaping species) against oracle/refresh_np.py to 1e-13 relative, the
  bar of the reference's own in-loop-vs-host refresh test
  (test_outer_loop_atm_refresh.py:56, REFRESH_RTOL);
* in the loop (SteadyStateRunner, update_frq = 3, molecular diffusion and
  top fluxes on): after every attempt the host snapshots the state, and
  whenever the device flagged a column (an accepted step whose
  pre-increment count is a multiple of 3, outer_loop.py:813-820) the host
  recomputes the refresh with the oracle from the clipped step output and
  compares the atmosphere record the next step reads (g, dzi, Hpi, top
  fluxes) and the geometry to 1e-13;
* replaying the loop from one HIP graph gives the eager run's state bit
  for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KB, NAVO = 1.38064852e-16, 6.02214086e23


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _static(rng, Tco, pico, ni, ms, Dzz_top, esc, pref, Rp=8.0e9, gs=2140.0, zpref=0.0):
    from vulcan_amd.atm_refresh import AtmRefreshStatic
    return AtmRefreshStatic(Tco=Tco, pico=pico, mol_mass=rng.uniform(1.0, 50.0, ni), ms=ms,
                            Dzz_top=Dzz_top, diff_esc_idx=np.asarray(esc, np.int32),
                            pref_indx=pref, zco_pref=zpref, gs=gs, Rp=Rp, max_flux=1e13,
                            nz=len(Tco), ni=ni)


def test_standalone_refresh_matches_oracle():
    from oracle import refresh_np as O
    from vulcan_amd import atm_refresh as A, synth
    rng = np.random.default_rng(8)
    nz, ni = 60, 69
    P = np.logspace(9.0, 2.0, nz)
    pico = synth.compute_pico(P)
    Tco = 1200.0 + 300.0 * np.sin(np.linspace(0.0, 3.0, nz))
    ms = rng.uniform(1.0, 60.0, ni)
    st = _static(rng, Tco, pico, ni, ms, rng.uniform(1e8, 1e10, ni), [0, 3], 12,
                 zpref=3.0e6)
    ymix = rng.uniform(0.0, 1.0, (nz, ni)) ** 4
    ymix /= ymix.sum(axis=1, keepdims=True)
    out = [t.cpu().numpy() for t in A.update_mu_dz_jax(ymix, st)]
    ref = O.update_mu_dz(ymix, Tco, pico, st.mol_mass, 12, 3.0e6, st.gs, st.Rp, KB, NAVO)
    for name, a, b in zip(("mu", "g", "Hp", "dz", "zco", "dzi", "Hpi"), out, ref):
        assert _rel(a, b) < 1e-13, name
    y = ymix * 1e12
    top_in = -rng.uniform(1e3, 1e5, ni)
    flux = A.update_phi_esc_jax(y, ref[1], ref[2], top_in, st).cpu().numpy()
    fref = O.update_phi_esc(y, ref[1], ref[2], top_in, Tco, st.Dzz_top, ms, [0, 3], KB, NAVO,
                            1e13)
    assert _rel(flux, fref) < 1e-13
    assert flux[1] == top_in[1] and flux[0] != top_in[0]


def _columns(seed=5, n=3):
    from vulcan_amd import synth
    cols, statics = [], []
    rng = np.random.default_rng(seed)
    for c, (nz, ni) in enumerate(((24, 14), (30, 20), (20, 14))[:n]):
        col = synth.make_column(seed * 10 + c, nz=nz, ni=ni, n_rev=3 * ni, dt=1e-8,
                                p_bot=1e6, p_top=1e1, consistent=True)
        a = col.atm
        Dzz = a.Kzz[:, None] * rng.uniform(0.05, 0.5, (nz - 1, ni))
        col.atm = a._replace(Dzz=Dzz, ms=rng.uniform(2.0, 40.0, ni),
                             top_flux=np.zeros(ni), use_topflux=True)
        cols.append(col)
        statics.append(_static(rng, a.Tco, synth.compute_pico(col.P), ni, col.atm.ms,
                               Dzz[-1], [1, 2], nz // 3))
    return cols, statics


def _clip_and_mix(sol, ymix_old, cfg, gas):
    """The controller's clip (outer_loop.py:196-224) and ymix_new."""
    yc = np.where((sol < cfg.pos_cut) & (sol >= cfg.nega_cut), 0.0, sol)
    yc = np.where((ymix_old < cfg.mtol) & (yc < 0), 0.0, yc)
    ys = np.where(gas[None, :], yc, 0.0).sum(axis=1, keepdims=True)
    return yc, yc / ys


def test_in_loop_refresh_matches_oracle_and_graph_replay():
    import torch
    from oracle import refresh_np as O
    from vulcan_amd import core
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols, statics = _columns()
    cfg = RunConfig(update_frq=3, count_max=400)
    r = SteadyStateRunner(cols, cfg, refresh=statics)
    assert r.refresh is not None
    b = r.batch
    checked = 0
    for it in range(60):
        ymix_before = [t.cpu().numpy().copy() for t in b.unpack_state(r.ymix)]
        acc_before = r.accept.cpu().numpy().copy()
        r.attempt()
        torch.cuda.synchronize()
        acc = r.accept.cpu().numpy()
        sols = [t.cpu().numpy() for t in b.unpack_state(r.sol)]
        atm = r.atm.cpu().numpy()
        for c, (col, st) in enumerate(zip(cols, statics)):
            if not (acc[c] > acc_before[c] and acc_before[c] % cfg.update_frq == 0):
                continue
            nz, ni = col.nz, col.ni
            yc, mixn = _clip_and_mix(sols[c], ymix_before[c], cfg, np.ones(ni, bool))
            ref = O.update_mu_dz(mixn, st.Tco, st.pico, st.mol_mass, st.pref_indx, st.zco_pref,
                                 st.gs, st.Rp, KB, NAVO)
            geo = {k: v.cpu().numpy() for k, v in r.refresh.geometry(c).items()}
            for name, i in (("mu", 0), ("g", 1), ("Hp", 2), ("dz", 3), ("zco", 4)):
                assert _rel(geo[name], ref[i]) < 1e-13, (it, c, name)
            total, o = core.atm_record_layout(nz, ni)
            rec = atm[int(b.off_a[c]):int(b.off_a[c]) + total]
            assert _rel(rec[o[6]:o[6] + nz], ref[1]) < 1e-13          # g
            assert _rel(rec[o[1]:o[1] + nz - 1], ref[5]) < 1e-13      # dzi
            assert _rel(rec[o[3]:o[3] + nz - 1], ref[6]) < 1e-13      # Hpi
            top = O.update_phi_esc(yc, ref[1], ref[2], np.zeros(ni), st.Tco, st.Dzz_top, st.ms,
                                   st.diff_esc_idx, KB, NAVO, 1e13)
            assert _rel(rec[o[10]:o[10] + ni][[1, 2]], top[[1, 2]]) < 1e-13
            checked += 1
    print(f"refreshes checked against the oracle: {checked}")
    assert checked >= 6
    assert not r.refresh.due.any()
    # graph replay of the same attempts from the initial state
    eager = (r.y.clone(), r.atm.clone(), r.dt.clone(), r.accept.clone())
    r.reset_to_initial()
    r.capture(20)
    for _ in range(3):
        r.graph[0].replay()
    torch.cuda.synchronize()
    assert torch.equal(r.accept, eager[3])
    assert torch.equal(r.y, eager[0]) and torch.equal(r.atm, eager[1])
    assert torch.equal(r.dt, eager[2])


def test_photolysis_reads_the_refreshed_layer_thickness():
    """With photolysis and the refresh both on, the photolysis dz follows
    the refreshed geometry (the reference's photo branch reads the carried
    dz, outer_loop.py:285, 841), and the photolysis rows then differ from a
    run without the refresh."""
    import torch
    from test_gpu_photo import _photo_col
    from vulcan_amd import synth
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = _photo_col(seed=43, nz=20, ni=14, n_rev=40, n_photo=5, nbin=200)
    rng = np.random.default_rng(4)
    st = _static(rng, c.atm.Tco, synth.compute_pico(c.P), c.ni, c.atm.ms, np.zeros(c.ni), [],
                 5)
    r = SteadyStateRunner([c], RunConfig(update_frq=2, photo_frq=1), refresh=[st])
    plain = SteadyStateRunner([c], RunConfig(photo_frq=1))
    assert torch.equal(r.photo.setup.dz, plain.photo.setup.dz)
    for _ in range(12):
        r.attempt()
        plain.attempt()
    torch.cuda.synchronize()
    geo_dz = r.refresh.geometry(0)["dz"]
    assert torch.equal(r.photo.setup.dz, geo_dz)
    assert not torch.equal(r.photo.setup.dz, plain.photo.setup.dz)
    rows = torch.as_tensor(np.asarray(c.photo.branch_re_idx), device="cuda")
    kr = r.k.view(c.k.shape[0], -1)[rows]
    kp = plain.k.view(c.k.shape[0], -1)[rows]
    assert not torch.equal(kr, kp)
