"""GPU: a real VULCAN network through the device path.

The reference's HD 189733b network (thermo/NCHO_photo_network.txt: 69
species, 439 reactions in 878 slots, three-body, Lindemann and Troe rows,
48 photolysis branches) is parsed from its text file, its rates are built
on the device from the T and M of an HD 189733b-like column with the
reference's NASA-9 fits, and the chemistry, the Ros2 step and element
conservation are checked against the oracle:

* k against the oracle's build_rate_array (rtol 1e-11, same zero pattern);
* chem_rhs and its Jacobian against the oracle (1e-13 of each layer's
  scale, as tight as the synthetic workloads allow);
* one Ros2 step against the NumPy and C oracles (1e-10 on the layers the
  two oracles agree on to 1e-12, within 10x their spread elsewhere);
* without diffusion every layer's H, O, C, N and He atom content is
  conserved by the step (the Ros2 update preserves linear invariants of the
  chemistry; checked to 1e-9 of the layer's atom count).
Photolysis rows stay at k = 0 here (the rate build leaves them to the
photo module, rates.py:71-73).
"""
from __future__ import annotations

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
NET = GOLD / "networks" / "NCHO_photo_network.txt"
KB = 1.38064852e-16


def _column(nz=48, kzz=0.0, seed=3):
    from vulcan_amd import network as N, synth
    from vulcan_amd._types import AtmStatic
    from vulcan_amd.gibbs import load_nasa9
    net = N.parse_network(NET, extra_species=["CH4", "O2", "He", "N2", "H2"])
    ni = net.ni
    c9, present = load_nasa9(net.species, GOLD)
    P = np.logspace(np.log10(1e9), np.log10(1e3), nz)          # 1000 bar .. 1 mbar
    z = np.linspace(0.0, 1.0, nz)
    T = 1700.0 - 700.0 * z + 60.0 * np.sin(7.0 * z)
    geo = synth.compute_geometry(P, T)
    M = P / (KB * T)
    rng = np.random.default_rng(seed)
    mix = 10.0 ** rng.uniform(-14.0, -8.0, size=(nz, ni))
    for sp, f in (("H2", 0.84), ("He", 0.145), ("H2O", 1e-3), ("CO", 5e-4), ("CH4", 1e-4),
                  ("N2", 6e-5), ("NH3", 1e-6), ("H", 1e-6)):
        mix[:, net.species_idx[sp]] = f
    mix /= mix.sum(axis=1, keepdims=True)
    y = M[:, None] * mix
    atm = AtmStatic(
        Kzz=np.full(nz - 1, kzz), Dzz=np.zeros((nz - 1, ni)), dzi=geo["dzi"],
        vz=np.zeros(nz - 1), Hpi=geo["Hpi"], Ti=geo["Ti"], Tco=T, g=geo["g"],
        ms=np.full(ni, 2.3), alpha=np.zeros(ni), M=M, vm=np.zeros((nz, ni)),
        vs=np.zeros((nz - 1, ni)), top_flux=np.zeros(ni), bot_flux=np.zeros(ni),
        bot_vdep=np.zeros(ni), gas_indx_mask=np.ones(ni, bool))
    return net, c9, present, y, atm


def _np(t):
    return t.detach().cpu().numpy()


def _k(net, c9, atm):
    from vulcan_amd import rates
    cfg = SimpleNamespace(use_lowT_limit_rates=False, remove_list=None)
    return _np(rates.build_rate_array(cfg, net, atm, c9))


def test_rates_chemistry_and_step_on_the_hd189_network():
    from oracle import ros2_c, ros2_np as O, rates_np as R
    from vulcan_amd import chem, core, jax_step, network as N
    net, c9, present, y, atm = _column(kzz=1e8)
    assert present.sum() >= 60                     # NASA-9 fits for most species
    arrays = N.to_arrays(net)
    k = _k(net, c9, atm)
    with np.errstate(all="ignore"):
        ko = R.build_rate_array(net, atm.Tco, atm.M, c9)
    assert np.array_equal(k == 0.0, ko == 0.0)
    nzr = ko != 0.0
    err_k = np.max(np.abs(k[nzr] - ko[nzr]) / np.abs(ko[nzr]))
    assert err_k < 1e-11
    assert np.all(k[net.is_photo] == 0.0) and np.any(k[781] > 0.0)   # Troe row built
    f = _np(chem.chem_rhs(y, atm.M, k, arrays))
    fo = O.chem_rhs(y, atm.M, k, arrays)
    scale = np.abs(fo).max(axis=1, keepdims=True)
    err_f = np.max(np.abs(f - fo) / scale)
    J = _np(chem.chem_jac_analytical(y, atm.M, k, arrays))
    Jo = O.chem_jac(y, atm.M, k, arrays)
    err_J = np.max(np.abs(J - Jo) / np.abs(Jo).max(axis=(1, 2), keepdims=True))
    dt = 1e-4
    sol, _ = jax_step.jax_ros2_step(y, k, dt, atm, arrays)
    so, _ = O.ros2_step(y, k, dt, atm, arrays)
    rec = core.pack_atm_record(atm).numpy()[None]
    sc, _ = ros2_c.ros2_steps(y[None], k[None], rec, dt, arrays)
    sol = _np(sol)
    lay = lambda a, b: np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)
    e_gpu, e_floor = lay(sol, so), lay(sc[0], so)
    well = e_floor < 1e-12
    moved = np.max(np.abs(so - y) / np.abs(so).max(axis=1, keepdims=True))
    print(f"k {err_k:.2e}; chem_rhs {err_f:.2e}; J {err_J:.2e}; step moves y by {moved:.2e}; "
          f"well-determined layers {well.sum()}/{len(well)}, gpu there {e_gpu[well].max():.2e}, "
          f"worst {e_gpu.max():.2e} vs oracle spread {e_floor.max():.2e}")
    assert err_f < 1e-13 and err_J < 1e-13
    assert np.all(np.isfinite(sol)) and moved > 1e-8
    assert well.sum() >= len(well) // 2
    assert e_gpu[well].max() < 1e-10
    assert e_gpu.max() <= 10.0 * e_floor.max() + 1e-10


def test_step_conserves_elements_on_the_hd189_network():
    from vulcan_amd import jax_step, network as N
    net, c9, _, y, atm = _column(kzz=0.0, seed=4)
    atoms, elements, _ = N.load_compositions(GOLD / "all_compose.txt", net.species)
    keep = [e for e, name in enumerate(elements) if atoms[:, e].any()]
    k = _k(net, c9, atm)
    sol = y
    for _ in range(3):
        sol = _np(jax_step.jax_ros2_step(sol, k, 1e-2, atm, N.to_arrays(net))[0])
    before = y @ atoms[:, keep]
    after = sol @ atoms[:, keep]
    rel = np.abs(after - before) / before
    print(f"elements {[elements[e] for e in keep]}: worst relative change {rel.max():.2e}; "
          f"y moved by {np.max(np.abs(sol - y) / np.abs(y).max(axis=1, keepdims=True)):.2e}")
    assert np.all(np.isfinite(sol))
    assert rel.max() < 1e-9


def test_low_temperature_caps_on_the_hd189_network():
    """A cold column (120-420 K) with use_lowT_limit_rates: the device build
    equals the oracle's (rtol 1e-11), and the three cap reactions (slots 665,
    671, 673) carry the capped values where T is below each threshold
    (rates.py:149-165)."""
    from oracle import rates_np as R
    from vulcan_amd import network as N, rates
    from vulcan_amd.gibbs import load_nasa9
    net = N.parse_network(NET, extra_species=["CH4", "O2", "He", "N2", "H2"])
    c9, _ = load_nasa9(net.species, GOLD)
    nz = 40
    T = np.linspace(120.0, 420.0, nz)
    M = np.logspace(19.0, 14.0, nz)
    cfg = SimpleNamespace(use_lowT_limit_rates=True, remove_list=None)
    k = _np(rates.build_rate_array(cfg, net, SimpleNamespace(Tco=T, M=M), c9))
    with np.errstate(all="ignore"):
        ko = R.build_rate_array(net, T, M, c9, use_lowT_limit_rates=True)
    assert np.array_equal(k == 0.0, ko == 0.0)
    nzr = (ko != 0.0) & np.isfinite(ko)
    assert np.max(np.abs(k[nzr] - ko[nzr]) / np.abs(ko[nzr])) < 1e-11
    cold = T <= 300.0
    assert np.all(k[671][cold] == 3.7e-30) and np.all(k[671][~cold] != 3.7e-30)
    assert np.all(k[673][T <= 200.0] == 2.49e-27)
    kinf = 2.06e-10 * T ** -0.4
    cap = 6.0e-29 / (1.0 + 6.0e-29 * M / kinf)
    m = T <= 277.5
    assert np.max(np.abs(k[665][m] - cap[m]) / cap[m]) < 1e-13


def _to_steady_state(kzz, cfg, conv=False):
    from vulcan_amd import _lib, network as N, synth
    from vulcan_amd.outer_loop import ColumnConv, SteadyStateRunner
    net, c9, _, y, atm = _column(nz=40, kzz=kzz, seed=3)
    k = _k(net, c9, atm)
    arr = N.to_arrays(net)
    col = SimpleNamespace(nz=40, ni=net.ni, y=y, k=k, atm=atm, net=arr, dt=1e-10, photo=None)
    lib = _lib.load()
    try:
        for sl in range(1, 8):
            _lib.check(lib.vk_tune_set(0, sl, 6), "vk_tune_set")   # cyclic reduction
        P = np.logspace(np.log10(1e9), np.log10(1e3), 40)
        hp = [ColumnConv(Hp=synth.compute_geometry(P, atm.Tco)["Hp"])] if conv else None
        r = SteadyStateRunner([col], cfg, conv=hp)
        r.run(max_attempts=2 * cfg.count_max, check_every=100, use_graph=True)
    finally:
        for sl in range(1, 8):
            lib.vk_tune_set(0, sl, 0)
    return r, net, arr, k, atm


def test_real_column_reaches_thermochemical_equilibrium():
    """Kzz = 0, no photolysis: every layer of the HD 189733b-like column runs
    to a steady state on the device (device-built rates, cyclic-reduction
    factor), which must be thermochemical equilibrium, since the reverse
    rates come from the NASA-9 Gibbs energies: each reversible pair with a
    rate above 1e-6 of its layer's largest runs forward and backward at the
    same rate.  The median pair balances to 1e-8 of its gross rate; the
    hydrostatic rebalance (every accepted state rescaled to n0) keeps the
    fixed point a little off exact equilibrium, so the worst pair is held
    to 1e-2 and chem_rhs to 1e-3 of each layer's gross rate."""
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import ST_CONVERGED, RunConfig
    r, net, arr, k, atm = _to_steady_state(0.0, RunConfig(count_max=20000))
    res = r.results()
    assert res["status"][0] & ST_CONVERGED, res["status"]
    ys = res["y"][0]
    ypad = np.concatenate([ys, np.ones((40, 1))], axis=1)
    rate = k.T * np.prod(ypad[:, arr.reactant_idx] ** arr.reactant_stoich[None], axis=2)
    rate = np.where(arr.is_three_body[None, :], rate * atm.M[:, None], rate)
    fw = np.arange(1, net.stop_rev_indx, 2)
    Rf, Rr = rate[:, fw], rate[:, fw + 1]
    both = (k[fw] > 0) & (k[fw + 1] > 0)
    big = both.T & (np.maximum(Rf, Rr) > 1e-6 * rate.max(axis=1, keepdims=True))
    rel = np.abs(Rf - Rr)[big] / np.maximum(Rf, Rr)[big]
    f = O.chem_rhs(ys, atm.M, k, arr)
    resid = np.max(np.abs(f).max(axis=1) / np.abs(rate).sum(axis=1))
    print(f"accepted {res['accept'][0]}, t {res['t'][0]:.3e}; {big.sum()} pairs: net/gross "
          f"median {np.median(rel):.2e}, max {rel.max():.2e}; |f|/gross {resid:.2e}")
    assert big.sum() > 500
    assert np.median(rel) < 1e-8 and rel.max() < 1e-2 and resid < 1e-3


def test_real_column_stops_by_the_reference_rule():
    """Kzz = 1e9: the same column under the reference's stopping rule
    (longdy against a 500-state ring) stops converged, by the strict or the
    loose criterion, well inside count_max."""
    from vulcan_amd.outer_loop import ST_CONVERGED, RunConfig
    cfg = RunConfig(conv_mode="longdy", count_max=20000)
    r, *_ = _to_steady_state(1e9, cfg, conv=True)
    res = r.results()
    ld, ldt = r.conv_state()[0, :2]
    print(f"accepted {res['accept'][0]}, longdy {ld:.3e}, longdydt {ldt:.3e}")
    assert res["status"][0] & ST_CONVERGED
    assert res["accept"][0] < 10000
    assert (ld < cfg.yconv_cri and ldt < cfg.slope_cri) or (ld < cfg.yconv_min and ldt < 1e-8)
