"""Pins of the rate-construction oracle (oracle/rates_np.py).

The reference's own rate tests compare against a sibling code base that is
not shipped (test_rates.py:27-33, test_gibbs.py:24-31), so the oracle is
pinned here by: the known-answer cases of test_read_rate.py (the three
low-T caps, the literal remove list), limits of the falloff forms,
standard Gibbs energies of formation (JANAF tables) reproduced from the
reference's own NASA-9 data files (tests/golden/NASA9, copied unchanged by
tests/golden/copy_nasa9.py), and equilibrium constants of two reactions in
both unit conventions (dn = 0 and dn = 1).  Each is chosen so that a
dropped term, a wrong sign or a wrong branch fails it.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import rates_np as R
from vulcan_amd import synth
from vulcan_amd._types import RateNetwork

GOLD = Path(__file__).resolve().parent / "golden" / "NASA9"
SPECIES = ("H", "H2", "O", "O2", "OH", "H2O", "CO", "CO2", "CH3", "CH4")
RGAS = 8.314462618e-3     # kJ/(mol K)


def _coeffs(species=SPECIES):
    return R.load_nasa9([GOLD / f"{s}.txt" for s in species])


def _tiny_net(rows, species=SPECIES, stop_rev=None):
    """A RateNetwork from (equation, reactants, products, form) rows; each
    row is a forward reaction at 2q+1 with its reverse at 2q+2."""
    ni = len(species)
    idx = {s: j for j, s in enumerate(species)}
    n_rev = len(rows)
    nr = 2 * n_rev
    T = 3
    r_idx = np.full((nr + 1, T), ni, np.int64)
    p_idx = np.full((nr + 1, T), ni, np.int64)
    r_st = np.zeros((nr + 1, T))
    p_st = np.zeros((nr + 1, T))
    z = lambda: np.zeros(nr + 1)
    b = lambda: np.zeros(nr + 1, bool)
    a, n, E, a_inf, n_inf, E_inf = z(), z(), z(), z(), z(), z()
    fwd, kinf, special = b(), b(), b()
    Rf = {}
    for q, (eq, reac, prod, form) in enumerate(rows):
        i = 2 * q + 1
        fwd[i] = True
        Rf[i] = eq
        for s, (sp, st) in enumerate(reac.items()):
            r_idx[i, s], r_st[i, s] = idx[sp], st
            p_idx[i + 1, s], p_st[i + 1, s] = idx[sp], st
        for s, (sp, st) in enumerate(prod.items()):
            p_idx[i, s], p_st[i, s] = idx[sp], st
            r_idx[i + 1, s], r_st[i + 1, s] = idx[sp], st
        a[i], n[i], E[i] = form[:3]
        if len(form) == 6:
            kinf[i] = kinf[i + 1] = True
            a_inf[i], n_inf[i], E_inf[i] = form[3:]
    return RateNetwork(species=tuple(species), ni=ni, nr=nr, reactant_idx=r_idx,
                       product_idx=p_idx, reactant_stoich=r_st, product_stoich=p_st, a=a, n=n,
                       E=E, a_inf=a_inf, n_inf=n_inf, E_inf=E_inf, is_forward=fwd,
                       has_kinf=kinf, is_special=special, is_photo=b(), is_ion=b(),
                       is_conden=b(), is_radiative=b(),
                       stop_rev_indx=nr + 1 if stop_rev is None else stop_rev, Rf=Rf)


CAP_ROWS = [
    ("H + CH3 + M -> CH4 + M", {"H": 1, "CH3": 1}, {"CH4": 1}, (1e-10, 0.0, 0.0)),
    ("H + C2H4 + M -> C2H5 + M", {"H": 1, "H2": 1}, {"H2O": 1}, (1e-10, 0.0, 0.0)),
    ("H + C2H5 + M -> C2H6 + M", {"H": 1, "O": 1}, {"OH": 1}, (1e-10, 0.0, 0.0)),
    ("O + H2 -> OH + H", {"O": 1, "H2": 1}, {"OH": 1, "H": 1}, (8.5e-20, 2.67, 3160.0)),
]


# ---------------------------------------------------------------------------
# low-T caps and the remove list: known answers of test_read_rate.py
# ---------------------------------------------------------------------------

def test_lowT_caps_fire_all_three():
    """test_read_rate.py:71-128: at T = 150, 250, 290, 350 K the CH3 cap
    (T <= 277.5) takes the first two layers, the C2H4 cap (T <= 300) the
    first three, the C2H5 cap (T <= 200) the first; other rows untouched."""
    net = _tiny_net(CAP_ROWS)
    T = np.array([150.0, 250.0, 290.0, 350.0])
    M = np.array([1e18, 1e16, 1e14, 1e12])
    k_in = np.full((net.nr + 1, 4), 7.77e-12)
    k_in[0] = 0.0
    k = R.apply_lowT_caps(net, k_in, T, M)
    kinf = 2.06e-10 * T ** -0.4
    cap = 6.0e-29 / (1.0 + 6.0e-29 * M / kinf)
    assert np.array_equal(k[1], [cap[0], cap[1], 7.77e-12, 7.77e-12])
    assert np.array_equal(k[3], [3.7e-30, 3.7e-30, 3.7e-30, 7.77e-12])
    assert np.array_equal(k[5], [2.49e-27, 7.77e-12, 7.77e-12, 7.77e-12])
    for i in (2, 4, 6, 7, 8):
        assert np.array_equal(k[i], k_in[i])
    # the CH3 cap is the Lindemann form: its low-pressure limit is k0
    assert R.apply_lowT_caps(net, k_in, T[:1], np.array([1.0]))[1, 0] == pytest.approx(6.0e-29,
                                                                                   rel=1e-12)


def test_lowT_caps_no_op_above_thresholds():
    """test_read_rate.py:131-148."""
    net = _tiny_net(CAP_ROWS)
    k_in = np.random.RandomState(0).uniform(1e-15, 1e-10, size=(net.nr + 1, 5))
    k = R.apply_lowT_caps(net, k_in, np.full(5, 1500.0), np.full(5, 1e15))
    assert np.array_equal(k, k_in)


def test_caps_need_the_exact_equation_text():
    """The caps match the equation text exactly (rates.py:154-165): a
    padded string takes no cap."""
    rows = [(" " + CAP_ROWS[1][0], *CAP_ROWS[1][1:])]
    net = _tiny_net(rows)
    k_in = np.full((net.nr + 1, 1), 5e-12)
    assert np.array_equal(R.apply_lowT_caps(net, k_in, np.array([100.0]), np.array([1e15])),
                          k_in)


def test_remove_list_literal_and_paired():
    """test_read_rate.py:151-189: apply_remove_list zeroes exactly rows
    1, 3, 4, 6 of [1, 3, 4, 6]; fill_reverse_k's removal takes the partner
    row too (gibbs.py:141-146), but only below stop_rev_indx."""
    net = _tiny_net(CAP_ROWS, stop_rev=7)
    rng = np.random.RandomState(1)
    k_in = rng.uniform(1.0, 2.0, size=(net.nr + 1, 4))
    k_in[0] = 0.0
    k = R.apply_remove_list(net, k_in, [1, 3, 4, 6])
    for i in (1, 3, 4, 6):
        assert np.all(k[i] == 0.0)
    for i in (2, 5, 7):
        assert np.array_equal(k[i], k_in[i])
    assert np.array_equal(R.apply_remove_list(net, k_in, None), k_in)
    assert np.array_equal(R.apply_remove_list(net, k_in, []), k_in)
    K = np.full_like(k_in, 2.0)
    f = R.fill_reverse_k(net, k_in, K, remove_list=[1, 6, 7])
    assert np.all(f[1] == 0) and np.all(f[2] == 0)          # pair (1, 2) via 1
    assert np.all(f[5] == 0) and np.all(f[6] == 0)          # pair (5, 6) via 6
    assert np.array_equal(f[7], k_in[7])                    # 7 >= stop_rev: not paired
    assert np.array_equal(f[4], k_in[3] / 2.0)
    assert np.all(f[8] == 0)                                # past stop_rev, odd offset


def test_fill_reverse_k_sentinels():
    """K <= 0 (and NaN) give a zero reverse rate; rows past stop_rev_indx
    at odd offsets are zeroed, the others kept (gibbs.py:147-152)."""
    net = _tiny_net(CAP_ROWS, stop_rev=6)
    k_in = np.arange(1.0, 1.0 + (net.nr + 1) * 3).reshape(net.nr + 1, 3)
    K = np.ones_like(k_in)
    K[1] = [0.0, -1.0, np.nan]
    f = R.fill_reverse_k(net, k_in, K)
    assert np.array_equal(f[2], [0.0, 0.0, 0.0])
    assert np.array_equal(f[4], k_in[3])
    assert np.array_equal(f[6], k_in[6])      # == stop_rev: left as is
    assert np.all(f[7] == 0.0)
    assert np.array_equal(f[8], k_in[8])


# ---------------------------------------------------------------------------
# forward forms
# ---------------------------------------------------------------------------

def test_arrhenius_known_value():
    """O + H2 -> OH + H, a = 8.5e-20, n = 2.67, E = 3160 K at 1000 K:
    8.5e-20 * 1000**2.67 * exp(-3.16) (logarithms of each factor)."""
    net = _tiny_net(CAP_ROWS)
    k = R.compute_forward_k(net, np.array([1000.0]), np.array([1e18]))
    ref = np.exp(np.log(8.5e-20) + 2.67 * np.log(1000.0) - 3.16)
    assert k[7, 0] == pytest.approx(ref, rel=1e-13)
    assert np.all(k[0::2] == 0.0)


def test_lindemann_limits():
    """k0 / (1 + k0 M / kinf): k0 at low M, kinf / M at high M, k0 / 2 where
    k0 M = kinf (rates.py:115-118)."""
    rows = [("A + B -> C", {"H": 1, "O": 1}, {"OH": 1}, (3e-30, -1.5, 200.0, 2e-11, 0.3, 50.0))]
    net = _tiny_net(rows)
    T = np.full(3, 400.0)
    k0 = 3e-30 * 400.0 ** -1.5 * np.exp(-200.0 / 400.0)
    kinf = 2e-11 * 400.0 ** 0.3 * np.exp(-50.0 / 400.0)
    M = np.array([1e-10, kinf / k0, 1e40])
    k = R.compute_forward_k(net, T, M)[1]
    assert k[0] == pytest.approx(k0, rel=1e-14)
    assert k[1] == pytest.approx(k0 / 2.0, rel=1e-14)
    assert k[2] == pytest.approx(kinf / M[2], rel=1e-14)


def test_troe_oh_ch3_limits():
    """The hard-coded Troe reaction: broadening factor -> 1 in both pressure
    limits, so k -> k0 at low M and k -> kinf / M at high M; Fc at the
    centre k0 M = kinf (rates.py:41-58, Jasper 2017 parameters)."""
    T = 600.0
    k0 = 1.932e3 * T ** -9.88 * np.exp(-7544.0 / T) + 5.109e-11 * T ** -6.25 * np.exp(-1433.0 / T)
    kinf = 1.031e-10 * T ** -0.018 * np.exp(16.74 / T)
    Fc = 0.1855 * np.exp(-T / 155.8) + 0.8145 * np.exp(-T / 1675.0) + np.exp(-4531.0 / T)
    Ts = np.full(3, T)
    lo = R.troe_oh_ch3(Ts, np.array([kinf / k0 * 1e-200, kinf / k0, kinf / k0 * 1e200]))
    assert lo[0] == pytest.approx(k0, rel=1e-6)
    assert lo[1] == pytest.approx(k0 / 2.0 * Fc, rel=1e-13)
    assert lo[2] == pytest.approx(kinf / (kinf / k0 * 1e200), rel=1e-6)
    assert 0.0 < Fc < 1.0


def test_forward_row_classes():
    """Reverse, photo, ion, condensation and radiative rows stay zero; a
    special row is the Troe form when its stripped text is the OH + CH3
    reaction, Arrhenius when it has a non-zero a, else zero
    (rates.py:102-113)."""
    rng = np.random.default_rng(5)
    from vulcan_amd import synth as S
    base, _, _ = S.random_network(rng, 20, 12, 4)
    net = S.random_rate_network(rng, base, 12, 4)
    T = np.linspace(150.0, 2500.0, 7)
    M = np.logspace(10, 19, 7)
    k = R.compute_forward_k(net, T, M)
    assert np.all(k[0::2] == 0.0)
    assert np.allclose(k[1], R.troe_oh_ch3(T, M), rtol=0, atol=0)
    assert np.all(k[3] > 0.0)
    assert np.all(k[5] == 0.0)
    for i in (13, 15, 25, 27):
        assert np.all(k[i] == 0.0)


# ---------------------------------------------------------------------------
# NASA-9 Gibbs energies and equilibrium constants (JANAF values)
# ---------------------------------------------------------------------------

def _dfg(species, coeffs, T, recipe):
    g = dict(zip(species, R.gibbs_sp_vector(coeffs, np.atleast_1d(T))))
    return sum(c * g[s] for s, c in recipe.items()) * RGAS * np.atleast_1d(T)


@pytest.mark.parametrize("T,ref", [(298.15, -228.582), (1000.0, -192.590), (1500.0, -164.376),
                                   (2000.0, -135.528)])
def test_gibbs_water_formation_matches_janaf(T, ref):
    """Delta_f G(H2O, g) = g(H2O) - g(H2) - g(O2)/2 in kJ/mol against the
    JANAF table, both NASA-9 ranges (gibbs.py:41-74)."""
    c = _coeffs()
    v = _dfg(SPECIES, c, T, {"H2O": 1.0, "H2": -1.0, "O2": -0.5})[0]
    assert v == pytest.approx(ref, abs=0.1)


@pytest.mark.parametrize("T,recipe,ref,tol", [
    (298.15, {"CO2": 1, "CO": -1, "O2": -0.5}, -394.389 + 137.163, 0.1),
    (1000.0, {"CO2": 1, "CO": -1, "O2": -0.5}, -395.886 + 200.275, 0.1),
    (298.15, {"H": 1, "H2": -0.5}, 203.276, 0.1),
    (1000.0, {"H": 1, "H2": -0.5}, 165.485, 0.1),
    (298.15, {"O": 1, "O2": -0.5}, 231.736, 0.1),
    (298.15, {"CO2": 1, "H2O": 2, "CH4": -1, "O2": -2}, -394.389 - 2 * 228.582 + 50.768, 0.5),
])
def test_gibbs_reaction_energies_match_janaf(T, recipe, ref, tol):
    c = _coeffs()
    assert _dfg(SPECIES, c, T, recipe)[0] == pytest.approx(ref, abs=tol)


def test_gibbs_elements_are_minus_entropy():
    """H(298.15) = 0 for an element in its reference state, so g/RT = -S/R:
    S(H2) = 130.680, S(O2) = 205.152 J/(mol K)."""
    g = R.gibbs_sp_vector(_coeffs(("H2", "O2")), np.array([298.15]))
    assert g[0, 0] == pytest.approx(-130.680 / 8.314462618, abs=2e-3)
    assert g[1, 0] == pytest.approx(-205.152 / 8.314462618, abs=2e-3)


def test_gibbs_branch_switch_at_1000K():
    """The low range strictly below 1000 K, the high range from 1000 K on;
    the fits meet there to a few parts in 1e6."""
    c = _coeffs()
    lo = R.gibbs_sp_vector(c, np.array([1000.0 - 1e-9]))
    hi = R.gibbs_sp_vector(c, np.array([1000.0]))
    swapped = c[:, ::-1, :]
    assert np.array_equal(R.gibbs_sp_vector(swapped, np.array([999.0])),
                          R.gibbs_sp_vector(c[:, [1, 1], :], np.array([999.0])))
    assert np.allclose(lo, hi, rtol=1e-5, atol=1e-5)
    assert not np.array_equal(lo, hi)


def test_missing_nasa9_file_gives_zeros(tmp_path):
    c = R.load_nasa9([GOLD / "H2.txt", None])
    assert np.all(c[1] == 0.0) and np.any(c[0] != 0.0)


def test_keq_water_gas_shift_dn0():
    """CO + H2O -> CO2 + H2 at 1000 K: K = exp(-Delta_r G / RT) with the
    JANAF Delta_r G = -3.021 kJ/mol, no unit factor (dn = 0)."""
    net = _tiny_net([("CO + H2O -> CO2 + H2", {"CO": 1, "H2O": 1}, {"CO2": 1, "H2": 1},
                      (1e-12, 0.0, 0.0))])
    T = np.array([1000.0])
    K = R.K_eq_array(net, R.gibbs_sp_vector(_coeffs(), T), T)
    ref = np.exp(3.021 / (RGAS * 1000.0))
    assert K[1, 0] == pytest.approx(ref, rel=0.015)
    assert K[2, 0] == 1.0


def test_keq_dn1_concentration_units():
    """2 H2 + O2 -> 2 H2O at 1500 K: K_c = exp(-Delta_r G / RT) (k_B T / P0)
    with dn = 3 - 2 = 1 (gibbs.py:122-126), Delta_r G = 2 Delta_f G(H2O)."""
    net = _tiny_net([("2 H2 + O2 -> 2 H2O", {"H2": 2, "O2": 1}, {"H2O": 2},
                      (1e-12, 0.0, 0.0))])
    T = np.array([1500.0])
    K = R.K_eq_array(net, R.gibbs_sp_vector(_coeffs(), T), T)
    ref = np.exp(2 * 164.376 / (RGAS * 1500.0)) * (1.38064852e-16 * 1500.0 / 1.0e6)
    assert K[1, 0] == pytest.approx(ref, rel=0.03)


def test_keq_sentinel_rows():
    """K = 1 on reverse rows, on rows at or past stop_rev_indx - 1 and on
    photo / ion / condensation / radiative rows (gibbs.py:96-105)."""
    net = _tiny_net(CAP_ROWS, stop_rev=6)
    net = net._replace(is_radiative=net.is_radiative.copy())
    net.is_radiative[3] = True
    T = np.array([500.0, 1500.0])
    K = R.K_eq_array(net, R.gibbs_sp_vector(_coeffs(), T), T)
    assert np.all(K[0::2] == 1.0)
    assert np.all(K[3] == 1.0) and np.all(K[5] == 1.0) and np.all(K[7] == 1.0)
    assert np.all(K[1] != 1.0)


def test_build_equals_compute_all_k_without_caps():
    """build_rate_array without caps or removal equals compute_all_k
    (rates.py:185-208 vs gibbs.py:157-170), and every reverse row below
    stop_rev_indx is k_f / K."""
    rng = np.random.default_rng(9)
    base, _, _ = synth.random_network(rng, 30, 40, 6)
    net = synth.random_rate_network(rng, base, 40, 6)
    c = synth.random_nasa9(rng, 30)
    T = np.linspace(150.0, 2500.0, 9)
    M = np.logspace(9, 19, 9)
    a = R.build_rate_array(net, T, M, c)
    b = R.compute_all_k(net, T, M, c)
    assert np.array_equal(a, b)
    K = R.K_eq_array(net, R.gibbs_sp_vector(c, T), T)
    i = np.arange(2, net.stop_rev_indx, 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        assert np.array_equal(a[i], np.where(K[i - 1] > 0, a[i - 1] / K[i - 1], 0.0))
