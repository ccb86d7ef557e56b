"""Pins of the atmosphere-refresh oracle (oracle/refresh_np.py).

The reference's own refresh test compares its in-loop branch with a sibling
code base that is not shipped (test_outer_loop_atm_refresh.py:36-45).  The
oracle is pinned here by closed forms: an isothermal column of one gas under
constant gravity is z = H ln(p_ref / p) above and below the reference
layer; with finite Rp each layer's g is gs (Rp / (Rp + z))^2 at the
interface the reference names and the column grows taller; interface
averages; the escape flux of a hand-computed case, its floor and the
untouched species.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import refresh_np as R

KB, NAVO = 1.38064852e-16, 6.02214086e23


def _column(nz=8, T=1200.0):
    pico = np.logspace(8.0, 2.0, nz + 1)
    return np.full(nz, T), pico


def test_isothermal_constant_gravity_closed_form():
    nz, pref = 8, 3
    Tco, pico = _column(nz)
    ymix = np.zeros((nz, 2))
    ymix[:, 0] = 1.0
    mass = np.array([2.3, 44.0])
    gs, Rp, zpref = 2140.0, 1e40, 5.0e6
    mu, g, Hp, dz, zco, dzi, Hpi = R.update_mu_dz(ymix, Tco, pico, mass, pref, zpref, gs, Rp,
                                                  KB, NAVO)
    H = KB * 1200.0 / (2.3 / NAVO * gs)
    assert np.all(mu == 2.3) and np.all(g == gs)
    assert np.allclose(Hp, H, rtol=1e-15)
    ref = zpref + H * np.log(pico[pref] / pico)
    assert np.allclose(zco, ref, rtol=1e-13, atol=1e-6)
    assert zco[pref] == zpref and np.all(np.diff(zco) > 0)
    assert np.allclose(dzi, 0.5 * (dz[:-1] + dz[1:]), rtol=0) and np.allclose(Hpi, H, rtol=1e-15)


def test_inverse_square_gravity_at_the_named_interface():
    nz, pref = 10, 4
    Tco, pico = _column(nz, 1500.0)
    rng = np.random.default_rng(2)
    ymix = rng.uniform(0.1, 1.0, (nz, 3))
    ymix /= ymix.sum(axis=1, keepdims=True)
    mass = np.array([2.0, 4.0, 18.0])
    gs, Rp = 2140.0, 8.0e9
    mu, g, Hp, dz, zco, _, _ = R.update_mu_dz(ymix, Tco, pico, mass, pref, 0.0, gs, Rp, KB, NAVO)
    assert np.allclose(mu, ymix @ mass, rtol=1e-15)
    assert g[pref] == gs
    for i in range(pref + 1, nz):
        assert g[i] == gs * (Rp / (Rp + zco[i])) ** 2
    for i in range(pref):
        assert g[i] == gs * (Rp / (Rp + zco[i + 1])) ** 2
    # below the reference layer g uses the layer's upper interface, so the
    # layer just under it sits at zco_pref and keeps gs
    assert np.all(g[pref + 1:] < gs) and np.all(g[:pref - 1] > gs) and g[pref - 1] == gs
    assert np.allclose(dz, Hp * np.log(pico[:-1] / pico[1:]), rtol=1e-15)
    assert np.allclose(Hp, KB * Tco / (mu / NAVO * g), rtol=1e-15)
    # lower gravity aloft makes the column above the reference layer taller
    # than under constant g; higher gravity below makes the part beneath shallower
    _, _, _, _, zc, _, _ = R.update_mu_dz(ymix, Tco, pico, mass, pref, 0.0, gs, 1e40, KB, NAVO)
    assert zco[-1] > zc[-1] and 0.0 > zco[0] > zc[0]


def test_escape_flux_hand_case_floor_and_untouched_species():
    nz, ni = 5, 4
    Tco = np.full(nz, 1000.0)
    y = np.ones((nz, ni))
    y[-1] = [1e8, 2e9, 3e7, 5e10]
    g = np.full(nz, 900.0)
    Hp = np.full(nz, 3.0e7)
    Dzz_top = np.array([1e9, 2e9, 5e8, 1e12])
    ms = np.array([1.008, 2.016, 16.0, 1.008])
    top_in = np.array([-1.0, -2.0, -3.0, -4.0])
    out = R.update_phi_esc(y, g, Hp, top_in, Tco, Dzz_top, ms, [0, 3], KB, NAVO, 1e13)
    f0 = -1e9 * 1e8 * (1.0 / 3.0e7 - 1.008 * 900.0 / (NAVO * KB * 1000.0))
    assert out[0] == pytest.approx(f0, rel=1e-15) and out[0] < 0.0
    assert out[3] == -1e13                                  # floored at -max_flux
    assert out[1] == -2.0 and out[2] == -3.0                # not escaping: unchanged
    # a heavy species whose mass term beats 1/Hp would get a positive (inward) flux
    heavy = R.update_phi_esc(y, g, Hp, top_in, Tco, Dzz_top, np.full(ni, 400.0), [2], KB,
                             NAVO, 1e13)
    assert heavy[2] > 0.0
