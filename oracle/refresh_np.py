"""CPU oracle for the in-loop atmosphere refresh — TEST INFRASTRUCTURE ONLY.

A NumPy restatement, written from the reference source read as text, of
atm_refresh.update_mu_dz_jax (atm_refresh.py:36-98): mean molecular mass,
then the hydrostatic column integrated upward and downward from the
reference layer (gravity falling off as (Rp / (Rp + z))^2), interface
averages; and atm_refresh.update_phi_esc_jax (atm_refresh.py:101-122): the
diffusion-limited escape flux at the top, floored at -max_flux.  Only
``tests/`` may import this module, and only to check against it.
Pins: tests/test_oracle_refresh.py.
"""
from __future__ import annotations

import numpy as np


def update_mu_dz(ymix, Tco, pico, mol_mass, pref_indx, zco_pref, gs, Rp, kb, Navo):
    """(mu, g, Hp, dz, zco, dzi, Hpi) from the mixing ratios."""
    ymix = np.asarray(ymix, np.float64)
    nz = ymix.shape[0]
    mu = ymix @ np.asarray(mol_mass, np.float64)
    g = np.zeros(nz)
    Hp = np.zeros(nz)
    dz = np.zeros(nz)
    zco = np.zeros(nz + 1)
    zco[pref_indx] = zco_pref
    z = float(zco_pref)
    for i in range(pref_indx, nz):           # upward (atm_refresh.py:56-64)
        g[i] = gs if i == pref_indx else gs * (Rp / (Rp + z)) ** 2
        Hp[i] = kb * Tco[i] / (mu[i] / Navo * g[i])
        dz[i] = Hp[i] * np.log(pico[i] / pico[i + 1])
        z = z + dz[i]
        zco[i + 1] = z
    z = float(zco_pref)
    for i in range(pref_indx - 1, -1, -1):   # downward (atm_refresh.py:72-78)
        g[i] = gs * (Rp / (Rp + z)) ** 2
        Hp[i] = kb * Tco[i] / (mu[i] / Navo * g[i])
        dz[i] = Hp[i] * np.log(pico[i] / pico[i + 1])
        z = z - dz[i]
        zco[i] = z
    dzi = 0.5 * (dz[:-1] + dz[1:])
    Hpi = 0.5 * (Hp[:-1] + Hp[1:])
    return mu, g, Hp, dz, zco, dzi, Hpi


def update_phi_esc(y, g, Hp, top_flux_in, Tco, Dzz_top, ms, diff_esc_idx, kb, Navo, max_flux):
    """top_flux with the escaping species' entries replaced."""
    out = np.array(top_flux_in, np.float64)
    idx = np.asarray(diff_esc_idx, np.int64)
    flux = -np.asarray(Dzz_top)[idx] * np.asarray(y)[-1, idx] * (
        1.0 / Hp[-1] - np.asarray(ms)[idx] * g[-1] / (Navo * kb * Tco[-1]))
    out[idx] = np.maximum(flux, -max_flux)
    return out
