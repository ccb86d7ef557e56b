"""CPU oracle for the rate construction k(T, M) — TEST INFRASTRUCTURE ONLY.

A plain NumPy restatement, written from the reference source read as text,
of the reference's host-side rate build: forward rates (modified Arrhenius,
Lindemann falloff, the hard-coded Troe reaction), the three low-T caps,
NASA-9 Gibbs energies, equilibrium constants and the reverse fill.  Only
``tests/`` may import this module, and only to check against it.

The network argument is duck-typed on the attributes the reference's
``network.Network`` carries for this (nr, species, is_forward, has_kinf,
is_special, is_photo, is_ion, is_conden, is_radiative, a, n, E, a_inf,
n_inf, E_inf, Rf, reactant_idx, product_idx, reactant_stoich,
product_stoich, stop_rev_indx).  Pins: tests/test_oracle_rates.py.
"""
from __future__ import annotations

import numpy as np

KB = 1.38064852e-16         # phy_const.py:3
P0 = 1.0e6                  # gibbs.py:12
CORR = KB / P0              # gibbs.py:13

SPECIAL_OH_CH3 = "OH + CH3 + M -> CH3OH + M"      # rates.py:27
LOWT_CH3 = "H + CH3 + M -> CH4 + M"               # rates.py:31
LOWT_C2H4 = "H + C2H4 + M -> C2H5 + M"            # rates.py:32
LOWT_C2H5 = "H + C2H5 + M -> C2H6 + M"            # rates.py:33


def arrhenius(a, n, E, T):
    """k = a T^n exp(-E/T) (rates.py:36-38)."""
    return a * T ** n * np.exp(-E / T)


def troe_oh_ch3(T, M):
    """The hard-coded Troe falloff of OH + CH3 + M -> CH3OH + M
    (rates.py:41-58)."""
    k0 = (1.932e3 * T ** -9.88 * np.exp(-7544.0 / T)
          + 5.109e-11 * T ** -6.25 * np.exp(-1433.0 / T))
    kinf = 1.031e-10 * T ** -0.018 * np.exp(16.74 / T)
    Fc = (0.1855 * np.exp(-T / 155.8) + 0.8145 * np.exp(-T / 1675.0)
          + np.exp(-4531.0 / T))
    nn = 0.75 - 1.27 * np.log(Fc)
    ff = np.exp(np.log(Fc) / (1.0 + (np.log(k0 * M / kinf) / nn) ** 2))
    return k0 / (1.0 + k0 * M / kinf) * ff


def compute_forward_k(net, T, M):
    """Forward rates on the grid, reverse slots zero (rates.py:61-122)."""
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    k = np.zeros((net.nr + 1, T.shape[0]))
    for i in range(1, net.nr + 1, 2):
        if not net.is_forward[i]:
            continue
        if net.is_photo[i] or net.is_ion[i] or net.is_conden[i] or net.is_radiative[i]:
            continue
        if net.is_special[i]:
            if net.Rf.get(i, "").strip() == SPECIAL_OH_CH3:
                k[i] = troe_oh_ch3(T, M)
            elif net.a[i] != 0.0:
                k[i] = arrhenius(net.a[i], net.n[i], net.E[i], T)
            continue
        if net.has_kinf[i]:
            k0 = arrhenius(net.a[i], net.n[i], net.E[i], T)
            kinf = arrhenius(net.a_inf[i], net.n_inf[i], net.E_inf[i], T)
            k[i] = k0 / (1.0 + k0 * M / kinf)
        else:
            k[i] = arrhenius(net.a[i], net.n[i], net.E[i], T)
    return k


def apply_lowT_caps(net, k, T, M):
    """The three low-T caps on forward rows (rates.py:142-167)."""
    k = np.array(k, np.float64)
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    for i in range(1, net.nr + 1, 2):
        if not net.is_forward[i]:
            continue
        rf = net.Rf.get(i, "")
        if rf == LOWT_CH3:
            k0 = 6.0e-29
            kinf = 2.06e-10 * T ** -0.4
            k[i] = np.where(T <= 277.5, k0 / (1.0 + k0 * M / kinf), k[i])
        elif rf == LOWT_C2H4:
            k[i] = np.where(T <= 300.0, 3.7e-30, k[i])
        elif rf == LOWT_C2H5:
            k[i] = np.where(T <= 200.0, 2.49e-27, k[i])
    return k


def load_nasa9(paths):
    """(ni, 2, 10) coefficients from NASA-9 text files, low-T range first
    (gibbs.py:16-38); a missing file gives zeros."""
    out = np.zeros((len(paths), 2, 10))
    for j, p in enumerate(paths):
        if p is None:
            continue
        flat = np.loadtxt(p).flatten()
        out[j, 0] = flat[0:10]
        out[j, 1] = flat[10:20]
    return out


def gibbs_sp_vector(coeffs, T):
    """g/(RT) per species and layer from the NASA-9 fits: H/RT - S/R, low
    range below 1000 K (gibbs.py:41-74)."""
    T = np.asarray(T, np.float64)[None, :]

    def branch(a):
        h = (-a[:, 0:1] / T ** 2 + a[:, 1:2] * np.log(T) / T + a[:, 2:3] + a[:, 3:4] * T / 2.0
             + a[:, 4:5] * T ** 2 / 3.0 + a[:, 5:6] * T ** 3 / 4.0 + a[:, 6:7] * T ** 4 / 5.0
             + a[:, 8:9] / T)
        s = (-a[:, 0:1] / T ** 2 / 2.0 - a[:, 1:2] / T + a[:, 2:3] * np.log(T) + a[:, 3:4] * T
             + a[:, 4:5] * T ** 2 / 2.0 + a[:, 5:6] * T ** 3 / 3.0 + a[:, 6:7] * T ** 4 / 4.0
             + a[:, 9:10])
        return h - s

    return np.where(T < 1000.0, branch(coeffs[:, 0, :]), branch(coeffs[:, 1, :]))


def K_eq_array(net, g_sp, T):
    """Equilibrium constants of the forward rows; 1 where there is no thermal
    reverse (gibbs.py:77-128)."""
    T = np.asarray(T, np.float64)
    nz = T.shape[0]
    K = np.ones((net.nr + 1, nz))
    g = np.concatenate([g_sp, np.zeros((1, nz))], axis=0)
    for i in range(1, net.nr + 1, 2):
        if not net.is_forward[i] or i + 1 >= net.stop_rev_indx:
            continue
        if net.is_photo[i] or net.is_ion[i] or net.is_conden[i] or net.is_radiative[i]:
            continue
        rs = np.zeros(nz)
        for s in range(net.reactant_idx.shape[1]):
            if net.reactant_stoich[i, s] != 0.0:
                rs += net.reactant_stoich[i, s] * g[net.reactant_idx[i, s]]
        ps = np.zeros(nz)
        for s in range(net.product_idx.shape[1]):
            if net.product_stoich[i, s] != 0.0:
                ps += net.product_stoich[i, s] * g[net.product_idx[i, s]]
        dn = float(net.reactant_stoich[i].sum()) - float(net.product_stoich[i].sum())
        K[i] = np.exp(rs - ps)
        if dn != 0.0:
            K[i] *= (CORR * T) ** dn
    return K


def fill_reverse_k(net, k, K_eq, remove_list=None):
    """Reverse rows k_f / K; a removed row takes its partner with it; rows past
    stop_rev_indx (odd offsets) are zeroed (gibbs.py:131-154)."""
    k = np.array(k, np.float64)
    remove = set(remove_list or [])
    for i in range(2, net.stop_rev_indx, 2):
        f = i - 1
        if i in remove or f in remove:
            k[i] = 0.0
            k[f] = 0.0
            continue
        with np.errstate(divide="ignore", invalid="ignore"):
            k[i] = np.where(K_eq[f] > 0, k[f] / K_eq[f], 0.0)
    for i in range(net.stop_rev_indx + 1, net.nr + 1, 2):
        k[i] = 0.0
    return k


def apply_remove_list(net, k, remove_list):
    """Zero the listed rows, no pairing (rates.py:170-182)."""
    k = np.array(k, np.float64)
    for i in remove_list or []:
        if 0 <= int(i) <= net.nr:
            k[int(i)] = 0.0
    return k


def build_rate_array(net, T, M, coeffs, use_lowT_limit_rates=False, remove_list=None):
    """Forward, optional low-T caps, reverse fill, then the remove list
    without pairing (rates.py:185-208)."""
    k = compute_forward_k(net, T, M)
    if use_lowT_limit_rates:
        k = apply_lowT_caps(net, k, T, M)
    K = K_eq_array(net, gibbs_sp_vector(coeffs, T), T)
    k = fill_reverse_k(net, k, K, remove_list=None)
    return apply_remove_list(net, k, remove_list)


def compute_all_k(net, T, M, coeffs, remove_list=None):
    """Forward rates plus the reverse fill with paired removal
    (gibbs.py:157-170)."""
    k = compute_forward_k(net, T, M)
    K = K_eq_array(net, gibbs_sp_vector(coeffs, T), T)
    return fill_reverse_k(net, k, K, remove_list=remove_list)
