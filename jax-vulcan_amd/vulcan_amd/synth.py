"""Seeded synthetic columns for the five Ros2 workloads.

Every array is float64 / int64 in cgs units and row-major.  The generators
follow the reference's setup code for geometry and table layout:

* pressure interfaces: ``atm_setup.compute_pico`` (atm_setup.py:23-38);
* heights, gravity, scale height, ``dzi``: ``compute_mu_dz_g`` with its two
  sequential scans (atm_setup.py:253-375), reference pressure 1 bar;
* reaction tables: forward record at an odd row, its reverse on the next even
  row with reactants and products swapped, PAD = ni, duplicate species
  collapsed into one slot with stoichiometry 2 (network.py:270-284, 366-419);
* forward rate ``A * T**b * exp(-E/T)`` (rates.py:36-38) and reverse rate
  ``k_f / K_eq`` (gibbs.py:147-150).

The workloads (all FP64):

1. one column, nz=100, ni=50, 300 reversible 2->2 reactions, dt=1e-6 s;
2. one column, nz=150, ni=93, 1200 reversible reactions + 60 photolysis
   branches, 2500 wavelength bins, two-stream photolysis;
3. a batch of uneven columns, nz in U{60..200}, ni in {50, 69, 93},
   13 ni reversible reactions (26 ni rate rows);
4. gradients at a converged workload-2 state;
5. a Poisson stream of workload-3-sized requests.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from ._types import AtmStatic, NetworkArrays, RateNetwork

# Physical constants (phy_const.py:3-9).
KB = 1.38064852e-16
NAVO = 6.02214086e23
HC = 1.98644582e-9
AU = 1.49597871e13
R_SUN = 6.957e10
R_JUP = 7.1492e9

# Planet / star used for geometry and the stellar flux (vulcan_cfg.py:41-45,81).
GS = 2140.0
RP = 1.138 * R_JUP
MU_MEAN = 2.3
R_STAR = 0.805
ORBIT_AU = 0.03142
SL_ANGLE = 48.0 / 180.0 * 3.14159
EDD = 0.5


@dataclass
class PhotoSetup:
    """Cross sections and stellar flux for the photolysis branches.

    Field meanings follow photo.PhotoData / PhotoJData (photo.py:15-32,
    248-266); there are no temperature-dependent cross sections here.
    """
    bins: np.ndarray          # (nbin,) nm
    dbin1: float
    dbin2: float
    din12_indx: int
    sflux_top: np.ndarray     # (nbin,) erg cm-2 s-1 nm-1 at the top
    absp_idx: np.ndarray      # (n_abs,) species index
    absp_cross: np.ndarray    # (n_abs, nbin) cm2
    scat_idx: np.ndarray      # (n_scat,)
    scat_cross: np.ndarray    # (n_scat, nbin)
    cross_J: np.ndarray       # (n_br, nbin)
    branch_re_idx: np.ndarray  # (n_br,) k row written by each branch
    branch_active: np.ndarray  # (n_br,) bool
    mu_zenith: float = float(np.cos(SL_ANGLE))
    edd: float = EDD
    ag0: float = 0.0
    hc: float = HC
    f_diurnal: float = 1.0
    # temperature-dependent absorbers and branches (photo.py:107-111, 334-347);
    # their per-layer cross sections are per column (Column.photo_T)
    absp_T_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    branch_T_re_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    branch_T_active: np.ndarray = field(default_factory=lambda: np.zeros(0, bool))


@dataclass
class Column:
    """One synthetic column: state, rates, atmosphere and network."""
    name: str
    nz: int
    ni: int
    y: np.ndarray             # (nz, ni) number densities cm-3
    k: np.ndarray             # (nr+1, nz)
    atm: AtmStatic
    net: NetworkArrays
    dt: float
    P: np.ndarray             # (nz,) dyn cm-2
    dz: np.ndarray            # (nz,)
    photo: Optional[PhotoSetup] = None
    meta: dict = field(default_factory=dict)
    photo_T: Optional[dict] = None   # absp_T_cross (n_T, nz, nbin), cross_J_T (n_brT, nz, nbin)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------

def compute_pico(pco: np.ndarray) -> np.ndarray:
    """Interface pressures from layer-centre pressures (atm_setup.py:23-38)."""
    pco = np.asarray(pco, dtype=np.float64)
    interior = np.sqrt(pco[1:] * pco[:-1])
    bottom = pco[0] ** 1.5 * pco[1] ** -0.5
    top = pco[-1] ** 1.5 * pco[-2] ** -0.5
    return np.concatenate([[bottom], interior, [top]])


def compute_geometry(pco: np.ndarray, Tco: np.ndarray, mu: float = MU_MEAN,
                     gs: float = GS, Rp: float = RP):
    """Hydrostatic heights, gravity and scale heights (atm_setup.py:253-375).

    The reference layer is the interface closest to 1 bar; the column is
    integrated upward from it and downward below it.
    """
    nz = Tco.shape[0]
    pico = compute_pico(pco)
    pref = int(np.argmin(np.abs(np.log10(pico) - 6.0)))
    pref = min(pref, nz - 1)
    gz = np.zeros(nz)
    Hp = np.zeros(nz)
    dz = np.zeros(nz)
    zco = np.zeros(nz + 1)
    z = 0.0
    for n_i, i in enumerate(range(pref, nz)):
        g_i = gs if n_i == 0 else gs * (Rp / (Rp + z)) ** 2
        H_i = KB * Tco[i] / (mu / NAVO * g_i)
        d_i = H_i * np.log(pico[i] / pico[i + 1])
        z = z + d_i
        gz[i], Hp[i], dz[i], zco[i + 1] = g_i, H_i, d_i, z
    z = 0.0
    for i in range(pref - 1, -1, -1):
        g_i = gs * (Rp / (Rp + z)) ** 2
        H_i = KB * Tco[i] / (mu / NAVO * g_i)
        d_i = H_i * np.log(pico[i] / pico[i + 1])
        z = z - d_i
        gz[i], Hp[i], dz[i], zco[i] = g_i, H_i, d_i, z
    dzi = 0.5 * (dz[1:] + dz[:-1])
    Ti = 0.5 * (Tco[1:] + Tco[:-1])
    Hpi = 0.5 * (Hp[1:] + Hp[:-1])
    return dict(pico=pico, g=gz, Hp=Hp, dz=dz, dzi=dzi, zco=zco, Ti=Ti, Hpi=Hpi)


def smooth_temperature(rng, P: np.ndarray, lo=500.0, hi=2000.0, n_modes=4):
    """Smooth random T(z): low-order cosine modes in log P mapped to [lo, hi]."""
    x = (np.log10(P) - np.log10(P).min()) / max(np.ptp(np.log10(P)), 1e-30)
    prof = np.zeros_like(x)
    for m in range(n_modes):
        prof += rng.normal() / (m + 1) * np.cos(np.pi * m * x + rng.uniform(0, 2 * np.pi))
    prof = (prof - prof.min()) / max(np.ptp(prof), 1e-30)
    return lo + (hi - lo) * prof


# --------------------------------------------------------------------------
# network
# --------------------------------------------------------------------------

def _collapse(species):
    out = {}
    for sp in species:
        out[sp] = out.get(sp, 0.0) + 1.0
    return out


def random_network(rng, ni: int, n_rev: int, n_photo: int = 0):
    """Reversible 2->2 reactions plus optional A -> B + C photolysis rows.

    Returns (net, photo_parent, photo_rows).  Forward reaction q sits at row
    2q+1, its reverse at 2q+2; photolysis branches follow at rows
    2*n_rev+1 ... (single rows, no reverse).
    """
    nr = 2 * n_rev + n_photo
    T = 2
    PAD = ni
    r_idx = np.full((nr + 1, T), PAD, dtype=np.int64)
    p_idx = np.full((nr + 1, T), PAD, dtype=np.int64)
    r_st = np.zeros((nr + 1, T))
    p_st = np.zeros((nr + 1, T))
    for q in range(n_rev):
        reac = _collapse(rng.integers(0, ni, size=2).tolist())
        prod = _collapse(rng.integers(0, ni, size=2).tolist())
        i, ir = 2 * q + 1, 2 * q + 2
        for s, (sp, st) in enumerate(reac.items()):
            r_idx[i, s], r_st[i, s] = sp, st
            p_idx[ir, s], p_st[ir, s] = sp, st
        for s, (sp, st) in enumerate(prod.items()):
            p_idx[i, s], p_st[i, s] = sp, st
            r_idx[ir, s], r_st[ir, s] = sp, st
    parents = np.zeros(n_photo, dtype=np.int64)
    rows = np.zeros(n_photo, dtype=np.int64)
    for b in range(n_photo):
        i = 2 * n_rev + 1 + b
        a = int(rng.integers(0, ni))
        prod = _collapse(rng.integers(0, ni, size=2).tolist())
        r_idx[i, 0], r_st[i, 0] = a, 1.0
        for s, (sp, st) in enumerate(prod.items()):
            p_idx[i, s], p_st[i, s] = sp, st
        parents[b], rows[b] = a, i
    net = NetworkArrays(ni=ni, nr=nr, reactant_idx=r_idx, product_idx=p_idx,
                        reactant_stoich=r_st, product_stoich=p_st,
                        is_three_body=np.zeros(nr + 1, dtype=bool))
    return net, parents, rows


def random_rate_params(rng, n_rev: int, net: NetworkArrays | None = None,
                       consistent: bool = False):
    """(A, b, E, log10 K_eq) per reversible reaction.

    consistent=False draws log10 K_eq ~ U[-5, 5] per reaction.
    consistent=True draws a potential g_i ~ U[-1.25, 1.25] per species and sets
    log10 K_eq = sum_products g - sum_reactants g (detailed balance, so the
    chemistry has one equilibrium; the span of log10 K stays [-5, 5])."""
    A = 10.0 ** rng.uniform(-14.0, -9.0, size=n_rev)
    b = rng.uniform(-2.0, 2.0, size=n_rev)
    E = rng.uniform(0.0, 2.0e4, size=n_rev)
    logK = rng.uniform(-5.0, 5.0, size=n_rev)
    if consistent:
        ni = net.ni
        g = np.concatenate([rng.uniform(-1.25, 1.25, size=ni), [0.0]])
        for q in range(n_rev):
            i = 2 * q + 1
            logK[q] = (np.sum(g[net.product_idx[i]] * net.product_stoich[i])
                       - np.sum(g[net.reactant_idx[i]] * net.reactant_stoich[i]))
    return A, b, E, logK


def rates_from_params(params, Tco: np.ndarray, n_photo: int = 0):
    """k (nr+1, nz): forward A*T^b*exp(-E/T), reverse k_f / 10**logK."""
    A, b, E, logK = params
    n_rev = A.shape[0]
    nz = Tco.shape[0]
    k = np.zeros((2 * n_rev + n_photo + 1, nz))
    for q in range(n_rev):
        kf = A[q] * Tco ** b[q] * np.exp(-E[q] / Tco)
        k[2 * q + 1] = kf
        k[2 * q + 2] = kf / 10.0 ** logK[q]
    return k


def random_rates(rng, Tco: np.ndarray, n_rev: int, n_photo: int = 0, net=None,
                 consistent: bool = False):
    """k (nr+1, nz) with freshly drawn parameters."""
    return rates_from_params(random_rate_params(rng, n_rev, net, consistent), Tco, n_photo)


# --------------------------------------------------------------------------
# rate-construction networks
# --------------------------------------------------------------------------

SPECIAL_OH_CH3 = "OH + CH3 + M -> CH3OH + M"
LOWT_EQUATIONS = ("H + CH3 + M -> CH4 + M", "H + C2H4 + M -> C2H5 + M",
                  "H + C2H5 + M -> C2H6 + M")


def random_rate_network(rng, net: NetworkArrays, n_rev: int, n_photo: int = 0,
                        f_lindemann: float = 0.2) -> RateNetwork:
    """Rate parameters and row flags over the reaction tables of ``net``.

    Forward rows draw modified Arrhenius parameters, a fraction
    ``f_lindemann`` a Lindemann pair (k0, k_inf); among the first rows sit
    the hard-coded Troe reaction, a special row with and one without
    Arrhenius columns, the three low-T cap reactions, one radiative and one
    condensation row.  Photolysis rows follow the reversible block, which
    ends at stop_rev_indx."""
    nr = net.nr
    S = nr + 1
    z = lambda: np.zeros(S)
    b = lambda: np.zeros(S, dtype=bool)
    a, n, E, a_inf, n_inf, E_inf = z(), z(), z(), z(), z(), z()
    is_forward, has_kinf, is_special = b(), b(), b()
    is_photo, is_ion, is_conden, is_radiative = b(), b(), b(), b()
    Rf = {}
    for q in range(n_rev):
        i = 2 * q + 1
        is_forward[i] = True
        Rf[i] = f"R{q} forward"
        if rng.uniform() < f_lindemann:
            has_kinf[i] = has_kinf[i + 1] = True
            a[i] = 10.0 ** rng.uniform(-33.0, -28.0)
            n[i] = rng.uniform(-3.0, 0.5)
            E[i] = rng.uniform(0.0, 3.0e3)
            a_inf[i] = 10.0 ** rng.uniform(-12.0, -9.0)
            n_inf[i] = rng.uniform(-1.0, 1.0)
            E_inf[i] = rng.uniform(0.0, 2.0e3)
        else:
            a[i] = 10.0 ** rng.uniform(-14.0, -9.0)
            n[i] = rng.uniform(-2.0, 2.0)
            E[i] = rng.uniform(0.0, 2.0e4)
    special = [(1, SPECIAL_OH_CH3, 0.0), (3, "X + Y -> Z", 2.5e-11), (5, "X + Y -> W", 0.0)]
    for i, eq, av in special:
        if i < 2 * n_rev:
            is_special[i] = is_special[i + 1] = True
            has_kinf[i] = has_kinf[i + 1] = False
            Rf[i] = "  " + eq + " "
            a[i], n[i], E[i] = av, 0.3 if av else 0.0, 120.0 if av else 0.0
    for j, eq in enumerate(LOWT_EQUATIONS):
        i = 7 + 2 * j
        if i < 2 * n_rev:
            Rf[i] = eq
    if 15 < 2 * n_rev:
        is_radiative[13] = is_radiative[14] = True
        is_conden[15] = is_conden[16] = True
    for r in range(n_photo):
        i = 2 * n_rev + 1 + r
        is_photo[i] = True
        is_forward[i] = (i % 2 == 1)
    return RateNetwork(
        species=tuple(f"S{j}" for j in range(net.ni)), ni=net.ni, nr=nr,
        reactant_idx=net.reactant_idx, product_idx=net.product_idx,
        reactant_stoich=net.reactant_stoich, product_stoich=net.product_stoich,
        a=a, n=n, E=E, a_inf=a_inf, n_inf=n_inf, E_inf=E_inf, is_forward=is_forward,
        has_kinf=has_kinf, is_special=is_special, is_photo=is_photo, is_ion=is_ion,
        is_conden=is_conden, is_radiative=is_radiative,
        stop_rev_indx=2 * n_rev + 1 if n_photo else nr + 1, Rf=Rf)


def random_nasa9(rng, ni: int) -> np.ndarray:
    """(ni, 2, 10) NASA-9 coefficients of plausible magnitude (low-T range
    first); a7 unused, a8 and a9 the enthalpy and entropy constants."""
    c = np.zeros((ni, 2, 10))
    for r, s3 in ((0, 5e-3), (1, 1e-3)):
        c[:, r, 0] = rng.uniform(-5e4, 5e4, ni)
        c[:, r, 1] = rng.uniform(-1e3, 1e3, ni)
        c[:, r, 2] = rng.uniform(2.0, 8.0, ni)
        c[:, r, 3] = rng.uniform(-s3, s3, ni)
        c[:, r, 4] = rng.uniform(-1e-5, 1e-5, ni)
        c[:, r, 5] = rng.uniform(-1e-8, 1e-8, ni)
        c[:, r, 6] = rng.uniform(-1e-12, 1e-12, ni)
        c[:, r, 8] = rng.uniform(-6e4, 6e4, ni)
        c[:, r, 9] = rng.uniform(-20.0, 30.0, ni)
    return c


# --------------------------------------------------------------------------
# photolysis setup
# --------------------------------------------------------------------------

def wavelength_grid(nbin: int = 2500, lo: float = 10.0, hi: float = 700.0,
                    n1: Optional[int] = None):
    """Two-resolution grid: n1 bins at dbin1 from lo, the rest at dbin2 to hi.

    The J trapezoid is split at din12_indx = n1 as in photo.py:309-330.
    """
    if n1 is None:
        n1 = int(round(0.8 * nbin))
    dbin1 = 0.2 * (2500.0 / nbin)
    seg1 = lo + dbin1 * np.arange(n1)
    start2 = seg1[-1] + dbin1
    dbin2 = (hi - start2) / (nbin - n1 - 1)
    seg2 = start2 + dbin2 * np.arange(nbin - n1)
    return np.concatenate([seg1, seg2]), dbin1, dbin2, n1


def smooth_log_cross(rng, n: int, nbin: int, lo=-24.0, hi=-17.0, width=20):
    """Cross sections log-uniform in [10^lo, 10^hi], Gaussian-smoothed in log."""
    raw = rng.uniform(lo, hi, size=(n, nbin + 6 * width))
    x = np.arange(-3 * width, 3 * width + 1)
    ker = np.exp(-0.5 * (x / width) ** 2)
    ker /= ker.sum()
    sm = np.stack([np.convolve(r, ker, mode="valid")[:nbin] for r in raw])
    # smoothing narrows the spread; stretch back to the requested range
    mn = sm.min(axis=1, keepdims=True)
    mx = sm.max(axis=1, keepdims=True)
    sm = lo + (hi - lo) * (sm - mn) / np.maximum(mx - mn, 1e-30)
    return 10.0 ** sm


def blackbody_flux(bins_nm: np.ndarray, T_star=5800.0, r_star=R_STAR,
                   orbit_au=ORBIT_AU):
    """pi * B_lambda(T) diluted by (R*/a)^2, erg cm-2 s-1 nm-1."""
    h = 6.62607015e-27
    c = 2.99792458e10
    lam = bins_nm * 1e-7
    B = 2 * h * c ** 2 / lam ** 5 / np.expm1(h * c / (lam * KB * T_star))
    return np.pi * B * 1e-7 * (r_star * R_SUN / (orbit_au * AU)) ** 2


def random_photo(rng, ni, parents, rows, nbin=2500):
    bins, dbin1, dbin2, din12 = wavelength_grid(nbin)
    n_br = parents.shape[0]
    cross_J = smooth_log_cross(rng, n_br, nbin)
    absp_sp = np.unique(parents)
    absp_cross = np.stack([cross_J[parents == sp].sum(axis=0) for sp in absp_sp])
    scat_sp = np.array([s for s in range(ni) if s not in set(absp_sp.tolist())][:2],
                       dtype=np.int64)
    if scat_sp.size == 0:
        scat_sp = np.array([0], dtype=np.int64)
    sig0 = 10.0 ** rng.uniform(-23.0, -21.0, size=scat_sp.size)
    scat_cross = sig0[:, None] * (bins[None, :] / 100.0) ** -4
    return PhotoSetup(
        bins=bins, dbin1=dbin1, dbin2=dbin2, din12_indx=din12,
        sflux_top=blackbody_flux(bins),
        absp_idx=absp_sp.astype(np.int64), absp_cross=absp_cross,
        scat_idx=scat_sp, scat_cross=scat_cross,
        cross_J=cross_J, branch_re_idx=rows.astype(np.int64),
        branch_active=np.ones(n_br, dtype=bool),
    )


def with_t_dependent_photo(col: "Column", seed: int, n_abs_T: int = 2, n_br_T: int = 3):
    """The column with n_abs_T of its absorbers and n_br_T of its photolysis
    branches made temperature-dependent: sigma(lambda) (T_z / 1000 K)^a with
    a in U[-1.5, 1.5] per species or branch.  A T-dependent branch takes over
    the k row of the plain branch it replaces (which is deactivated)."""
    import dataclasses
    rng = np.random.default_rng(seed)
    ph = col.photo
    f = col.atm.Tco / 1000.0
    sel_a = np.sort(rng.choice(ph.absp_idx.size, size=n_abs_T, replace=False))
    keep = np.setdiff1d(np.arange(ph.absp_idx.size), sel_a)
    ea = rng.uniform(-1.5, 1.5, n_abs_T)
    absp_T_cross = ph.absp_cross[sel_a][:, None, :] * (f[None, :, None] ** ea[:, None, None])
    sel_b = np.sort(rng.choice(ph.cross_J.shape[0], size=n_br_T, replace=False))
    eb = rng.uniform(-1.5, 1.5, n_br_T)
    cross_J_T = ph.cross_J[sel_b][:, None, :] * (f[None, :, None] ** eb[:, None, None])
    active = ph.branch_active.copy()
    active[sel_b] = False
    ph2 = dataclasses.replace(
        ph, absp_idx=ph.absp_idx[keep], absp_cross=ph.absp_cross[keep],
        absp_T_idx=ph.absp_idx[sel_a], branch_active=active,
        branch_T_re_idx=ph.branch_re_idx[sel_b], branch_T_active=np.ones(n_br_T, bool))
    return dataclasses.replace(col, photo=ph2, photo_T={"absp_T_cross": absp_T_cross,
                                                       "cross_J_T": cross_J_T})


# --------------------------------------------------------------------------
# columns
# --------------------------------------------------------------------------

def make_column(seed: int, nz: int, ni: int, n_rev: int, n_photo: int = 0,
                nbin: int = 2500, dt: float = 1e-6, name: str = "column",
                p_bot: float = 1e9, p_top: float = 1e-2, consistent: bool = False,
                ymix_lo: float = -20.0) -> Column:
    """Draw one column with the workload-1 distributions (mixing ratios
    log-uniform 10^ymix_lo .. 1)."""
    rng = np.random.default_rng(seed)
    P = np.logspace(np.log10(p_bot), np.log10(p_top), nz)
    Tco = smooth_temperature(rng, P)
    geo = compute_geometry(P, Tco)
    M = P / (KB * Tco)
    ymix = 10.0 ** rng.uniform(ymix_lo, 0.0, size=(nz, ni))
    ymix /= ymix.sum(axis=1, keepdims=True)
    y = M[:, None] * ymix
    net, parents, rows = random_network(rng, ni, n_rev, n_photo)
    k = random_rates(rng, Tco, n_rev, n_photo, net, consistent)
    Kzz = 10.0 ** rng.uniform(5.0, 10.0, size=nz - 1)
    atm = AtmStatic(
        Kzz=Kzz, Dzz=np.zeros((nz - 1, ni)), dzi=geo["dzi"], vz=np.zeros(nz - 1),
        Hpi=geo["Hpi"], Ti=geo["Ti"], Tco=Tco, g=geo["g"],
        ms=np.full(ni, MU_MEAN), alpha=np.zeros(ni), M=M,
        vm=np.zeros((nz, ni)), vs=np.zeros((nz - 1, ni)),
        top_flux=np.zeros(ni), bot_flux=np.zeros(ni), bot_vdep=np.zeros(ni),
        gas_indx_mask=np.ones(ni, dtype=bool),
    )
    photo = random_photo(rng, ni, parents, rows, nbin) if n_photo > 0 else None
    return Column(name=name, nz=nz, ni=ni, y=y, k=k, atm=atm, net=net, dt=dt,
                  P=P, dz=geo["dz"], photo=photo,
                  meta=dict(seed=seed, n_rev=n_rev, n_photo=n_photo))


def columns_same_network(seed: int, B: int, nz: int, ni: int, n_rev: int,
                         n_photo: int = 0, dt: float = 1e-6, name: str = "col"):
    """B columns sharing one network and its rate parameters; each column
    draws its own T(z), Kzz and mixing ratios (workload-1 distributions)."""
    rng = np.random.default_rng(seed)
    net, parents, rows = random_network(rng, ni, n_rev, n_photo)
    params = random_rate_params(rng, n_rev)
    cols = []
    for c in range(B):
        r = np.random.default_rng([seed, c])
        P = np.logspace(9.0, -2.0, nz)
        Tco = smooth_temperature(r, P)
        geo = compute_geometry(P, Tco)
        M = P / (KB * Tco)
        ymix = 10.0 ** r.uniform(-20.0, 0.0, size=(nz, ni))
        ymix /= ymix.sum(axis=1, keepdims=True)
        atm = AtmStatic(
            Kzz=10.0 ** r.uniform(5.0, 10.0, size=nz - 1), Dzz=np.zeros((nz - 1, ni)),
            dzi=geo["dzi"], vz=np.zeros(nz - 1), Hpi=geo["Hpi"], Ti=geo["Ti"], Tco=Tco,
            g=geo["g"], ms=np.full(ni, MU_MEAN), alpha=np.zeros(ni), M=M,
            vm=np.zeros((nz, ni)), vs=np.zeros((nz - 1, ni)), top_flux=np.zeros(ni),
            bot_flux=np.zeros(ni), bot_vdep=np.zeros(ni), gas_indx_mask=np.ones(ni, dtype=bool))
        cols.append(Column(name=f"{name}_{c}", nz=nz, ni=ni, y=M[:, None] * ymix,
                           k=rates_from_params(params, Tco, n_photo), atm=atm, net=net, dt=dt,
                           P=P, dz=geo["dz"], meta=dict(seed=seed, index=c)))
    return cols


def config1(seed: int = 1001) -> Column:
    """Plain single Ros2 step: nz=100, ni=50, 300 reversible reactions."""
    return make_column(seed, nz=100, ni=50, n_rev=300, dt=1e-6, name="cfg1")


def config2(seed: int = 1002, nbin: int = 2500) -> Column:
    """Plain steady state: nz=150, ni=93, 1200 reversible + 60 photo branches."""
    return make_column(seed, nz=150, ni=93, n_rev=1200, n_photo=60, nbin=nbin,
                       dt=1e-10, name="cfg2")


def config3_sizes(seed: int = 1003, ncols: int = 512):
    """(nz, ni, n_rev) for the uneven batch: nz~U{60..200}, ni in {50,69,93}."""
    rng = np.random.default_rng(seed)
    nz = rng.integers(60, 201, size=ncols)
    ni = rng.choice(np.array([50, 69, 93]), size=ncols)
    n_rev = 13 * ni
    return [(int(a), int(b), int(c)) for a, b, c in zip(nz, ni, n_rev)]


def config3(seed: int = 1003, ncols: int = 512, dt: float = 1e-10):
    """Uneven batch of independently drawn columns."""
    sizes = config3_sizes(seed, ncols)
    return [make_column(seed * 100003 + c, nz=a, ni=b, n_rev=r, dt=dt,
                        name=f"cfg3_{c}") for c, (a, b, r) in enumerate(sizes)]


def poisson_arrivals(seed: int = 1005, n: int = 10000, lam: float = 200.0):
    """Arrival times (s) of a Poisson stream with rate lam."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / lam, size=n))


# --------------------------------------------------------------------------
# workload 4: an open column with a unique steady state
# --------------------------------------------------------------------------

@dataclass
class OpenColumn:
    """The differentiation workload: one nz=150 column of 92 trace species
    and a background gas (ni = 93) under 1200 reactions, open at the bottom.

    Reactions q = 0 .. 1199 (forward row 2q+1, reverse row 2q+2):

    * q = 0: the background gas, made at K_BG * M (three-body row with no
      reactants) and lost at K_BG, so it settles at M and sets ysum;
    * q = 1 .. 92: first-order sinks s -> 0 of every trace (irreversible);
    * q = 93 .. 112: sources 0 -> s of 20 traces at k M (irreversible);
    * q = 113 .. 612: reversible isomerisations s_a + M <-> s_b + M;
    * q = 613 .. 1199: reversible 2->2 reactions among the traces, with the
      workload-1 Arrhenius distributions.

    Both reversible kinds take K_eq from per-species potentials (detailed
    balance).  Ten traces with no volume source get bottom fluxes
    (log-uniform 1e5..1e12 cm-2 s-1); every trace deposits at the bottom.
    The pressure grid and T(z), Kzz and geometry follow workload 1.

    Parameters of the gradient: one log multiplier per reaction (shared by
    its two rows), log T(z) per layer (entering k(T), M = P/kT, Tco, Ti,
    Hpi, g and the hydrostatic dzi) and the log bottom fluxes.
    """
    nz: int
    ni: int
    net: NetworkArrays
    P: np.ndarray
    T0: np.ndarray
    Kzz: np.ndarray
    kind: np.ndarray           # (nq,) 0 bg, 1 sink, 2 source, 3 isomerisation, 4 two-body
    A: np.ndarray
    b: np.ndarray
    E: np.ndarray
    logK: np.ndarray
    flux_sp: np.ndarray        # (10,) species with bottom fluxes
    logbf0: np.ndarray         # (10,) natural log of the fluxes
    vdep: np.ndarray           # (ni,) deposition velocities
    loss_sp: np.ndarray        # (5,) species of the loss
    bg: int
    T_ref: float = 1000.0
    K_BG: float = 1.0

    @property
    def nq(self) -> int:
        return int(self.kind.shape[0])

    def leaves(self, T, logm=None, logbf=None, isobaric: bool = False):
        """(k_arr, AtmStatic) at temperatures T, multipliers exp(logm), fluxes exp(logbf).

        The number density M (and with it the background gas, hence ysum)
        stays at the reference profile P / (k T0) unless ``isobaric``, when
        it follows P / (k T).
        """
        T = np.asarray(T, dtype=np.float64)
        logm = np.zeros(self.nq) if logm is None else np.asarray(logm, dtype=np.float64)
        logbf = self.logbf0 if logbf is None else np.asarray(logbf, dtype=np.float64)
        M = self.P / (KB * (T if isobaric else self.T0))
        k = np.zeros((2 * self.nq + 1, self.nz))
        m = np.exp(logm)[:, None]
        kf = m * self.A[:, None] * (T[None] / self.T_ref) ** self.b[:, None] \
            * np.exp(-self.E[:, None] / T[None])
        kr = np.where((self.kind >= 3)[:, None], kf / 10.0 ** self.logK[:, None], 0.0)
        bgq = self.kind == 0
        kf[bgq] = m[bgq] * self.K_BG
        kr[bgq] = m[bgq] * self.K_BG
        k[1::2] = kf
        k[2::2] = kr
        geo = compute_geometry(self.P, T)
        bf = np.zeros(self.ni)
        bf[self.flux_sp] = np.exp(logbf)
        atm = AtmStatic(
            Kzz=self.Kzz, Dzz=np.zeros((self.nz - 1, self.ni)), dzi=geo["dzi"],
            vz=np.zeros(self.nz - 1), Hpi=geo["Hpi"], Ti=geo["Ti"], Tco=T, g=geo["g"],
            ms=np.full(self.ni, MU_MEAN), alpha=np.zeros(self.ni), M=M,
            vm=np.zeros((self.nz, self.ni)), vs=np.zeros((self.nz - 1, self.ni)),
            top_flux=np.zeros(self.ni), bot_flux=bf, bot_vdep=self.vdep,
            gas_indx_mask=np.ones(self.ni, dtype=bool), use_botflux=True)
        return k, atm

    def dlogk_dlogT(self, T):
        """d log k / d log T per reaction and layer (nq, nz): b + E/T for the
        Arrhenius rows (a reverse row shares its forward row's, K_eq being
        independent of T here), 0 for the background pair."""
        T = np.asarray(T, dtype=np.float64)
        s = self.b[:, None] + self.E[:, None] / T[None]
        s[self.kind == 0] = 0.0
        return s

    def initial_state(self):
        """A starting state for the steady-state solve: traces at 1e-12 of M."""
        y = np.empty((self.nz, self.ni))
        M = self.P / (KB * self.T0)
        y[:] = 1e-12 * M[:, None]
        y[:, self.bg] = M
        return y


def config4(seed: int = 1004, nz: int = 150, n_iso: int = 500, n_two: int = 587,
            n_src: int = 20, n_flux: int = 10) -> OpenColumn:
    """The differentiation workload (see OpenColumn)."""
    rng = np.random.default_rng(seed)
    nt = 92
    ni, bg = nt + 1, nt
    P = np.logspace(9.0, -2.0, nz)
    T0 = smooth_temperature(rng, P)
    Kzz = 10.0 ** rng.uniform(5.0, 10.0, size=nz - 1)
    pot = np.concatenate([rng.uniform(-1.25, 1.25, size=nt), [0.0]])
    perm = rng.permutation(nt)
    src_sp = np.sort(perm[:n_src])
    flux_sp = np.sort(perm[n_src:n_src + n_flux])
    loss_sp = np.sort(rng.choice(nt, size=5, replace=False))
    nq = 1 + nt + n_src + n_iso + n_two
    S = 2 * nq + 1
    PAD = ni
    r_idx = np.full((S, 2), PAD, dtype=np.int64)
    p_idx = np.full((S, 2), PAD, dtype=np.int64)
    r_st = np.zeros((S, 2))
    p_st = np.zeros((S, 2))
    is3 = np.zeros(S, dtype=bool)
    kind = np.zeros(nq, dtype=np.int64)
    A = np.zeros(nq)
    b = np.zeros(nq)
    E = np.zeros(nq)
    logK = np.zeros(nq)

    def put(q, reac, prod):
        f, r = 2 * q + 1, 2 * q + 2
        for s, (sp, st) in enumerate(_collapse(reac).items()):
            r_idx[f, s], r_st[f, s] = sp, st
            p_idx[r, s], p_st[r, s] = sp, st
        for s, (sp, st) in enumerate(_collapse(prod).items()):
            p_idx[f, s], p_st[f, s] = sp, st
            r_idx[r, s], r_st[r, s] = sp, st

    q = 0
    put(q, [], [bg])                     # 0 -> BG at K_BG M; BG -> 0 at K_BG
    is3[1] = True
    q += 1
    for s in range(nt):                  # sinks
        put(q, [s], [])
        kind[q], A[q] = 1, 10.0 ** rng.uniform(-5.0, -2.0)
        b[q], E[q] = rng.uniform(-1.0, 1.0), rng.uniform(0.0, 1000.0)
        q += 1
    for s in src_sp:                     # sources at k M
        put(q, [], [int(s)])
        is3[2 * q + 1] = True
        kind[q], A[q] = 2, 10.0 ** rng.uniform(-22.0, -18.0)
        b[q], E[q] = rng.uniform(-1.0, 1.0), rng.uniform(0.0, 1000.0)
        q += 1
    for _ in range(n_iso):               # s_a + M <-> s_b + M
        a_, b_ = rng.choice(nt, size=2, replace=False)
        put(q, [int(a_)], [int(b_)])
        is3[2 * q + 1] = is3[2 * q + 2] = True
        kind[q], A[q] = 3, 10.0 ** rng.uniform(-22.0, -17.0)
        b[q], E[q] = rng.uniform(-2.0, 2.0), rng.uniform(0.0, 2.0e4)
        logK[q] = pot[b_] - pot[a_]
        q += 1
    for _ in range(n_two):               # two-body 2 -> 2 among the traces
        reac = rng.integers(0, nt, size=2).tolist()
        prod = rng.integers(0, nt, size=2).tolist()
        put(q, reac, prod)
        kind[q], A[q] = 4, 10.0 ** rng.uniform(-14.0, -9.0)
        b[q], E[q] = rng.uniform(-2.0, 2.0), rng.uniform(0.0, 2.0e4)
        logK[q] = pot[prod].sum() - pot[reac].sum()
        q += 1
    assert q == nq
    net = NetworkArrays(ni=ni, nr=S - 1, reactant_idx=r_idx, product_idx=p_idx,
                        reactant_stoich=r_st, product_stoich=p_st, is_three_body=is3)
    vdep = np.full(ni, 1.0)
    vdep[bg] = 0.0
    logbf0 = np.log(10.0 ** rng.uniform(5.0, 12.0, size=n_flux))
    return OpenColumn(nz=nz, ni=ni, net=net, P=P, T0=T0, Kzz=Kzz, kind=kind, A=A, b=b, E=E,
                      logK=logK, flux_sp=flux_sp, logbf0=logbf0, vdep=vdep, loss_sp=loss_sp,
                      bg=bg)
