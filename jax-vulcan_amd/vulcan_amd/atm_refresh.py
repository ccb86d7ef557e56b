"""Atmosphere refresh on the GPU: mean molecular mass, hydrostatic column
and interface averages from the current mixing ratios, and the
diffusion-limited escape flux at the top.

Same names, arguments and results as atm_refresh.py of the reference:
``AtmRefreshStatic`` (atm_refresh.py:17-33), ``update_mu_dz_jax``
(atm_refresh.py:36-98) and ``update_phi_esc_jax`` (atm_refresh.py:101-122),
each one launch of the device kernel.  ``RefreshSet`` holds the refresh
records of a whole batch; ``SteadyStateRunner`` uses it to refresh
columns inside the attempt loop after every ``update_frq``-th accepted
step, as the reference's body_fn does (outer_loop.py:809-843).
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple, Sequence

import numpy as np
import torch

from . import _lib
from ._types import AtmStatic, NetworkArrays
from .core import F64, Batch, device, device_network, pack_atm_record

__all__ = ["AtmRefreshStatic", "RefreshSet", "update_mu_dz_jax", "update_phi_esc_jax",
           "refresh_record"]

KB = 1.38064852e-16          # phy_const.py
NAVO = 6.02214086e23


class AtmRefreshStatic(NamedTuple):
    """Static inputs of the refresh (atm_refresh.py:17-33)."""
    Tco: np.ndarray            # (nz,)
    pico: np.ndarray           # (nz+1,) interface pressures
    mol_mass: np.ndarray       # (ni,) g/mol
    ms: np.ndarray             # (ni,)
    Dzz_top: np.ndarray        # (ni,) molecular diffusion at the top interface
    diff_esc_idx: np.ndarray   # (n_esc,) species escaping at the top
    pref_indx: int
    zco_pref: float
    gs: float
    Rp: float
    kb: float = KB
    Navo: float = NAVO
    max_flux: float = 1e13
    nz: int = 0
    ni: int = 0


class RefreshState(C.Structure):
    _fields_ = [("rec", C.c_void_p), ("off_r", C.c_void_p), ("geo", C.c_void_p),
                ("off_g", C.c_void_p), ("due", C.c_void_p)]


def _np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def refresh_record(st: AtmRefreshStatic) -> np.ndarray:
    """One column's refresh record in the order of vk_refresh_record_layout."""
    nz = int(np.shape(st.Tco)[0])
    ni = int(np.shape(st.mol_mass)[0])
    pref = int(st.pref_indx)
    if not 0 <= pref < nz:
        raise ValueError(f"pref_indx {pref} outside 0..{nz - 1}")
    pico = np.asarray(_np(st.pico), np.float64).reshape(-1)
    if pico.shape[0] != nz + 1:
        raise ValueError(f"pico has {pico.shape[0]} entries, expected {nz + 1}")
    esc = np.zeros(ni)
    idx = np.asarray(_np(st.diff_esc_idx), np.int64).reshape(-1)
    if idx.size and (idx.min() < 0 or idx.max() >= ni):
        raise ValueError("diff_esc_idx out of range")
    esc[idx] = 1.0
    rec = np.concatenate([pico, np.asarray(_np(st.mol_mass), np.float64), esc,
                          [float(pref), float(st.zco_pref), float(st.gs), float(st.Rp),
                           float(st.max_flux), float(st.kb), float(st.Navo)]])
    total = int(_lib.load().vk_refresh_record_layout(nz, ni, None))
    assert rec.size == total
    return rec


class RefreshSet:
    """Refresh records, geometry outputs and due flags of a batch."""

    def __init__(self, batch: Batch, statics: Sequence[AtmRefreshStatic]):
        if len(statics) != batch.ncols:
            raise ValueError(f"need one AtmRefreshStatic per column ({batch.ncols})")
        dev = device()
        recs = []
        for c, st in enumerate(statics):
            if int(np.shape(st.Tco)[0]) != int(batch.nz[c]) or \
                    int(np.shape(st.mol_mass)[0]) != int(batch.ni[c]):
                raise ValueError(f"column {c}: refresh static does not match (nz, ni)")
            recs.append(refresh_record(st))
        sizes_r = np.array([r.size for r in recs], np.int64)
        sizes_g = 5 * batch.nz.astype(np.int64) + 1
        ex = lambda s: np.concatenate([[0], np.cumsum(s)[:-1]]).astype(np.int64)
        self.off_r_h, self.off_g_h = ex(sizes_r), ex(sizes_g)
        self.nz = batch.nz.copy()
        self.rec = torch.tensor(np.concatenate(recs), dtype=F64, device=dev)
        self.off_r = torch.tensor(self.off_r_h, dtype=torch.int64, device=dev)
        self.off_g = torch.tensor(self.off_g_h, dtype=torch.int64, device=dev)
        self.geo = torch.zeros(int(sizes_g.sum()), dtype=F64, device=dev)
        self.due = torch.zeros(batch.ncols, dtype=torch.int32, device=dev)
        self.batch = batch

    def state(self, use_due: bool) -> RefreshState:
        return RefreshState(self.rec.data_ptr(), self.off_r.data_ptr(), self.geo.data_ptr(),
                            self.off_g.data_ptr(), self.due.data_ptr() if use_due else None)

    def run(self, atm: torch.Tensor, mode: int = 3, ymix=None, y=None, sol=None, params=None,
            use_due: bool = False):
        st = self.state(use_due)
        _lib.check(_lib.load().vk_atm_refresh_f64(
            self.batch.handle, int(mode), _lib.ptr(ymix), _lib.ptr(y), _lib.ptr(sol),
            None if params is None else C.byref(params), atm.data_ptr(), C.byref(st),
            _lib.stream_ptr()), "vk_atm_refresh_f64")
        return self

    def geometry(self, c: int) -> dict:
        """mu, g, Hp, dz (nz) and zco (nz+1) of column c."""
        nz = int(self.nz[c])
        o = int(self.off_g_h[c])
        g = self.geo[o:o + 5 * nz + 1]
        return dict(mu=g[:nz], g=g[nz:2 * nz], Hp=g[2 * nz:3 * nz], dz=g[3 * nz:4 * nz],
                    zco=g[4 * nz:])


_ONE: dict = {}


def _one_column(st: AtmRefreshStatic, top_flux=None):
    """A one-column batch whose atmosphere record carries what the refresh
    reads (Tco, ms, the top Dzz row, top_flux)."""
    nz = int(np.shape(st.Tco)[0])
    ni = int(np.shape(st.mol_mass)[0])
    net = _ONE.get(ni)
    if net is None:
        pad = np.full((1, 1), ni, np.int64)
        net = NetworkArrays(ni, 0, pad, pad.copy(), np.zeros((1, 1)), np.zeros((1, 1)),
                            np.zeros(1, bool))
        _ONE[ni] = net
    b = Batch.uniform_of(1, nz, device_network(net))
    Dzz = np.zeros((nz - 1, ni))
    Dzz[-1] = _np(st.Dzz_top)
    z1 = np.zeros(nz - 1)
    atm = AtmStatic(Kzz=z1, Dzz=Dzz, dzi=z1, vz=z1, Hpi=z1, Ti=z1,
                    Tco=np.asarray(_np(st.Tco), np.float64), g=np.zeros(nz),
                    ms=np.asarray(_np(st.ms), np.float64), alpha=np.zeros(ni), M=np.zeros(nz),
                    vm=np.zeros((nz, ni)), vs=np.zeros((nz - 1, ni)),
                    top_flux=np.zeros(ni) if top_flux is None else _np(top_flux),
                    bot_flux=np.zeros(ni), bot_vdep=np.zeros(ni),
                    gas_indx_mask=np.ones(ni, bool))
    rec = pack_atm_record(atm).to(device())
    return b, rec, RefreshSet(b, [st])


def update_mu_dz_jax(ymix, st: AtmRefreshStatic):
    """(mu, g, Hp, dz, zco, dzi, Hpi) from the mixing ratios
    (atm_refresh.py:36-98)."""
    from .core import atm_record_layout
    b, rec, rs = _one_column(st)
    ymix = torch.as_tensor(_np(ymix), dtype=F64, device=device()).contiguous()
    nz = int(b.nz[0])
    rs.run(rec, mode=1, ymix=ymix)
    geo = rs.geometry(0)
    _, off = atm_record_layout(nz, int(b.ni[0]))
    dzi = rec[off[1]:off[1] + nz - 1].clone()
    Hpi = rec[off[3]:off[3] + nz - 1].clone()
    return (geo["mu"].clone(), geo["g"].clone(), geo["Hp"].clone(), geo["dz"].clone(),
            geo["zco"].clone(), dzi, Hpi)


def update_phi_esc_jax(y, g, Hp, top_flux_in, st: AtmRefreshStatic):
    """top_flux_in with the escaping species' diffusion-limited fluxes,
    floored at -max_flux (atm_refresh.py:101-122)."""
    from .core import atm_record_layout
    b, rec, rs = _one_column(st, top_flux=top_flux_in)
    nz, ni = int(b.nz[0]), int(b.ni[0])
    geo = rs.geometry(0)
    geo["g"].copy_(torch.as_tensor(_np(g), dtype=F64))
    geo["Hp"].copy_(torch.as_tensor(_np(Hp), dtype=F64))
    y = torch.as_tensor(_np(y), dtype=F64, device=device()).contiguous()
    rs.run(rec, mode=2, y=y)
    _, off = atm_record_layout(nz, ni)
    return rec[off[10]:off[10] + ni].clone()
