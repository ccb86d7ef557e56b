"""Photolysis on the GPU: optical depth, two-stream flux, J-rates, k update.

Reference-named entry points with the argument order of photo.py:
``compute_tau_jax(y, dz, photo)`` (photo.py:84-123),
``compute_flux_jax(tau, sflux_top, ymix, photo, bins, mu_zenith, edd, ag0,
hc, dflux_u_prev, ag0_is_zero)`` (photo.py:126-245),
``compute_J_jax_flat(aflux, cross_J, cross_J_T, din12_indx, dbin1, dbin2)``
(photo.py:377-392) and ``update_k_with_J(...)`` (photo.py:455-478).
``PhotoData`` keeps the fields of photo.PhotoData (photo.py:15-32) and
``photo_data_from_static`` (photo.py:35-81) builds it.  Temperature-dependent
cross sections (absp_T_cross (n_T, nz, nbin), cross_J_T (n_brT, nz, nbin);
photo.py:107-111, 156-160, 334-347) are carried per column.

``PhotoBatch`` is the in-loop photolysis of ``SteadyStateRunner``: all of
tau, flux, J and the k rows for every column whose photo_due flag is set.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .core import F64, Batch, device, device_network
from .solver import _solver_network


class PhotoData(NamedTuple):
    absp_idx: np.ndarray
    absp_cross: np.ndarray
    absp_T_idx: np.ndarray
    absp_T_cross: np.ndarray
    scat_idx: np.ndarray
    scat_cross: np.ndarray


class PhotoSetupC(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nbin", "n_abs", "n_scat", "n_br", "din12", "ag0_zero")]
                + [(n, C.c_double) for n in ("dbin1", "dbin2", "mu_zenith", "edd", "ag0", "hc",
                                             "f_diurnal", "flux_atol")]
                + [(n, C.c_void_p) for n in ("bins", "sflux_top", "absp_cross", "scat_cross",
                                             "cross_J", "dz", "absp_idx", "scat_idx",
                                             "branch_re_idx", "branch_active", "off_p", "off_l")]
                + [(n, C.c_int32) for n in ("n_absT", "n_brT")]
                + [(n, C.c_void_p) for n in ("absp_T_cross", "cross_J_T", "absp_T_idx",
                                             "branch_T_re_idx", "branch_T_active")])


class PhotoStateC(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("tau", "w0", "sflux", "dflux_d", "dflux_u", "aflux",
                                          "J", "aflux_change", "due", "status", "J_T")]


def _t(a, dtype=F64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device())


class _Setup:
    """Device copy of one shared photolysis setup plus its C struct."""

    def __init__(self, ph, nzs, dzs, flux_atol=1.0, branch_re_idx=None, branch_active=None,
                 cross_J=None, tcross=None):
        """tcross: per column None or a dict with the column's
        ``absp_T_cross`` (n_T, nz, nbin) and ``cross_J_T`` (n_brT, nz, nbin)
        for the indices ph['absp_T_idx'] / ph['branch_T_re_idx']."""
        self.nbin = int(np.asarray(ph["bins"]).shape[0])
        self.bins = _t(ph["bins"])
        self.sflux_top = _t(ph["sflux_top"])
        self.absp_idx = _t(ph["absp_idx"], torch.int32)
        self.absp_cross = _t(np.asarray(ph["absp_cross"]).reshape(-1, self.nbin))
        self.scat_idx = _t(ph["scat_idx"], torch.int32)
        self.scat_cross = _t(np.asarray(ph["scat_cross"]).reshape(-1, self.nbin))
        cj = np.asarray(ph["cross_J"] if cross_J is None else cross_J).reshape(-1, self.nbin)
        self.cross_J = _t(cj)
        n_br = cj.shape[0]
        re = np.zeros(n_br, np.int32) if branch_re_idx is None else np.asarray(branch_re_idx)
        act = np.zeros(n_br, np.int32) if branch_active is None else np.asarray(branch_active)
        self.branch_re_idx = _t(re, torch.int32)
        self.branch_active = _t(np.asarray(act, dtype=np.int32), torch.int32)
        nzs = np.asarray(nzs, dtype=np.int64)
        self.off_l_np = np.concatenate([[0], np.cumsum(nzs)[:-1]]).astype(np.int64)
        self.off_p_np = np.concatenate([[0], np.cumsum((nzs + 1) * self.nbin)[:-1]]).astype(np.int64)
        self.off_l = _t(self.off_l_np, torch.int64)
        self.off_p = _t(self.off_p_np, torch.int64)
        self.dz = _t(np.concatenate([np.asarray(d, dtype=np.float64) for d in dzs]))
        self.plen = int(((nzs + 1) * self.nbin).sum())
        self.llen = int(nzs.sum())
        self.n_br = n_br
        # temperature-dependent cross sections, per column at off_l * n * nbin
        aT_idx = np.asarray(ph.get("absp_T_idx", np.zeros(0)), dtype=np.int32).reshape(-1)
        bT_re = np.asarray(ph.get("branch_T_re_idx", np.zeros(0)), dtype=np.int32).reshape(-1)
        bT_act = np.asarray(ph.get("branch_T_active", np.ones(bT_re.size)), dtype=np.int32).reshape(-1)
        self.n_absT, self.n_brT = int(aT_idx.size), int(bT_re.size)
        tcross = list(tcross) if tcross is not None else [None] * len(nzs)
        if (self.n_absT or self.n_brT) and (len(tcross) != len(nzs) or any(t is None for t in tcross)):
            raise ValueError("temperature-dependent cross sections need one set per column")

        def per_col(key, n):
            parts = []
            for t, nz in zip(tcross, nzs):
                a = np.asarray(t[key], dtype=np.float64)
                if a.shape != (n, int(nz), self.nbin):
                    raise ValueError(f"{key} must be ({n}, {int(nz)}, {self.nbin}), got {a.shape}")
                parts.append(a.reshape(-1))
            return _t(np.concatenate(parts))
        self.absp_T_idx = _t(aT_idx, torch.int32)
        self.absp_T_cross = per_col("absp_T_cross", self.n_absT) if self.n_absT else _t(np.zeros(1))
        self.branch_T_re_idx = _t(bT_re, torch.int32)
        self.branch_T_active = _t(bT_act, torch.int32)
        self.cross_J_T = per_col("cross_J_T", self.n_brT) if self.n_brT else _t(np.zeros(1))
        c = PhotoSetupC()
        c.nbin, c.n_abs, c.n_scat, c.n_br = self.nbin, int(self.absp_idx.numel()), \
            int(self.scat_idx.numel()), n_br
        c.din12 = int(ph["din12_indx"])
        ag0 = float(ph.get("ag0", 0.0))
        c.ag0_zero = int(ag0 == 0.0)
        c.dbin1, c.dbin2 = float(ph["dbin1"]), float(ph["dbin2"])
        c.mu_zenith, c.edd, c.ag0 = float(ph["mu_zenith"]), float(ph["edd"]), ag0
        c.hc, c.f_diurnal, c.flux_atol = float(ph["hc"]), float(ph.get("f_diurnal", 1.0)), float(flux_atol)
        for f in ("bins", "sflux_top", "absp_cross", "scat_cross", "cross_J", "dz", "absp_idx",
                  "scat_idx", "branch_re_idx", "branch_active", "off_p", "off_l", "absp_T_cross",
                  "cross_J_T", "absp_T_idx", "branch_T_re_idx", "branch_T_active"):
            setattr(c, f, getattr(self, f).data_ptr())
        c.n_absT, c.n_brT = self.n_absT, self.n_brT
        self.c = c


def _ph_dict(p):
    """PhotoSetup dataclass (synth) -> dict of fields."""
    if isinstance(p, dict):
        return p
    return {f: getattr(p, f) for f in p.__dataclass_fields__}


class _State:
    def __init__(self, plen, llen, n_br, ncols, n_brT=0):
        z = lambda n: torch.zeros(n, dtype=F64, device=device())
        self.tau, self.w0, self.sflux, self.dflux_d, self.dflux_u, self.aflux = (z(plen) for _ in range(6))
        self.J = z(max(llen * n_br, 1))
        self.J_T = z(max(llen * n_brT, 1))
        self.aflux_change = z(ncols)

    def c(self, due=None, status=None):
        s = PhotoStateC()
        for f in ("tau", "w0", "sflux", "dflux_d", "dflux_u", "aflux", "J", "aflux_change", "J_T"):
            setattr(s, f, getattr(self, f).data_ptr())
        s.due = None if due is None else due.data_ptr()
        s.status = None if status is None else status.data_ptr()
        return s


class PhotoBatch:
    """In-loop photolysis for a SteadyStateRunner (outer_loop.py:253-352)."""

    def __init__(self, runner, columns):
        phs = [c.photo for c in columns]
        if any(p is None for p in phs):
            raise ValueError("every column of a photolysis batch needs a PhotoSetup")
        ph0 = _ph_dict(phs[0])
        for p in phs[1:]:
            if p is not phs[0]:
                raise ValueError("the columns of one batch must share one PhotoSetup")
        self.runner = runner
        tcross = [getattr(c, "photo_T", None) for c in columns]
        self.setup = _Setup(ph0, [c.nz for c in columns], [c.dz for c in columns],
                            flux_atol=1.0, branch_re_idx=ph0["branch_re_idx"],
                            branch_active=ph0["branch_active"],
                            tcross=tcross if any(t is not None for t in tcross) else None)
        self.state = _State(self.setup.plen, self.setup.llen, self.setup.n_br, len(columns),
                            self.setup.n_brT)
        self.cstate = self.state.c(due=runner.photo_due, status=runner.status)

    def reset(self):
        for t in (self.state.tau, self.state.w0, self.state.sflux, self.state.dflux_d,
                  self.state.dflux_u, self.state.aflux, self.state.J, self.state.J_T,
                  self.state.aflux_change):
            t.zero_()

    def state_tensors(self) -> dict:
        """The device arrays the next photolysis update reads or keeps."""
        return {f: getattr(self.state, f) for f in ("tau", "w0", "sflux", "dflux_d", "dflux_u",
                                                     "aflux", "J", "J_T", "aflux_change")}

    def update(self):
        r = self.runner
        _lib.check(_lib.load().vk_photo_f64(r.batch.handle, 7, C.byref(self.setup.c),
                                            C.byref(self.cstate), r.y.data_ptr(), r.k.data_ptr(),
                                            _lib.stream_ptr()), "vk_photo_f64")


# ---------------------------------------------------------------------------
# reference-named single-column functions
# ---------------------------------------------------------------------------

def _one_col_batch(nz, ni):
    return Batch.uniform_of(1, nz, _solver_network(ni))


def _run(parts, nz, ni, ph, dz, y=None, k=None, state=None, branch=None, tcross=None):
    b = _one_col_batch(nz, ni)
    setup = _Setup(ph, [nz], [dz], tcross=None if tcross is None else [tcross], **(branch or {}))
    st = state or _State(setup.plen, setup.llen, setup.n_br, 1, setup.n_brT)
    cst = st.c()
    _lib.check(_lib.load().vk_photo_f64(b.handle, parts, C.byref(setup.c), C.byref(cst),
                                        None if y is None else y.data_ptr(),
                                        None if k is None else k.data_ptr(), _lib.stream_ptr()),
               "vk_photo_f64")
    return setup, st


def _photo_dict(photo: PhotoData, nbin, bins=None, extra=None):
    d = dict(bins=np.zeros(nbin) if bins is None else bins, sflux_top=np.zeros(nbin),
             absp_idx=np.asarray(photo.absp_idx), absp_cross=np.asarray(photo.absp_cross),
             scat_idx=np.asarray(photo.scat_idx), scat_cross=np.asarray(photo.scat_cross),
             absp_T_idx=np.asarray(photo.absp_T_idx).reshape(-1),
             cross_J=np.zeros((0, nbin)), din12_indx=1, dbin1=1.0, dbin2=1.0,
             mu_zenith=1.0, edd=0.5, hc=1.0)
    if extra:
        d.update(extra)
    return d


def _tcross(photo: PhotoData, nz, nbin):
    """The T-dependent absorber cross sections of one column, or None."""
    n = np.asarray(photo.absp_T_idx).size
    if n == 0:
        return None
    return {"absp_T_cross": np.asarray(photo.absp_T_cross, dtype=np.float64).reshape(n, nz, nbin)}


def photo_data_from_static(static, species_list) -> PhotoData:
    """PhotoData from the static photolysis inputs (photo.py:35-81): the
    absorbers without a temperature-dependent cross section, the
    temperature-dependent ones (their (n_T, nz, nbin) cross sections passed
    through) and the scatterers, as species indices into species_list."""
    names = list(species_list)
    t_names = list(static.absp_T_sp)
    t_set = set(t_names)
    nbin = int(static.nbin)
    plain = [n for n in static.absp_sp if n not in t_set]
    row_of = {n: i for i, n in enumerate(static.absp_sp)}
    absp_cross = np.asarray(static.absp_cross, dtype=np.float64)
    return PhotoData(
        absp_idx=np.array([names.index(n) for n in plain], dtype=np.int64),
        absp_cross=(absp_cross[[row_of[n] for n in plain]] if plain
                    else np.zeros((0, nbin), dtype=np.float64)),
        absp_T_idx=np.array([names.index(n) for n in t_names], dtype=np.int64),
        absp_T_cross=np.asarray(static.absp_T_cross, dtype=np.float64),
        scat_idx=np.array([names.index(n) for n in static.scat_sp], dtype=np.int64),
        scat_cross=(np.asarray(static.scat_cross, dtype=np.float64) if len(static.scat_sp)
                    else np.zeros((0, nbin), dtype=np.float64)))


def compute_tau_jax(y, dz, photo: PhotoData) -> torch.Tensor:
    """tau (nz+1, nbin), top-down cumulative; photo.py:84-123."""
    y = torch.as_tensor(y, dtype=F64, device=device()).contiguous()
    nz, ni = y.shape
    nbin = np.asarray(photo.absp_cross).shape[-1]
    _, st = _run(1, nz, ni, _photo_dict(photo, nbin), np.asarray(dz), y=y,
                 tcross=_tcross(photo, nz, nbin))
    return st.tau.view(nz + 1, nbin).clone()


def compute_flux_jax(tau, sflux_top, ymix, photo: PhotoData, bins, mu_zenith, edd, ag0, hc,
                     dflux_u_prev, ag0_is_zero=True):
    """(aflux, sflux, dflux_d, dflux_u); photo.py:126-245."""
    tau = torch.as_tensor(tau, dtype=F64, device=device())
    nz = tau.shape[0] - 1
    nbin = tau.shape[1]
    ymix = torch.as_tensor(ymix, dtype=F64, device=device()).contiguous()
    ni = ymix.shape[1]
    ph = _photo_dict(photo, nbin, np.asarray(bins), dict(
        sflux_top=np.asarray(sflux_top), mu_zenith=float(mu_zenith), edd=float(edd),
        ag0=0.0 if ag0_is_zero else float(ag0), hc=float(hc)))
    tc = _tcross(photo, nz, nbin)
    st = _State((nz + 1) * nbin, nz, 0, 1)
    # w0 from ymix (tau from this pass is replaced by the caller's tau)
    _run(1, nz, ni, ph, np.ones(nz), y=ymix, state=st, tcross=tc)
    st.tau.copy_(tau.reshape(-1))
    st.dflux_u.copy_(torch.as_tensor(dflux_u_prev, dtype=F64, device=device()).reshape(-1))
    _run(2, nz, ni, ph, np.ones(nz), state=st, tcross=tc)
    v = lambda t, n: t.view(n, nbin).clone()
    return v(st.aflux, nz + 1)[:nz], v(st.sflux, nz + 1), v(st.dflux_d, nz + 1), v(st.dflux_u, nz + 1)


def compute_J_jax_flat(aflux, cross_J, cross_J_T, din12_indx, dbin1, dbin2):
    """(J_br (n_br, nz), J_br_T (n_brT, nz)); photo.py:309-351, 377-392.  The
    (n_br x nbin) . (nbin x nz) contraction runs on FP64 MFMA."""
    aflux = torch.as_tensor(aflux, dtype=F64, device=device())
    nz, nbin = aflux.shape
    cj = np.asarray(cross_J, dtype=np.float64).reshape(-1, nbin)
    cjt = np.asarray(cross_J_T, dtype=np.float64)
    n_brT = cjt.shape[0] if cjt.size else 0
    ph = dict(bins=np.ones(nbin), sflux_top=np.zeros(nbin), absp_idx=np.zeros(0, np.int32),
              absp_cross=np.zeros((0, nbin)), scat_idx=np.zeros(0, np.int32),
              scat_cross=np.zeros((0, nbin)), cross_J=cj, din12_indx=int(din12_indx),
              dbin1=float(dbin1), dbin2=float(dbin2), mu_zenith=1.0, edd=0.5, hc=1.0,
              branch_T_re_idx=np.zeros(n_brT, np.int32), branch_T_active=np.zeros(n_brT, np.int32))
    st = _State((nz + 1) * nbin, nz, cj.shape[0], 1, n_brT)
    st.aflux[: nz * nbin].copy_(aflux.reshape(-1))
    _run(4, nz, 1, ph, np.ones(nz), state=st,
         tcross={"cross_J_T": cjt.reshape(n_brT, nz, nbin)} if n_brT else None)
    J = st.J[: cj.shape[0] * nz].view(cj.shape[0], nz).clone()
    JT = st.J_T[: n_brT * nz].view(n_brT, nz).clone()
    return J, JT


def update_k_with_J(k_arr, J_br, J_br_T, branch_re_idx, branch_active, branch_T_re_idx,
                    branch_T_active, f_diurnal):
    """k_arr[re_idx] = J * f_diurnal for active branches; photo.py:455-478."""
    k = torch.as_tensor(k_arr, dtype=F64, device=device()).clone()
    J = torch.as_tensor(J_br, dtype=F64, device=device())
    idx = torch.as_tensor(np.asarray(branch_re_idx), device=device()).long()
    act = torch.as_tensor(np.asarray(branch_active), device=device()).bool()
    if act.any():
        k[idx[act]] = J[act] * float(f_diurnal)
    if np.asarray(J_br_T).size and np.asarray(branch_T_active).any():
        JT = torch.as_tensor(J_br_T, dtype=F64, device=device())
        idxT = torch.as_tensor(np.asarray(branch_T_re_idx), device=device()).long()
        actT = torch.as_tensor(np.asarray(branch_T_active), device=device()).bool()
        k[idxT[actT]] = JT[actT] * float(f_diurnal)
    return k
