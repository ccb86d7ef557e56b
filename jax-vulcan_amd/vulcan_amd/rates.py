"""Rate construction k(T, M) on the GPU.

Same names, arguments and meaning as rates.py of the reference:
``compute_forward_k`` (rates.py:61-122), ``apply_lowT_caps``
(rates.py:142-167), ``apply_remove_list`` (rates.py:170-182),
``build_rate_array`` (rates.py:185-208), ``k_dict_from_array`` and
``k_array_from_dict`` (rates.py:125-139).  Results are float64 tensors of
shape (nr+1, nz) on the GPU, row 0 unused, as in the reference.

The host reads the network once (``rate_tables``): the forward form of
each row (modified Arrhenius, Lindemann falloff, the hard-coded Troe
reaction, or zero), which of the three low-T caps it takes, and whether
its reverse is filled from the equilibrium constant.  The device does the
arithmetic: ``build_rate_array`` is a single launch, ``build_rate_batch``
builds the rates of a whole batch of columns (each with its own T(z),
M(z)) in one launch, and ``build_rate_array_differentiable`` carries
gradients with respect to T and M through the device's vector-Jacobian
product.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Iterable, Sequence

import numpy as np
import torch

from . import _lib
from ._types import RateNetwork
from .core import F64, Batch, device

__all__ = ["RateNetwork", "rate_tables", "DeviceRates", "compute_forward_k", "apply_lowT_caps",
           "apply_remove_list", "build_rate_array", "build_rate_array_differentiable",
           "build_rate_batch", "k_dict_from_array", "k_array_from_dict"]

# forms, caps and modes of include/vulcan_ros2.h
KF_ZERO, KF_ARRHENIUS, KF_LINDEMANN, KF_TROE_OH_CH3 = 0, 1, 2, 3
LOWT_NONE, LOWT_CH3, LOWT_C2H4, LOWT_C2H5 = 0, 1, 2, 3
RATES_FORWARD, RATES_CAPS, RATES_GIBBS, RATES_KEQ, RATES_FILL, RATES_BUILD = range(6)

SPECIAL_OH_CH3 = "OH + CH3 + M -> CH3OH + M"                  # rates.py:27
LOWT_RXNS = {"H + CH3 + M -> CH4 + M": LOWT_CH3,              # rates.py:31-33
             "H + C2H4 + M -> C2H5 + M": LOWT_C2H4,
             "H + C2H5 + M -> C2H6 + M": LOWT_C2H5}


def _np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def rate_tables(net) -> dict:
    """Per-row classification of a network for the device rate build.

    form: which expression compute_forward_k takes for the row
    (rates.py:102-120; special rows test the stripped equation text, the
    caps the exact text, rates.py:109 and 154); lowt: the cap of
    apply_lowT_caps (rates.py:151-165); has_rev: rows whose K_eq is not the
    sentinel 1 (gibbs.py:96-105)."""
    nr = int(net.nr)
    S = nr + 1
    form = np.zeros(S, np.int32)
    lowt = np.zeros(S, np.int32)
    has_rev = np.zeros(S, np.int32)
    par = np.zeros((S, 6), np.float64)
    fwd = _np(net.is_forward)
    flags = [_np(getattr(net, f)) for f in ("is_photo", "is_ion", "is_conden", "is_radiative")]
    a, n, E = _np(net.a), _np(net.n), _np(net.E)
    a_inf, n_inf, E_inf = _np(net.a_inf), _np(net.n_inf), _np(net.E_inf)
    kinf, special = _np(net.has_kinf), _np(net.is_special)
    stop_rev = int(net.stop_rev_indx)
    for i in range(1, nr + 1, 2):
        if not fwd[i]:
            continue
        rf = net.Rf.get(i, "")
        lowt[i] = LOWT_RXNS.get(rf, LOWT_NONE)
        if any(f[i] for f in flags):
            continue
        if i + 1 < stop_rev:
            has_rev[i] = 1
        if special[i]:
            if rf.strip() == SPECIAL_OH_CH3:
                form[i] = KF_TROE_OH_CH3
            elif a[i] != 0.0:
                form[i] = KF_ARRHENIUS
                par[i, :3] = (a[i], n[i], E[i])
            continue
        if kinf[i]:
            form[i] = KF_LINDEMANN
            par[i] = (a[i], n[i], E[i], a_inf[i], n_inf[i], E_inf[i])
        else:
            form[i] = KF_ARRHENIUS
            par[i, :3] = (a[i], n[i], E[i])
    return dict(form=form, lowt=lowt, has_rev=has_rev, par=par, stop_rev=stop_rev)


def remove_mask(net, remove_list, paired: bool) -> np.ndarray | None:
    """Rows zeroed by a remove list: literally the listed rows
    (rates.py:178-181), or, paired, both rows of every reversible pair one
    of whose rows is listed (gibbs.py:141-146)."""
    if not remove_list:
        return None
    nr = int(net.nr)
    m = np.zeros(nr + 1, np.uint8)
    if paired:
        remove = set(int(i) for i in remove_list)
        for i in range(2, int(net.stop_rev_indx), 2):
            if i in remove or (i - 1) in remove:
                m[i] = m[i - 1] = 1
    else:
        for i in remove_list:
            if 0 <= int(i) <= nr:
                m[int(i)] = 1
    return m if m.any() else None


class DeviceRates:
    """One vk_rates handle: a network's rate tables, optionally NASA-9
    coefficients (ni, 2, 10) and a mask of rows held at zero."""

    def __init__(self, net, nasa9=None, zero_rows=None):
        lib = _lib.load()
        self.ni = int(net.ni) if net is not None else int(np.shape(nasa9)[0])
        self.nr = int(net.nr) if net is not None else 0
        self.has_nasa9 = nasa9 is not None
        keep = []

        def buf(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        if net is not None:
            t = rate_tables(net)
            r_idx = _np(net.reactant_idx)
            self.T = int(r_idx.shape[1])
            args = (buf(t["form"], np.int32), buf(t["lowt"], np.int32), buf(t["has_rev"], np.int32),
                    buf(t["par"], np.float64), buf(r_idx, np.int64),
                    buf(_np(net.reactant_stoich), np.float64), buf(_np(net.product_idx), np.int64),
                    buf(_np(net.product_stoich), np.float64))
            stop_rev = t["stop_rev"]
        else:
            self.T = 1
            args = (None,) * 8
            stop_rev = 1
        if nasa9 is not None:
            c = np.ascontiguousarray(_np(nasa9), dtype=np.float64)
            if c.shape != (self.ni, 2, 10):
                raise ValueError(f"NASA-9 coefficients have shape {c.shape}, "
                                 f"expected ({self.ni}, 2, 10)")
            n9 = buf(c, np.float64)
        else:
            n9 = None
        zr = buf(zero_rows, np.uint8) if zero_rows is not None else None
        h = C.c_void_p()
        _lib.check(lib.vk_rates_create(self.nr, self.ni, self.T, stop_rev, *args, n9, zr,
                                       C.byref(h)), "vk_rates_create")
        self.handle = h
        self._fin = weakref.finalize(self, lib.vk_rates_destroy, h)

    def eval(self, what: int, T, M=None, lowt_on=False, in_a=None, in_b=None,
             out: torch.Tensor | None = None) -> torch.Tensor:
        dev = device()
        T = torch.as_tensor(T, dtype=F64, device=dev).contiguous()
        nz = int(T.shape[0])
        if M is not None:
            M = torch.as_tensor(M, dtype=F64, device=dev).contiguous()
        rows = self.ni if what == RATES_GIBBS else self.nr + 1
        if out is None:
            out = torch.empty((rows, nz), dtype=F64, device=dev)
        elif (tuple(out.shape) != (rows, nz) or out.dtype != F64 or not out.is_contiguous()
              or out.device != T.device):
            raise ValueError(f"out must be a contiguous float64 ({rows}, {nz}) tensor on {dev}")
        _lib.check(_lib.load().vk_rates_eval_f64(
            self.handle, what, nz, T.data_ptr(), _lib.ptr(M), int(bool(lowt_on)),
            _lib.ptr(in_a), _lib.ptr(in_b), out.data_ptr(), _lib.stream_ptr()),
            "vk_rates_eval_f64")
        return out

    def vjp(self, T, M, k_bar, lowt_on=False):
        """(T_bar, M_bar) = k_bar . dk/dT, k_bar . dk/dM of the full build."""
        dev = device()
        T = torch.as_tensor(T, dtype=F64, device=dev).contiguous()
        M = torch.as_tensor(M, dtype=F64, device=dev).contiguous()
        k_bar = torch.as_tensor(k_bar, dtype=F64, device=dev).contiguous()
        nz = int(T.shape[0])
        if tuple(k_bar.shape) != (self.nr + 1, nz):
            raise ValueError(f"k_bar has shape {tuple(k_bar.shape)}, expected "
                             f"({self.nr + 1}, {nz})")
        Tb = torch.empty(nz, dtype=F64, device=dev)
        Mb = torch.empty(nz, dtype=F64, device=dev)
        _lib.check(_lib.load().vk_rates_vjp_f64(
            self.handle, nz, T.data_ptr(), M.data_ptr(), int(bool(lowt_on)), k_bar.data_ptr(),
            Tb.data_ptr(), Mb.data_ptr(), _lib.stream_ptr()), "vk_rates_vjp_f64")
        return Tb, Mb


class _RateBuild(torch.autograd.Function):
    """k = build(T, M) with its T and M cotangents from vk_rates_vjp_f64."""

    @staticmethod
    def forward(ctx, T, M, rates_handle, lowt_on):
        ctx.rates = rates_handle
        ctx.lowt = lowt_on
        ctx.save_for_backward(T, M)
        return rates_handle.eval(RATES_BUILD, T, M, lowt_on=lowt_on)

    @staticmethod
    def backward(ctx, k_bar):
        T, M = ctx.saved_tensors
        Tb, Mb = ctx.rates.vjp(T, M, k_bar, ctx.lowt)
        return Tb.to(T.dtype), Mb.to(M.dtype), None, None


# Handles are cached per (network object, coefficients, mask) so repeated
# builds (an atmosphere refresh, every column of a batch) upload nothing.
_CACHE: dict = {}


def device_rates(net, nasa9=None, zero_rows=None) -> DeviceRates:
    c = None if nasa9 is None else np.ascontiguousarray(_np(nasa9), dtype=np.float64)
    key = (id(net), None if c is None else hash(c.tobytes()),
           None if zero_rows is None else zero_rows.tobytes())
    hit = _CACHE.get(key)
    if hit is not None and hit[0] is net:
        return hit[1]
    if len(_CACHE) > 64:
        _CACHE.clear()
    d = DeviceRates(net, c, zero_rows)
    _CACHE[key] = (net, d)
    return d


def _grid(T, M):
    T = _np(T).astype(np.float64) if not isinstance(T, torch.Tensor) else T
    M = _np(M).astype(np.float64) if not isinstance(M, torch.Tensor) else M
    if tuple(M.shape) != tuple(T.shape) or len(T.shape) != 1:
        raise ValueError(f"T and M must have the same shape; got {tuple(T.shape)} and "
                         f"{tuple(M.shape)}")
    return T, M


def _rows(net, k, nz):
    k = torch.as_tensor(k, dtype=F64, device=device()).contiguous()
    if tuple(k.shape) != (int(net.nr) + 1, nz):
        raise ValueError(f"k has shape {tuple(k.shape)}, expected ({int(net.nr) + 1}, {nz})")
    return k


def compute_forward_k(net, T, M) -> torch.Tensor:
    """Forward rates (nr+1, nz); reverse, photo, ion, condensation and
    radiative rows zero (rates.py:61-122)."""
    T, M = _grid(T, M)
    return device_rates(net).eval(RATES_FORWARD, T, M)


def apply_lowT_caps(net, k, T, M) -> torch.Tensor:
    """The three low-T caps on the forward rows of k (rates.py:142-167)."""
    T, M = _grid(T, M)
    k = _rows(net, k, int(T.shape[0]))
    return device_rates(net).eval(RATES_CAPS, T, M, in_a=k)


def apply_remove_list(net, k, remove_list: Iterable[int] | None) -> torch.Tensor:
    """Zero exactly the listed rows, no pairing (rates.py:170-182)."""
    k = torch.as_tensor(k, dtype=F64, device=device()).clone()
    if not remove_list:
        return k
    rows = [int(i) for i in remove_list if 0 <= int(i) <= int(net.nr)]
    if rows:
        k[torch.tensor(rows, device=k.device)] = 0.0
    return k


def build_rate_array(cfg, net, atm, nasa9_coeffs) -> torch.Tensor:
    """Forward rates, the low-T caps when ``cfg.use_lowT_limit_rates``, the
    reverse fill from NASA-9 K_eq, then ``cfg.remove_list`` without pairing
    (rates.py:185-208), in one launch."""
    T, M = _grid(atm.Tco, atm.M)
    mask = remove_mask(net, getattr(cfg, "remove_list", None), paired=False)
    lowt = bool(getattr(cfg, "use_lowT_limit_rates", False))
    return device_rates(net, nasa9_coeffs, mask).eval(RATES_BUILD, T, M, lowt_on=lowt)


def build_rate_array_differentiable(net, T, M, nasa9_coeffs, use_lowT_limit_rates=False,
                                    remove_list=None) -> torch.Tensor:
    """build_rate_array on torch tensors T, M (nz,) that carries gradients:
    backward returns k_bar . dk/dT and k_bar . dk/dM from the device, exact
    derivatives of the same expressions (no counterpart in the reference,
    whose NumPy rate build is outside its autodiff)."""
    dev = device()
    T = torch.as_tensor(T, dtype=F64, device=dev)
    M = torch.as_tensor(M, dtype=F64, device=dev)
    if T.shape != M.shape or T.dim() != 1:
        raise ValueError(f"T and M must have the same 1-D shape; got {tuple(T.shape)} and "
                         f"{tuple(M.shape)}")
    mask = remove_mask(net, remove_list, paired=False)
    d = device_rates(net, nasa9_coeffs, mask)
    return _RateBuild.apply(T, M, d, bool(use_lowT_limit_rates))


def build_rate_batch(batch: Batch, nets: Sequence, nasa9: Sequence, atm_packed: torch.Tensor,
                     use_lowT_limit_rates: bool = False,
                     remove_lists: Sequence | None = None, out: torch.Tensor | None = None
                     ) -> torch.Tensor:
    """build_rate_array for every column of a batch in one launch.

    ``nets[j]`` / ``nasa9[j]`` describe the batch's j-th network (the order
    of ``batch.nets``); each column reads T and M from its packed atmosphere
    record; the result is the batch's packed k buffer (``batch.k_len``)."""
    if len(nets) != len(batch.nets) or len(nasa9) != len(batch.nets):
        raise ValueError(f"need one rate network and NASA-9 table per batch network "
                         f"({len(batch.nets)})")
    lib = _lib.load()
    handles = []
    for j, (net, c) in enumerate(zip(nets, nasa9)):
        rl = remove_lists[j] if remove_lists is not None else None
        handles.append(device_rates(net, c, remove_mask(net, rl, paired=False)))
    key = tuple(id(h) for h in handles)
    if getattr(batch, "_rates_key", None) != key:
        arr = (C.c_void_p * len(handles))(*[h.handle for h in handles])
        _lib.check(lib.vk_batch_set_rates(batch.handle, arr, len(handles)), "vk_batch_set_rates")
        batch._rates_key = key
        batch._rates_keep = handles
    if out is None:
        out = torch.empty(batch.k_len, dtype=F64, device=device())
    if out.numel() < batch.k_len or atm_packed.numel() < batch.atm_len:
        raise ValueError("k or atmosphere buffer smaller than the batch layout")
    _lib.check(lib.vk_rates_build_batch_f64(batch.handle, int(bool(use_lowT_limit_rates)),
                                            atm_packed.data_ptr(), out.data_ptr(),
                                            _lib.stream_ptr()), "vk_rates_build_batch_f64")
    return out


def k_dict_from_array(net, k_arr) -> dict:
    """{i: k[i]} for i = 1..nr (rates.py:125-130)."""
    return {i: k_arr[i].clone() if isinstance(k_arr, torch.Tensor) else np.asarray(k_arr[i]).copy()
            for i in range(1, int(net.nr) + 1)}


def k_array_from_dict(net, k_dict: dict, nz: int) -> torch.Tensor:
    """(nr+1, nz) from {i: k_i}; keys outside 1..nr ignored (rates.py:133-139)."""
    out = torch.zeros((int(net.nr) + 1, nz), dtype=F64, device=device())
    for i, v in k_dict.items():
        if 1 <= i <= int(net.nr):
            out[i] = torch.as_tensor(v, dtype=F64, device=out.device)
    return out
