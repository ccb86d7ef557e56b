"""Device networks, batch tables and packed buffers for the C ABI.

A ``Batch`` owns one ``vk_batch`` handle (column descriptor table on the
device) and knows the packed offsets of every per-column array, so that
host code can pack NumPy/torch inputs into the flat device buffers the
kernels read and unpack results again.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._types import ATM_FLAGS, ATM_RECORD_FIELDS, AtmStatic, NetworkArrays

F64 = torch.float64


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("vulcan_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


class DeviceNetwork:
    """A reaction network uploaded once (replaces chem.to_jax, chem.py:70-80)."""

    def __init__(self, net: NetworkArrays):
        lib = _lib.load()
        self.ni = int(net.ni)
        self.nr = int(net.nr)
        r_idx = np.ascontiguousarray(_np(net.reactant_idx), dtype=np.int64)
        self.T = int(r_idx.shape[1])
        p_idx = np.ascontiguousarray(_np(net.product_idx), dtype=np.int64)
        r_st = np.ascontiguousarray(_np(net.reactant_stoich), dtype=np.float64)
        p_st = np.ascontiguousarray(_np(net.product_stoich), dtype=np.float64)
        is3 = np.ascontiguousarray(_np(net.is_three_body), dtype=np.uint8)
        if r_idx.shape != (self.nr + 1, self.T):
            raise ValueError(f"reactant_idx has shape {r_idx.shape}, expected ({self.nr + 1}, T)")
        h = C.c_void_p()
        rc = lib.vk_network_create(self.ni, self.nr, self.T, r_idx.ctypes.data, p_idx.ctypes.data,
                                   r_st.ctypes.data, p_st.ctypes.data, is3.ctypes.data,
                                   C.byref(h))
        _lib.check(rc, "vk_network_create")
        self.handle = h
        self._fin = weakref.finalize(self, lib.vk_network_destroy, h)

    @property
    def S(self) -> int:
        return self.nr + 1


_NET_CACHE: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()
_NET_BY_ID: dict = {}


def device_network(net) -> DeviceNetwork:
    """Upload (once) and return the device copy of a network."""
    if isinstance(net, DeviceNetwork):
        return net
    key = id(net)
    hit = _NET_BY_ID.get(key)
    if hit is not None and hit[0] is net:
        return hit[1]
    dn = DeviceNetwork(net)
    _NET_BY_ID[key] = (net, dn)
    return dn


def block_doubles(ni: int) -> int:
    sl = (ni + 15) // 16
    return sl * sl * 256


def atm_record_layout(nz: int, ni: int):
    lib = _lib.load()
    off = (C.c_int64 * 18)()
    total = lib.vk_atm_record_layout(nz, ni, off)
    return int(total), list(off)


class Batch:
    """Packed layout of ncols columns plus its device descriptor table."""

    def __init__(self, nzs: Sequence[int], nets: Sequence[DeviceNetwork],
                 net_ids: Sequence[int] | None = None):
        lib = _lib.load()
        self.nets = list(nets)
        ncols = len(nzs)
        if net_ids is None:
            if len(self.nets) != 1:
                raise ValueError("net_ids is required with more than one network")
            net_ids = [0] * ncols
        self.ncols = ncols
        self.nz = np.asarray(nzs, dtype=np.int32)
        self.net_ids = np.asarray(net_ids, dtype=np.int32)
        self.ni = np.array([self.nets[i].ni for i in self.net_ids], dtype=np.int32)
        self.S = np.array([self.nets[i].S for i in self.net_ids], dtype=np.int32)
        sizes_s = self.nz.astype(np.int64) * self.ni
        sizes_k = self.S.astype(np.int64) * self.nz
        sizes_a = np.array([atm_record_layout(int(z), int(i))[0]
                            for z, i in zip(self.nz, self.ni)], dtype=np.int64)
        sizes_b = self.nz.astype(np.int64) * np.array([block_doubles(int(i)) for i in self.ni])
        ex = lambda s: np.concatenate([[0], np.cumsum(s)[:-1]]).astype(np.int64)
        self.off_s, self.off_k, self.off_a, self.off_b = ex(sizes_s), ex(sizes_k), ex(sizes_a), ex(sizes_b)
        self.state_len = int(sizes_s.sum())
        self.k_len = int(sizes_k.sum())
        self.atm_len = int(sizes_a.sum())
        self.blk_len = int(sizes_b.sum())
        self.uniform = (len(set(self.nz.tolist())) == 1 and len(set(self.net_ids.tolist())) == 1)
        h = C.c_void_p()
        if self.uniform:
            rc = lib.vk_batch_uniform(ncols, int(self.nz[0]), self.nets[int(self.net_ids[0])].handle,
                                      C.byref(h))
        else:
            handles = (C.c_void_p * len(self.nets))(*[n.handle for n in self.nets])
            rc = lib.vk_batch_create(ncols, self.nz.ctypes.data, self.ni.ctypes.data,
                                     self.S.ctypes.data, self.net_ids.ctypes.data,
                                     self.off_s.ctypes.data, self.off_k.ctypes.data,
                                     self.off_a.ctypes.data, self.off_b.ctypes.data, handles,
                                     len(self.nets), C.byref(h))
        _lib.check(rc, "vk_batch_create")
        self.handle = h
        self._fin = weakref.finalize(self, lib.vk_batch_destroy, h)
        self._ws = None

    @classmethod
    def uniform_of(cls, B: int, nz: int, net: DeviceNetwork) -> "Batch":
        return cls([nz] * B, [net])

    # ---- packing helpers -------------------------------------------------
    def pack_state(self, arrays, dev=None) -> torch.Tensor:
        """Concatenate per-column (nz, ni) arrays into one flat device buffer."""
        dev = dev or device()
        parts = [torch.as_tensor(a, dtype=F64).reshape(-1) for a in arrays]
        return torch.cat(parts).to(dev)

    def unpack_state(self, flat: torch.Tensor):
        return [flat[o:o + int(z) * int(i)].reshape(int(z), int(i))
                for o, z, i in zip(self.off_s, self.nz, self.ni)]

    def pack_k(self, ks, dev=None) -> torch.Tensor:
        dev = dev or device()
        return torch.cat([torch.as_tensor(k, dtype=F64).reshape(-1) for k in ks]).to(dev)

    def pack_atm(self, atms: Sequence[AtmStatic], dev=None) -> torch.Tensor:
        dev = dev or device()
        return torch.cat([pack_atm_record(a) for a in atms]).to(dev)

    def workspace(self) -> torch.Tensor:
        lib = _lib.load()
        nbytes = int(lib.vk_ros2_workspace_bytes(self.handle))
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=device())
        return self._ws


def pack_atm_record(atm: AtmStatic) -> torch.Tensor:
    """One packed record (CPU tensor) in the order of vk_atm_record_layout."""
    nz, ni = atm.nz, atm.ni
    parts = []
    for f in ATM_RECORD_FIELDS:
        v = getattr(atm, f)
        t = torch.as_tensor(_np(v), dtype=F64).reshape(-1)
        parts.append(t)
    parts.append(torch.tensor([float(bool(getattr(atm, f))) for f in ATM_FLAGS], dtype=F64))
    rec = torch.cat(parts)
    total, _ = atm_record_layout(nz, ni)
    if rec.numel() != total:
        raise ValueError(f"atmosphere record has {rec.numel()} values, expected {total}")
    return rec


_BATCH_CACHE: dict = {}


def uniform_batch(B: int, nz: int, net) -> Batch:
    """Cached uniform batch (B identical columns) for the functional API."""
    dn = device_network(net)
    key = (B, nz, id(dn))
    b = _BATCH_CACHE.get(key)
    if b is None:
        b = Batch.uniform_of(B, nz, dn)
        if len(_BATCH_CACHE) > 64:
            _BATCH_CACHE.clear()
        _BATCH_CACHE[key] = b
    return b
