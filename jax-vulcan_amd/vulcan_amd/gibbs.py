"""NASA-9 Gibbs energies, equilibrium constants and the reverse-rate fill
on the GPU.

Same names, arguments and meaning as gibbs.py of the reference:
``load_nasa9`` (gibbs.py:16-38, host file reading), ``gibbs_sp_vector``
(gibbs.py:41-74), ``K_eq_array`` (gibbs.py:77-128), ``fill_reverse_k``
(gibbs.py:131-154) and ``compute_all_k`` (gibbs.py:157-170).  Results are
float64 tensors on the GPU.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import torch

from .core import F64, device
from .rates import (RATES_BUILD, RATES_FILL, RATES_GIBBS, RATES_KEQ, _grid, _rows,
                    device_rates, remove_mask)

__all__ = ["KB", "CORR", "load_nasa9", "gibbs_sp_vector", "K_eq_array", "fill_reverse_k",
           "compute_all_k"]

KB = 1.38064852e-16      # erg/K (phy_const.py:3)
P0 = 1.0e6               # gibbs.py:12
CORR = KB / P0           # gibbs.py:13


# a decimal number with an optional exponent, Fortran's D exponent included
_NUMBER = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?")


def _nasa9_fit(path: Path) -> np.ndarray:
    """The (2, 10) coefficients of one NASA-9 fit file: the low and the high
    temperature range, ten numbers each (a1..a7, an unused slot, b1, b2), in
    file order; anything after the first twenty numbers is ignored."""
    numbers = [float(tok.translate(_D_TO_E)) for tok in _NUMBER.findall(path.read_text())]
    if len(numbers) < 20:
        raise ValueError(f"{path}: two ranges of ten NASA-9 coefficients expected, "
                         f"{len(numbers)} numbers found")
    return np.asarray(numbers[:20], dtype=np.float64).reshape(2, 10)


_D_TO_E = str.maketrans("Dd", "Ee")


def load_nasa9(species, thermo_dir):
    """(coeffs (ni, 2, 10), present (ni,)) for the species names, read from
    the files ``<thermo_dir>/NASA9/<name>.txt``; a species without a file
    keeps zero coefficients and present False (gibbs.py:16-38)."""
    folder = Path(thermo_dir) / "NASA9"
    names = list(species)
    fits = [_nasa9_fit(folder / f"{n}.txt") if (folder / f"{n}.txt").is_file() else None
            for n in names]
    present = np.array([f is not None for f in fits], dtype=bool)
    coeffs = np.zeros((len(names), 2, 10), dtype=np.float64)
    for j, f in enumerate(fits):
        if f is not None:
            coeffs[j] = f
    return coeffs, present


def gibbs_sp_vector(coeffs, T) -> torch.Tensor:
    """g/(RT) per species and layer, (ni, nz): H/RT - S/R of the NASA-9 fit,
    the low range below 1000 K (gibbs.py:41-74)."""
    T = torch.as_tensor(T, dtype=F64)
    if T.dim() != 1:
        raise ValueError(f"T must be 1-D, got shape {tuple(T.shape)}")
    return device_rates(None, coeffs).eval(RATES_GIBBS, T)


def K_eq_array(net, gibbs_sp, T) -> torch.Tensor:
    """Equilibrium constants of the forward rows, (nr+1, nz); 1 on rows
    without a thermal reverse (gibbs.py:77-128)."""
    T = torch.as_tensor(T, dtype=F64)
    nz = int(T.shape[0])
    g = torch.as_tensor(gibbs_sp, dtype=F64, device=device()).contiguous()
    if tuple(g.shape) != (int(net.ni), nz):
        raise ValueError(f"gibbs_sp has shape {tuple(g.shape)}, expected ({int(net.ni)}, {nz})")
    return device_rates(net).eval(RATES_KEQ, T, in_a=g)


def fill_reverse_k(net, k, K_eq, remove_list=None) -> torch.Tensor:
    """Reverse rows k_f / K (0 where K <= 0), removal paired with the
    partner row, rows past stop_rev_indx at odd offsets zeroed
    (gibbs.py:131-154)."""
    k = torch.as_tensor(k, dtype=F64, device=device()).contiguous()
    nz = int(k.shape[1])
    k = _rows(net, k, nz)
    K = _rows(net, K_eq, nz)
    T = torch.ones(nz, dtype=F64)
    mask = remove_mask(net, remove_list, paired=True)
    return device_rates(net, None, mask).eval(RATES_FILL, T, in_a=k, in_b=K)


def compute_all_k(net, T, M, nasa9_coeffs, remove_list=None) -> torch.Tensor:
    """Forward rates plus the K_eq reverse fill, paired removal
    (gibbs.py:157-170), in one launch."""
    T, M = _grid(T, M)
    mask = remove_mask(net, remove_list, paired=True)
    return device_rates(net, nasa9_coeffs, mask).eval(RATES_BUILD, T, M)
