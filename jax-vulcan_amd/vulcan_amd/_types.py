"""Host-side containers shared by the API modules.

``NetworkArrays`` keeps the field names and meaning of chem.NetworkArrays
(chem.py:28-67); ``AtmStatic`` keeps those of jax_step.AtmStatic
(jax_step.py:26-49).  Arrays may be NumPy arrays or torch tensors.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np


class NetworkArrays(NamedTuple):
    """Padded reaction tables (network.py:366-451).

    Rows are 1-based (row 0 unused); PAD = ni marks an empty slot; a
    forward reaction at an odd row is followed by its reverse.
    """
    ni: int
    nr: int
    reactant_idx: np.ndarray     # (nr+1, T) int64
    product_idx: np.ndarray      # (nr+1, T) int64
    reactant_stoich: np.ndarray  # (nr+1, T) float64
    product_stoich: np.ndarray   # (nr+1, T) float64
    is_three_body: np.ndarray    # (nr+1,) bool


class RateNetwork(NamedTuple):
    """What the rate build reads of network.Network (network.py:44-95).

    Per-row flags and Arrhenius columns are (nr+1,) arrays indexed like the
    reaction tables; ``Rf`` maps a forward row to its equation text.  Any
    object with these attributes (the reference's ``Network`` included) is
    accepted where a ``RateNetwork`` is expected.
    """
    species: tuple
    ni: int
    nr: int
    reactant_idx: np.ndarray     # (nr+1, T) int64, PAD = ni
    product_idx: np.ndarray      # (nr+1, T) int64
    reactant_stoich: np.ndarray  # (nr+1, T) float64
    product_stoich: np.ndarray   # (nr+1, T) float64
    a: np.ndarray
    n: np.ndarray
    E: np.ndarray
    a_inf: np.ndarray
    n_inf: np.ndarray
    E_inf: np.ndarray
    is_forward: np.ndarray
    has_kinf: np.ndarray
    is_special: np.ndarray
    is_photo: np.ndarray
    is_ion: np.ndarray
    is_conden: np.ndarray
    is_radiative: np.ndarray
    stop_rev_indx: int
    Rf: dict


class AtmStatic(NamedTuple):
    """Per-column atmosphere held fixed through a step."""
    Kzz: np.ndarray       # (nz-1,)
    Dzz: np.ndarray       # (nz-1, ni)
    dzi: np.ndarray       # (nz-1,)
    vz: np.ndarray        # (nz-1,)
    Hpi: np.ndarray       # (nz-1,)
    Ti: np.ndarray        # (nz-1,)
    Tco: np.ndarray       # (nz,)
    g: np.ndarray         # (nz,)
    ms: np.ndarray        # (ni,)
    alpha: np.ndarray     # (ni,)
    M: np.ndarray         # (nz,)
    vm: np.ndarray        # (nz, ni)
    vs: np.ndarray        # (nz-1, ni)
    top_flux: np.ndarray  # (ni,)
    bot_flux: np.ndarray  # (ni,)
    bot_vdep: np.ndarray  # (ni,)
    gas_indx_mask: np.ndarray  # (ni,) bool
    use_vm_mol: bool = False
    use_settling: bool = False
    use_topflux: bool = False
    use_botflux: bool = False

    @property
    def nz(self) -> int:
        return int(self.Tco.shape[0])

    @property
    def ni(self) -> int:
        return int(self.ms.shape[0])


# order of the fields inside one packed atmosphere record (vk_atm_record_layout)
ATM_RECORD_FIELDS = (
    "Kzz", "dzi", "vz", "Hpi", "Ti", "Tco", "g", "M", "ms", "alpha", "top_flux",
    "bot_flux", "bot_vdep", "gas_indx_mask", "Dzz", "vm", "vs",
)
ATM_FLAGS = ("use_vm_mol", "use_settling", "use_topflux", "use_botflux")
