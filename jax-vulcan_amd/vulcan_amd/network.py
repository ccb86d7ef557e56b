"""VULCAN reaction-network text files -> the tables the device path reads.

Same name, argument and result as network.parse_network of the reference
(network.py:197-491): a ``Network`` with the reference's fields, forward
reactions at odd slots 1, 3, 5, ... each followed by its reverse slot,
species in order of first appearance, then species named only by the run
configuration.  ``to_arrays`` gives the ``NetworkArrays`` the chemistry
kernels take (chem.to_jax's input, chem.py:70-80); the ``Network`` itself
is what the rate construction reads (``rates.build_rate_array``).

File format (network.py:103-154, 232-256): a reaction line is
``<id> [ A + 2*B + M -> C + D ] <numbers...> <free text>``; section markers
are comment lines starting ``# 3-body`` (Lindemann, with k_inf),
``# 3-body reactions without high-pressure rates``, ``# special``,
``# condensation``, ``# radiative``, ``# photo``, ``# ionisation`` (or
``# ionization``); ``# reverse stops`` marks stop_rev_indx; other ``#``
lines are comments.  Photo and ion lines carry ``<species> <branch>`` after
the bracket.
"""
from __future__ import annotations

import importlib
import re
from dataclasses import dataclass
from pathlib import Path
from typing import Iterable, Optional

import numpy as np

from ._types import NetworkArrays

__all__ = ["Network", "parse_network", "to_arrays", "summarize", "load_compositions"]

SECTIONS = ("two_body", "three_body_kinf", "three_body_no_kinf", "special", "conden",
            "radiative", "photo", "ion")

# (marker prefix, section), checked in this order: the longer 3-body marker
# before the plain one (network.py:133-154).
_MARKERS = (
    ("# 3-body reactions without high-pressure rates", "three_body_no_kinf"),
    ("# 3-body", "three_body_kinf"),
    ("# special", "special"),
    ("# condensation", "conden"),
    ("# radiative", "radiative"),
    ("# photo", "photo"),
    ("# ionisation", "ion"),
    ("# ionization", "ion"),
)
_LINE = re.compile(r"^\s*(\d+)\s*\[\s*([^\]]+)\s*\]\s*(.*)$")
# configuration fields whose species join the species list (network.py:157-194)
_CFG_SPECIES_FIELDS = ("diff_esc", "scat_sp", "condense_sp", "non_gas_sp", "fix_species",
                       "use_relax", "conver_ignore", "plot_spec", "T_cross_sp")


@dataclass(frozen=True)
class Network:
    """A parsed network; arrays are indexed by slot (row 0 unused)."""
    species: tuple
    species_idx: dict
    ni: int
    nr: int
    reactant_idx: np.ndarray      # (nr+1, T) int64, PAD = ni
    product_idx: np.ndarray
    reactant_stoich: np.ndarray   # (nr+1, T) float64
    product_stoich: np.ndarray
    a: np.ndarray                 # (nr+1,) Arrhenius columns of forward slots
    n: np.ndarray
    E: np.ndarray
    a_inf: np.ndarray
    n_inf: np.ndarray
    E_inf: np.ndarray
    is_forward: np.ndarray        # (nr+1,) bool
    is_three_body: np.ndarray
    has_kinf: np.ndarray
    is_special: np.ndarray
    is_conden: np.ndarray
    is_radiative: np.ndarray
    is_photo: np.ndarray
    is_ion: np.ndarray
    stop_rev_indx: int
    conden_indx: int
    radiative_indx: int
    photo_indx: int
    ion_indx: int
    photo_sp: tuple
    pho_rate_index: dict          # (species, branch) -> forward slot
    n_branch: dict
    ion_sp: tuple
    ion_rate_index: dict
    ion_branch: dict
    Rf: dict                      # forward slot -> equation text
    Rindx: dict                   # forward slot -> id in the file
    network_path: str

    def species_index(self, sp: str) -> int:
        return self.species_idx[sp]


def _term(text: str):
    """'2*H' -> (2, 'H'); 'H' -> (1, 'H')."""
    if "*" in text:
        c, sp = text.split("*", 1)
        return int(c), sp
    return 1, text


def _side(text: str):
    return [_term(t.strip()) for t in text.split("+") if t.strip()]


def _section_of(line: str) -> Optional[str]:
    s = line.lstrip()
    for prefix, sec in _MARKERS:
        if s.startswith(prefix):
            return sec
    return None


def _cfg_species() -> list:
    """Species named by an importable ``vulcan_cfg`` (network.py:157-194)."""
    try:
        cfg = importlib.import_module("vulcan_cfg")
    except Exception:
        return []
    out = []

    def add(values):
        for v in values or ():
            if isinstance(v, str) and v not in out:
                out.append(v)

    for name in ("const_mix", "use_fix_sp_bot"):
        d = getattr(cfg, name, {})
        if isinstance(d, dict):
            add(d.keys())
    for name in _CFG_SPECIES_FIELDS:
        add(getattr(cfg, name, []))
    return out


def parse_network(network_path, extra_species: Optional[Iterable[str]] = None) -> Network:
    """Read a VULCAN network file (network.py:197-491).

    ``extra_species`` are appended to the species list when absent from the
    reactions; None takes them from an importable ``vulcan_cfg`` module, as
    the reference does."""
    network_path = str(network_path)
    species: list = []
    index: dict = {}

    def intern(sp: str) -> int:
        if sp == "M":
            return -1
        if sp not in index:
            index[sp] = len(species)
            species.append(sp)
        return index[sp]

    section = "two_body"
    slot = 1
    first = {}                       # section -> first forward slot
    stop_rev = None
    recs = []
    with open(network_path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            if not line.strip():
                continue
            sec = _section_of(line)
            if sec is not None:
                section = sec
                first.setdefault(sec, slot)
                continue
            s = line.lstrip()
            if s.startswith("# reverse stops"):
                stop_rev = slot
                continue
            if s.startswith("#"):
                continue
            m = _LINE.match(line)
            if m is None:
                continue
            fid, eq, tail = int(m.group(1)), m.group(2).strip(), m.group(3)
            if "->" not in eq:
                raise ValueError(f"reaction {fid}: equation without '->': {eq!r}")
            lhs, rhs = eq.split("->", 1)
            reac, prod = _side(lhs), _side(rhs)
            for _, sp in reac + prod:
                intern(sp)

            def collapse(terms):
                out: dict = {}
                for c, sp in terms:
                    if sp != "M":
                        out[index[sp]] = out.get(index[sp], 0.0) + float(c)
                return out

            words = tail.split()
            nums = []
            for w in words:
                try:
                    nums.append(float(w))
                except ValueError:
                    break
            meta = None
            if section in ("photo", "ion") and len(words) >= 2 and words[1].isdigit():
                meta = (words[0], int(words[1]))
            recs.append(dict(slot=slot, fid=fid, section=section, eq=eq, reac=collapse(reac),
                             prod=collapse(prod),
                             m_reac=any(sp == "M" for _, sp in reac),
                             m_prod=any(sp == "M" for _, sp in prod),
                             nums=nums, meta=meta,
                             target=words[0] if words else eq.split()[0]))
            slot += 2
    extra = _cfg_species() if extra_species is None else list(extra_species)
    for sp in extra:
        intern(sp)

    nr = slot - 1
    ni = len(species)
    if stop_rev is None:
        stop_rev = first.get("photo", nr + 1)
    T = max([1] + [len(r["reac"]) for r in recs] + [len(r["prod"]) for r in recs])
    S = nr + 1
    r_idx = np.full((S, T), ni, np.int64)
    p_idx = np.full((S, T), ni, np.int64)
    r_st = np.zeros((S, T))
    p_st = np.zeros((S, T))
    col = {name: np.zeros(S) for name in ("a", "n", "E", "a_inf", "n_inf", "E_inf")}
    flag = {name: np.zeros(S, bool) for name in ("is_forward", "is_three_body", "has_kinf",
                                                 "is_special", "is_conden", "is_radiative",
                                                 "is_photo", "is_ion")}
    photo_sp, ion_sp = [], []
    pho_idx, n_branch, ion_idx, ion_branch = {}, {}, {}, {}
    for r in recs:
        i, sec = r["slot"], r["section"]
        flag["is_forward"][i] = True
        for k, (sp, st) in enumerate(r["reac"].items()):
            r_idx[i, k], r_st[i, k] = sp, st
            p_idx[i + 1, k], p_st[i + 1, k] = sp, st
        for k, (sp, st) in enumerate(r["prod"].items()):
            p_idx[i, k], p_st[i, k] = sp, st
            r_idx[i + 1, k], r_st[i + 1, k] = sp, st
        flag["is_three_body"][i] = r["m_reac"]
        flag["is_three_body"][i + 1] = r["m_prod"]
        both = {"three_body_kinf": "has_kinf", "special": "is_special", "conden": "is_conden",
                "radiative": "is_radiative"}
        if sec in both:
            flag[both[sec]][i] = flag[both[sec]][i + 1] = True
        if sec == "photo":
            flag["is_photo"][i] = True
        if sec == "ion":
            flag["is_ion"][i] = True
        nums = r["nums"]
        if sec in ("two_body", "three_body_kinf", "three_body_no_kinf") and len(nums) >= 3:
            col["a"][i], col["n"][i], col["E"][i] = nums[:3]
            if sec == "three_body_kinf" and len(nums) >= 6:
                col["a_inf"][i], col["n_inf"][i], col["E_inf"][i] = nums[3:6]
        if sec in ("photo", "ion"):
            names, rate_idx, branches = ((photo_sp, pho_idx, n_branch) if sec == "photo"
                                         else (ion_sp, ion_idx, ion_branch))
            if r["target"] not in names:
                names.append(r["target"])
            if r["meta"] is not None:
                rate_idx[r["meta"]] = i
                branches[r["target"]] = max(branches.get(r["target"], 0), r["meta"][1])
    return Network(
        species=tuple(species), species_idx=dict(index), ni=ni, nr=nr,
        reactant_idx=r_idx, product_idx=p_idx, reactant_stoich=r_st, product_stoich=p_st,
        **col, **flag, stop_rev_indx=int(stop_rev),
        conden_indx=int(first.get("conden", nr + 1)),
        radiative_indx=int(first.get("radiative", nr + 1)),
        photo_indx=int(first.get("photo", nr + 1)), ion_indx=int(first.get("ion", nr + 1)),
        photo_sp=tuple(photo_sp), pho_rate_index=pho_idx, n_branch=n_branch,
        ion_sp=tuple(ion_sp), ion_rate_index=ion_idx, ion_branch=ion_branch,
        Rf={r["slot"]: r["eq"] for r in recs}, Rindx={r["slot"]: r["fid"] for r in recs},
        network_path=network_path)


def to_arrays(net: Network) -> NetworkArrays:
    """The padded tables of the chemistry kernels (chem.NetworkArrays)."""
    return NetworkArrays(ni=net.ni, nr=net.nr, reactant_idx=net.reactant_idx,
                         product_idx=net.product_idx, reactant_stoich=net.reactant_stoich,
                         product_stoich=net.product_stoich, is_three_body=net.is_three_body)


def load_compositions(path, species) -> tuple:
    """(atoms (ni, n_elements), elements, masses (ni,)) from a composition
    table such as thermo/all_compose.txt: a header ``species <elements...>
    mass`` and one row per species."""
    with open(path) as fh:
        head = fh.readline().split()
        rows = {}
        for line in fh:
            w = line.split()
            if len(w) == len(head):
                rows[w[0]] = w[1:]
    elements = head[1:-1]
    atoms = np.zeros((len(species), len(elements)))
    mass = np.zeros(len(species))
    for j, sp in enumerate(species):
        if sp not in rows:
            raise KeyError(f"species {sp!r} missing from {path}")
        atoms[j] = [float(v) for v in rows[sp][:-1]]
        mass[j] = float(rows[sp][-1])
    return atoms, elements, mass


def summarize(net: Network) -> str:
    """A short human-readable description (network.py:494-513)."""
    fw = slice(1, None, 2)
    return "\n".join([
        f"Network: {net.network_path}",
        f"  ni = {net.ni} species",
        f"  nr = {net.nr} reaction slots ({net.nr // 2} forward + reverses)",
        f"  three-body reactions: {int(net.is_three_body[fw].sum())}",
        f"  with k_inf falloff:   {int(net.has_kinf[fw].sum())}",
        f"  special (hardcoded):  {int(net.is_special[fw].sum())}",
        f"  condensation:         {int(net.is_conden[fw].sum())}",
        f"  radiative recomb.:    {int(net.is_radiative[fw].sum())}",
        f"  photo dissociation:   {int(net.is_photo[fw].sum())}",
        f"  photo ionization:     {int(net.is_ion[fw].sum())}",
        f"  stop_rev_indx = {net.stop_rev_indx}",
        f"  photo_indx    = {net.photo_indx}",
    ])
