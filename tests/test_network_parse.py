"""Network ingestion (network.parse_network, network.py:197-491) on the
reference's HD 189733b network file (tests/golden/networks, copied
unchanged by tests/golden/copy_reference_data.py).

The reference's own parse test compares against a sibling code base that is
not shipped (test_network_parse.py:24-30).  Pins here: counts read off the
file with plain text tools (439 reaction lines, section markers at the
listed lines), values of individual lines read by eye, first-appearance
species order, and element balance of every parsed reaction against the
reference's composition table (tests/golden/all_compose.txt), which a
dropped term, a wrong stoichiometric factor or a swapped side breaks.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from vulcan_amd import network as N

GOLD = Path(__file__).resolve().parent / "golden"
NET = GOLD / "networks" / "NCHO_photo_network.txt"
HD189_EXTRA = ["CH4", "O2", "He", "N2", "H2", "H2", "He"]   # const_mix + scat_sp of its config


@pytest.fixture(scope="module")
def net():
    return N.parse_network(NET, extra_species=HD189_EXTRA)


def test_counts(net):
    """439 reaction lines -> 878 slots; 69 species; the photo section (line
    441 of the file) starts at slot 783, right after '# reverse stops'."""
    lines = [l for l in NET.read_text().splitlines()
             if l.strip()[:1].isdigit() and "[" in l and "->" in l]
    assert len(lines) == 439
    assert net.nr == 878 and net.ni == 69
    assert net.stop_rev_indx == 783 and net.photo_indx == 783
    assert net.conden_indx == net.radiative_indx == net.ion_indx == 879
    fw = slice(1, None, 2)
    assert int(net.is_photo[fw].sum()) == 48
    assert int(net.has_kinf[fw].sum()) == 55          # slots 657 .. 765
    assert int(net.is_special[fw].sum()) == 1
    assert int(net.is_ion.sum()) == int(net.is_conden.sum()) == 0
    assert net.reactant_idx.shape == (879, 3)
    assert np.all(net.is_forward[1::2]) and not np.any(net.is_forward[0::2])
    assert len(net.photo_sp) == 27 and not net.ion_sp


def test_lines_read_by_eye(net):
    i = 1       # "1 [ OH + H2 -> H2O + H ] 3.57E-16 1.520 1740.0"
    assert net.Rf[i] == "OH + H2 -> H2O + H" and net.Rindx[i] == 1
    assert (net.a[i], net.n[i], net.E[i]) == (3.57e-16, 1.52, 1740.0)
    assert (net.a[i + 1], net.n[i + 1], net.E[i + 1]) == (0.0, 0.0, 0.0)
    sp = net.species_idx
    assert [sp[s] for s in ("OH", "H2", "H2O", "H")] == [0, 1, 2, 3]
    assert list(net.reactant_idx[i, :2]) == [0, 1] and list(net.product_idx[i, :2]) == [2, 3]
    assert list(net.reactant_idx[i + 1, :2]) == [2, 3] and list(net.product_idx[i + 1, :2]) == [0, 1]
    i = 657     # "H + H + M -> H2 + M  2.70E-31 -0.600 0.0  3.31E-06 -1.000 0.0"
    assert net.Rf[i] == "H + H + M -> H2 + M"
    assert net.has_kinf[i] and net.has_kinf[i + 1] and net.is_three_body[i]
    assert (net.a[i], net.n[i], net.E[i]) == (2.70e-31, -0.6, 0.0)
    assert (net.a_inf[i], net.n_inf[i], net.E_inf[i]) == (3.31e-06, -1.0, 0.0)
    assert net.reactant_idx[i, 0] == sp["H"] and net.reactant_stoich[i, 0] == 2.0
    assert net.reactant_stoich[i, 1] == 0.0 and net.reactant_idx[i, 1] == net.ni
    i = 771     # "NH + M -> H + N + M  2.99E-10 0. 3.76E04" (no k_inf)
    assert (net.a[i], net.n[i], net.E[i]) == (2.99e-10, 0.0, 37600.0)
    assert not net.has_kinf[i] and net.is_three_body[i] and net.is_three_body[i + 1]
    i = 777     # "HNCO + M -> H + NCO": M on the left only
    assert net.is_three_body[i] and not net.is_three_body[i + 1]
    i = 781     # the special Troe reaction, no numbers
    assert net.is_special[i] and net.Rf[i] == "OH + CH3 + M -> CH3OH + M" and net.a[i] == 0.0
    assert net.pho_rate_index[("H2O", 1)] == 783 and net.pho_rate_index[("H2O", 3)] == 787
    assert net.n_branch["H2O"] >= 3 and net.n_branch["C6H6"] == 2
    assert net.is_photo[783] and not net.is_photo[784] and net.a[783] == 0.0


def test_species_order_and_extras():
    plain = N.parse_network(NET, extra_species=())
    assert plain.species[:8] == ("OH", "H2", "H2O", "H", "O", "CH", "C", "CH2")
    extra = N.parse_network(NET, extra_species=["He", "Ar", "Ar"])
    assert extra.species[:plain.ni] == plain.species
    assert extra.species[plain.ni:] == ("Ar",)          # He is in the network already
    assert extra.nr == plain.nr


def test_every_reaction_balances_elements(net):
    """Sum of atoms on the left equals the right for all 439 reactions and
    their reverse slots."""
    atoms, elements, mass = N.load_compositions(GOLD / "all_compose.txt", net.species)
    pad = np.vstack([atoms, np.zeros((1, atoms.shape[1]))])
    lhs = (net.reactant_stoich[:, :, None] * pad[net.reactant_idx]).sum(axis=1)
    rhs = (net.product_stoich[:, :, None] * pad[net.product_idx]).sum(axis=1)
    bad = [i for i in range(1, net.nr + 1) if not np.array_equal(lhs[i], rhs[i])]
    assert not bad, [(i, net.Rf.get(i - (1 - i % 2))) for i in bad[:5]]
    assert np.all(mass > 0)


def test_to_arrays_and_rate_classification(net):
    """The parsed network drives the chemistry tables and the rate build's
    host classification as it stands."""
    from vulcan_amd import rates
    arr = N.to_arrays(net)
    assert arr.ni == 69 and arr.nr == 878 and arr.is_three_body is net.is_three_body
    t = rates.rate_tables(net)
    assert t["form"][781] == rates.KF_TROE_OH_CH3
    assert t["form"][657] == rates.KF_LINDEMANN and t["form"][1] == rates.KF_ARRHENIUS
    assert t["form"][783] == rates.KF_ZERO and t["has_rev"][783] == 0
    assert t["has_rev"][1] == 1 and t["has_rev"][781] == 1
    assert t["stop_rev"] == 783
    # the three low-T cap reactions are slots 665, 671, 673 of this file
    assert (t["lowt"][665], t["lowt"][671], t["lowt"][673]) == (rates.LOWT_CH3, rates.LOWT_C2H4,
                                                                 rates.LOWT_C2H5)
    assert int((t["lowt"] != 0).sum()) == 3


def test_malformed_equation_raises(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1 [ A + B = C ] 1.0 0.0 0.0\n")
    with pytest.raises(ValueError):
        N.parse_network(p, extra_species=())


def test_oracle_chem_rhs_conserves_elements_on_the_real_network(net):
    """The oracle's chem_rhs on the parsed network with the oracle's rates:
    every layer's H, O, C, N and He production sums to zero against the
    composition table, to roundoff of the summed rate magnitudes.  Pins the
    oracle's use of the real tables (a wrong stoichiometric factor, a
    dropped slot or a swapped reverse row breaks it)."""
    from oracle import rates_np as R, ros2_np as O
    from vulcan_amd.gibbs import load_nasa9
    atoms, elements, _ = N.load_compositions(GOLD / "all_compose.txt", net.species)
    keep = [e for e in range(atoms.shape[1]) if atoms[:, e].any()]
    c9, _ = load_nasa9(net.species, GOLD)
    nz = 12
    T = np.linspace(1800.0, 700.0, nz)
    M = np.logspace(21.0, 15.0, nz)
    with np.errstate(all="ignore"):
        k = R.build_rate_array(net, T, M, c9)
    k = np.where(np.isfinite(k), k, 0.0)
    rng = np.random.default_rng(11)
    mix = rng.dirichlet(np.ones(net.ni), nz)
    y = M[:, None] * mix
    arr = N.to_arrays(net)
    f = O.chem_rhs(y, M, k, arr)
    # roundoff scale: every rate enters each element sum with |stoich * atoms|
    ys = np.concatenate([y, np.ones((nz, 1))], axis=1)
    rate = k.T * np.prod(ys[:, arr.reactant_idx] ** arr.reactant_stoich[None], axis=2)
    rate = np.where(arr.is_three_body[None, :], rate * M[:, None], rate)
    scale = np.abs(rate).sum(axis=1) * atoms.max()
    bal = f @ atoms[:, keep]
    assert np.all(np.abs(bal) / scale[:, None] < 1e-13), np.max(np.abs(bal) / scale[:, None])
    assert np.all(np.abs(f).max(axis=1) > 1e-6 * scale / atoms.max())      # not trivially 0
