"""The ``.vul`` output of a finished column in the reference's schema.

``vul_dicts`` builds the three top-level dictionaries the reference's writer
pickles (legacy_io._synthesize_save_dicts, legacy_io.py:429-539):
'variable' (species, nr, k as {row: k_row}, y, ymix, t, dt, longdy,
longdydt, y_ini, atom dictionaries, Rf; photolysis fields when the run had
photolysis), 'atm' (the atmosphere leaves plus Ti, gas_indx, pref_indx, gs
and the condensation dictionaries) and 'parameter' (step counters, delta,
the plotting palette, start_time), so downstream readers of ``.vul`` files
find the same keys.  ``save_vul`` writes them with pickle protocol 4 as
Output.save_out does (legacy_io.py:704-724).

longdy and longdydt are the reference's (against the ring's lookback
state, outer_loop.py:636-670) when the runner uses conv_mode "longdy";
under conv_mode "relmax" longdy is the largest relative change of the last
accepted step (the quantity of _f_dy, outer_loop.py:2163-2184) and
longdydt = longdy / dt.
"""
from __future__ import annotations

import pickle
from typing import Optional, Sequence

import numpy as np

__all__ = ["vul_dicts", "save_vul", "end_case", "TABLEAU20"]

# the normalised Tableau 20 palette published in the parameter dictionary
TABLEAU20 = [(r / 255.0, g / 255.0, b / 255.0) for r, g, b in (
    (31, 119, 180), (255, 127, 14), (44, 160, 44), (214, 39, 40), (148, 103, 189),
    (140, 86, 75), (227, 119, 194), (127, 127, 127), (188, 189, 34), (23, 190, 207),
    (174, 199, 232), (255, 187, 120), (152, 223, 138), (255, 152, 150), (197, 176, 213),
    (196, 156, 148), (247, 182, 210), (199, 199, 199), (219, 219, 141), (158, 218, 229))]


def end_case(accept_count: int, t: float, status: int, cfg) -> int:
    """The reference's end_case code of a finished run: 3 when the step cap
    was exceeded, 2 when the runtime was, else 1 (outer_loop.py:1938-1942,
    2326-2339).  Under conv_mode "longdy" the cap is the reference's own
    (accept_count > count_max); under "relmax" the runner stops at
    count_max accepted steps, so a column that stopped there without
    converging counts as capped."""
    from .outer_loop import ST_CONVERGED, ST_MAXSTEPS
    if cfg.conv_mode == "longdy":
        capped = int(accept_count) > int(cfg.count_max)
    else:
        capped = (int(accept_count) >= int(cfg.count_max) and bool(int(status) & ST_MAXSTEPS)
                  and not int(status) & ST_CONVERGED)
    if capped:
        return 3
    if float(t) > float(cfg.runtime):
        return 2
    return 1


def vul_dicts(runner, c: int, species: Sequence[str], atom_order: Sequence[str] = ("X",),
              Rf: Optional[dict] = None, y_ini=None, start_time: float = 0.0,
              pref_indx: int = 0, gs: float = 0.0):
    """(variable, atm, parameter) of column c of a SteadyStateRunner."""
    b = runner.batch
    res = runner.results()
    nz, ni = int(b.nz[c]), int(b.ni[c])
    S = int(b.S[c])
    if len(species) != ni:
        raise ValueError(f"{len(species)} species names for {ni} species")
    k = runner.k[int(b.off_k[c]):int(b.off_k[c]) + S * nz].view(S, nz).cpu().numpy()
    y = np.asarray(res["y"][c], np.float64)
    dt = float(res["dt"][c])
    na = len(atom_order)
    ai = runner.atom_ini.view(-1, 8)[c, :na].cpu().numpy()
    al = runner.atom_loss.view(-1, 8)[c, :na].cpu().numpy()
    compo = runner.compo[int(b.off_s[c]):int(b.off_s[c]) + ni * na].view(ni, na).cpu().numpy()
    asum = (y @ compo).sum(axis=0)
    if getattr(runner, "conv", None) is not None:
        longdy, longdydt = (float(v) for v in runner.conv_state()[c, :2])
    else:
        longdy = float(res["rel"][c])
        longdydt = longdy / dt if dt > 0 else 0.0
    var = {"species": list(species), "nr": S - 1,
           "k": {i: k[i].copy() for i in range(1, S)},
           "y": y, "ymix": np.asarray(res["ymix"][c], np.float64),
           "t": float(res["t"][c]), "dt": dt, "longdy": longdy, "longdydt": longdydt,
           "y_ini": np.asarray(runner.columns[c].y if y_ini is None else y_ini, np.float64),
           "atom_ini": {a: float(ai[j]) for j, a in enumerate(atom_order)},
           "atom_sum": {a: float(asum[j]) for j, a in enumerate(atom_order)},
           "atom_loss": {a: float(al[j]) for j, a in enumerate(atom_order)},
           "atom_conden": {a: 0.0 for a in atom_order},
           "Rf": dict(Rf or {})}
    if runner.photo is not None:
        st = runner.photo.state_tensors()
        setup = runner.photo.setup
        o, n = int(setup.off_p_np[c]), (nz + 1) * int(setup.nbin)
        for f in ("tau", "sflux", "aflux"):
            var[f] = st[f][o:o + n].view(nz + 1, -1).cpu().numpy()
        var["aflux_change"] = float(st["aflux_change"][c].item())
    atm = {f: np.asarray(getattr(runner.columns[c].atm, f))
           for f in runner.columns[c].atm._fields}
    total, off = _layout(nz, ni)
    rec = runner.atm[int(b.off_a[c]):int(b.off_a[c]) + total].cpu().numpy()
    for name, i, n in (("dzi", 1, nz - 1), ("Hpi", 3, nz - 1), ("g", 6, nz),
                       ("top_flux", 10, ni)):
        atm[name] = rec[off[i]:off[i] + n].copy()        # as the run last refreshed them
    gas = np.asarray(runner.columns[c].atm.gas_indx_mask, bool)
    atm.update({"gas_indx": [int(i) for i in np.nonzero(gas)[0]], "pref_indx": int(pref_indx),
                "gs": float(gs), "sat_p": {}, "sat_mix": {}, "r_p": {}, "rho_p": {},
                "fix_sp_indx": {}, "conden_min_lev": {}})
    counts = res["counts"]
    para = {"count": int(res["accept"][c]), "nega_count": int(counts["nega"][c]),
            "loss_count": int(counts["loss"][c]), "delta_count": int(counts["delta"][c]),
            "delta": float(res["delta"][c]), "small_y": 0.0, "nega_y": 0.0,
            "end_case": end_case(res["accept"][c], res["t"][c], res["status"][c], runner.cfg),
            "device_status": int(res["status"][c]), "solver_str": "solver",
            "switch_final_photo_frq": False, "where_varies_most": np.zeros_like(y),
            "pic_count": 0, "fix_species_start": False, "tableau20": list(TABLEAU20),
            "start_time": float(start_time)}
    return var, atm, para


def _layout(nz, ni):
    from .core import atm_record_layout
    return atm_record_layout(nz, ni)


def save_vul(path, runner, c: int, species: Sequence[str], **kw):
    """Write column c's .vul file (pickle protocol 4, legacy_io.py:716-724)."""
    var, atm, para = vul_dicts(runner, c, species, **kw)
    with open(path, "wb") as fh:
        pickle.dump({"variable": var, "atm": atm, "parameter": para}, fh, protocol=4)
