"""GPU: checkpoint / resume of a steady-state batch, and the .vul output.

* A batch (three uneven columns, atmosphere refresh on) runs 30 attempts,
  is snapshotted to an .npz file, runs 30 more; a fresh runner on the same
  columns loads the file and runs the same 30: every device array equals
  the first run's, bit for bit (the run state the reference packs and
  unpacks, outer_loop.py:1795-2013).
* The .vul dictionaries of a column carry the keys of the reference's
  writer (legacy_io.py:429-539): variable, atm and parameter; k as
  {row: array(nz)}; y, ymix and the counters as the runner holds them;
  the file round-trips through pickle.
"""
from __future__ import annotations

import pickle

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VAR_KEYS = {"species", "nr", "k", "y", "ymix", "t", "dt", "longdy", "longdydt", "y_ini",
            "atom_ini", "atom_sum", "atom_loss", "atom_conden", "Rf"}
ATM_EXTRA = {"Ti", "gas_indx", "pref_indx", "gs", "sat_p", "sat_mix", "r_p", "rho_p",
             "fix_sp_indx", "conden_min_lev"}
PARA_KEYS = {"count", "nega_count", "loss_count", "delta_count", "delta", "small_y", "nega_y",
             "end_case", "solver_str", "switch_final_photo_frq", "where_varies_most",
             "pic_count", "fix_species_start", "tableau20", "start_time"}


def test_checkpoint_resume_is_bitwise(tmp_path):
    import torch
    from test_gpu_refresh import _columns
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols, statics = _columns(seed=6)
    cfg = RunConfig(update_frq=4, count_max=400)
    a = SteadyStateRunner(cols, cfg, refresh=statics)
    for _ in range(30):
        a.attempt()
    path = tmp_path / "run.npz"
    a.save_checkpoint(path)
    for _ in range(30):
        a.attempt()
    end = a.snapshot()
    b = SteadyStateRunner(cols, cfg, refresh=statics)
    b.load_checkpoint(path)
    for _ in range(30):
        b.attempt()
    torch.cuda.synchronize()
    got = b.snapshot()
    assert set(got) == set(end)
    for f in end:
        assert np.array_equal(got[f], end[f]), f
    assert int(end["accept"].min()) > 0


def test_vul_schema_and_round_trip(tmp_path):
    from test_gpu_refresh import _columns
    from vulcan_amd import output
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols, _ = _columns(seed=7)
    r = SteadyStateRunner(cols, RunConfig(count_max=400))
    for _ in range(20):
        r.attempt()
    c = 1
    col = cols[c]
    names = [f"S{i}" for i in range(col.ni)]
    path = tmp_path / "col.vul"
    output.save_vul(path, r, c, names, Rf={1: "S0 + S1 -> S2 + S3"})
    with open(path, "rb") as fh:          # a file this test wrote itself
        d = pickle.load(fh)
    assert set(d) == {"variable", "atm", "parameter"}
    var, atm, para = d["variable"], d["atm"], d["parameter"]
    assert VAR_KEYS <= set(var) and PARA_KEYS <= set(para)
    assert ATM_EXTRA | set(col.atm._fields) <= set(atm)
    res = r.results()
    assert np.array_equal(var["y"], res["y"][c]) and np.array_equal(var["ymix"], res["ymix"][c])
    assert var["nr"] == col.net.nr and sorted(var["k"]) == list(range(1, col.net.nr + 1))
    assert np.array_equal(np.stack([var["k"][i] for i in range(1, col.net.nr + 1)]), col.k[1:])
    assert para["count"] == int(res["accept"][c]) and len(para["tableau20"]) == 20
    assert var["species"] == names and var["longdydt"] == var["longdy"] / var["dt"]
    # still running after 20 attempts: the reference's end_case 1 (outer_loop.py:2326-2339)
    assert para["end_case"] == 1 and para["device_status"] == int(res["status"][c])


def test_vul_end_case_of_capped_and_converged_columns():
    """end_case is the reference's code (1 converged, 3 step cap exceeded),
    not the device status word: one column capped at 10 accepted steps, and a
    small column run to convergence."""
    from test_gpu_refresh import _columns
    from test_gpu_steady import _small
    from vulcan_amd import output
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner, ST_CONVERGED, ST_MAXSTEPS
    cols, _ = _columns(seed=7)
    r = SteadyStateRunner(cols[:1], RunConfig(count_max=10))
    r.run(max_attempts=400, check_every=20)
    res = r.results()
    assert res["status"][0] & ST_MAXSTEPS and not res["status"][0] & ST_CONVERGED
    para = output.vul_dicts(r, 0, [f"S{i}" for i in range(cols[0].ni)])[2]
    assert para["end_case"] == 3 and para["count"] == 10
    c = _small(21)
    r = SteadyStateRunner([c], RunConfig())
    r.run(max_attempts=20000, check_every=100)
    res = r.results()
    assert res["status"][0] & ST_CONVERGED
    para = output.vul_dicts(r, 0, [f"S{i}" for i in range(c.ni)])[2]
    assert para["end_case"] == 1 and para["device_status"] & ST_CONVERGED
