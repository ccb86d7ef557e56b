"""GPU: constraints on accepted states inside the device controller
(outer_loop.py:855-895): ion charge balance, the fixed bottom layer,
pinned bottom mixing ratios and the H2/He pin, after the hydrostatic
balance.

Five columns run in one SteadyStateRunner, one constraint each (one
without).  After
every attempt the host snapshots (ymix before, sol after) and, for each
column the attempt accepted, recomputes the accepted state with the
oracle (clip, balance, constraint) and compares y to 1e-13 per layer; the
constraint itself must hold exactly on the device (bottom layer, pinned
species) or to rounding (e = -sum charge * y).  A column without
constraints gives the same steps bit for bit as a run without the
constraint table.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cols():
    from vulcan_amd import synth
    out = []
    # the serving test's first requests: columns whose adaptive runs converge
    for c, (nz, ni, n_rev) in enumerate(((18, 14, 40), (24, 20, 55), (16, 14, 40),
                                         (30, 20, 55), (22, 14, 40))):
        out.append(synth.make_column(500 + c, nz=nz, ni=ni, n_rev=n_rev, dt=1e-10, p_bot=1e5,
                                     p_top=1e-1, consistent=True))
    return out


def test_constraints_in_the_loop_match_oracle():
    import torch
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import ColumnConstraints, RunConfig, SteadyStateRunner
    cols = _cols()
    ni0, ni1, ni2 = cols[0].ni, cols[1].ni, cols[2].ni
    charge = np.zeros(ni0)
    charge[[2, 5]] = -1.0                       # two cations; e is the last species
    bottom = cols[1].y[0] * 1.01
    cons = [ColumnConstraints(charge=charge, e_idx=ni0 - 1),
            ColumnConstraints(bottom_n=bottom),
            ColumnConstraints(fix_sp_bot={0: 0.3, 4: 1e-3}),
            None,
            ColumnConstraints(h2he=(1, 3, 0.0))]
    cfg = RunConfig(count_max=400)
    r = SteadyStateRunner(cols, cfg, constraints=cons)
    plain = SteadyStateRunner([cols[3]], cfg)
    b = r.batch
    p = O.StepParams()
    checked = h2he_checked = 0
    pinned, mixes = False, (0.0, 0.0)
    for it in range(40):
        ymix_before = [t.cpu().numpy().copy() for t in b.unpack_state(r.ymix)]
        t_before = r.t.cpu().numpy().copy()
        pinned, mixes = O.h2he_pin_step(pinned, mixes, t_before[4], ymix_before[4], 1, 3, 0.0)
        yprev_before = [t.cpu().numpy().copy() for t in b.unpack_state(r.y_prev)]
        acc_before = r.accept.cpu().numpy().copy()
        retry_before = r.retry.cpu().numpy().copy()
        dt_before = r.dt.cpu().numpy().copy()
        r.attempt()
        plain.attempt()
        torch.cuda.synchronize()
        acc = r.accept.cpu().numpy()
        ys = [t.cpu().numpy() for t in b.unpack_state(r.y)]
        sols = [t.cpu().numpy() for t in b.unpack_state(r.sol)]
        for c in (0, 1, 2, 4):
            if acc[c] == acc_before[c]:
                continue
            col = cols[c]
            may_force = (retry_before[c] >= cfg.batch_max_retries or
                         dt_before[c] * cfg.dt_var_min < cfg.dt_min)
            if may_force and np.array_equal(ys[c], np.maximum(yprev_before[c], 0.0)):
                continue                        # a forced accept keeps y_prev (clipped at 0)
            yc, mixn, _, _ = O.clip(sols[c], ymix_before[c], p, np.ones(col.ni, bool))
            cc = cons[c]
            ref = O.balance_and_constrain(yc, mixn, col.atm.M, charge=cc.charge,
                                          e_idx=cc.e_idx, bottom_n=cc.bottom_n,
                                          fix_sp_bot=cc.fix_sp_bot,
                                          h2he=(1, 3) + mixes if c == 4 and pinned else None)
            err = np.max(np.abs(ys[c] - ref) / np.abs(ref).max(axis=1, keepdims=True))
            assert err < 1e-13, (it, c, err)
            if c == 0:
                q = -(ys[c] * charge[None, :]).sum(axis=1)
                assert np.max(np.abs(ys[c][:, ni0 - 1] - q) / q) < 1e-14
            if c == 1:
                assert np.array_equal(ys[c][0], bottom)
            if c == 2:
                assert ys[c][0, 0] == 0.3 * col.atm.M[0] and ys[c][0, 4] == 1e-3 * col.atm.M[0]
            if c == 4 and pinned:
                assert ys[c][0, 1] == mixes[0] * col.atm.M[0]
                assert ys[c][0, 3] == mixes[1] * col.atm.M[0]
                h2he_checked += 1
            checked += 1
        # the unconstrained column is unaffected by the others' constraints
        assert torch.equal(b.unpack_state(r.y)[3], plain.batch.unpack_state(plain.y)[0])
    print(f"accepted states checked: {checked}, with the H2/He pin active: {h2he_checked}")
    assert checked >= 30 and h2he_checked >= 5
