"""CPU pins of the oracle's steady-state stopping rule (oracle/conv_np.py,
outer_loop.py:636-713, 918-1003).

Closed forms and hand-built cases, each chosen so that a plausible slip
(a dropped mask, the wrong normalisation, the last slot not excluded, a
strict/loose inequality swapped, the 5 % rule dropped) fails one of them.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import conv_np as V


def test_lookback_picks_nearest_to_half_t_and_skips_the_last_slot():
    ring = np.array([1.0, 2.0, 4.0, 8.0, 16.0])
    # t = 16 -> target 8, slot 3; the just-written slot is 4 (accept count 5)
    assert V.lookback_index(ring, 16.0, 0.5, 5) == 3
    # target 16 is nearest the slot just written, which is excluded -> 8 (slot 3)
    assert V.lookback_index(ring, 32.0, 0.5, 5) == 3
    # ties go to the lowest index: target 3 is 1 from both 2 and 4
    assert V.lookback_index(ring, 6.0, 0.5, 1) == 1
    # a ring that wrapped: accept count 7 wrote slot 1 last, so target 2
    # falls to slot 0 (distance 1) rather than slot 1 (distance 0)
    assert V.lookback_index(ring, 4.0, 0.5, 7) == 0
    # a ring still filling: slot 2 is empty (t = 0) and nearer than slot 0
    assert V.lookback_index(np.array([5.0, 2.0, 0.0]), 4.0, 0.5, 2) == 2


def test_longdy_closed_forms():
    rng = np.random.default_rng(0)
    nz, ni = 6, 5
    n0 = rng.uniform(1e10, 1e12, nz)
    mix = rng.dirichlet(np.ones(ni), nz)
    y_old = n0[:, None] * mix
    # the same state: 0
    assert V.longdy(y_old, y_old, mix, n0, 0.0, 1e-20) == 0.0
    # twice the old state, ymix of the new one: |y - y/2| / n0 / (y / n0) = 1/2
    y = 2.0 * y_old
    assert V.longdy(y, y_old, y / n0[:, None], n0, 0.0, 1e-20) == pytest.approx(0.5, rel=1e-15)
    # against an empty ring entry: |y| / n0 / ymix = 1 where ymix = y / n0
    assert V.longdy(y_old, 0 * y_old, mix, n0, 0.0, 1e-20) == pytest.approx(1.0, rel=1e-15)


def test_longdy_masks():
    nz, ni = 3, 4
    n0 = np.full(nz, 100.0)
    mix = np.full((nz, ni), 0.25)
    y_old = n0[:, None] * mix
    y = y_old.copy()
    y[1, 2] *= 3.0            # only entry that moved: longdy = 2 * 25 / 100 / 0.25 = 2
    assert V.longdy(y, y_old, mix, n0, 0.0, 1e-20) == pytest.approx(2.0, rel=1e-15)
    ign = np.zeros(ni, bool)
    ign[2] = True
    assert V.longdy(y, y_old, mix, n0, 0.0, 1e-20, ignore=ign) == 0.0
    assert V.longdy(y, y_old, mix, n0, 100.0, 1e-20) == 0.0          # y below atol
    m2 = mix.copy()
    m2[1, 2] = 1e-21
    assert V.longdy(y, y_old, m2, n0, 0.0, 1e-20) == 0.0               # ymix below mtol_conv
    m3 = mix.copy()
    m3[1, 2] = 0.0
    assert V.longdy(y, y_old, m3, n0, 0.0, -1.0) == 0.0                # ymix == 0 -> ratio 0


def test_slope_min_clamp():
    Hp = np.array([1e5, 1e5, 1e5, 7.0])     # the last layer never counts
    assert V.slope_min(np.full(3, 1e5), Hp) == 1e-8                    # 1e5 / 1e8 = 1e-3 -> cap
    assert V.slope_min(np.full(3, 1e-3), Hp) == 1e-10                  # 1e-11 -> floor
    assert V.slope_min(np.array([5e-1, 1.0, 2.0]), Hp) == pytest.approx(5e-9, rel=1e-15)


def _tracker(**kw):
    return V.ConvTracker(2, 2, 4, trun_min=0.0, count_min=0, **kw)


def test_strict_and_loose_branches():
    Kzz, Hp = np.full(1, 1e5), np.full(2, 1e5)      # slope_min = 1e-8
    tr = _tracker()
    tr.longdy, tr.longdydt = 0.009, 0.9e-4
    assert tr.terminate(1.0, 1, Kzz, Hp) == (True, False)
    tr.longdy, tr.longdydt = 0.011, 0.9e-4            # strict fails, loose needs 1e-8
    assert tr.terminate(1.0, 1, Kzz, Hp) == (False, False)
    tr.longdydt = 0.9e-8
    assert tr.terminate(1.0, 1, Kzz, Hp) == (True, False)
    assert tr.terminate(1.0, 1, Kzz, Hp, aflux_change=0.2) == (False, False)
    tr.longdy = 0.1                                   # not < yconv_min
    assert tr.terminate(1.0, 1, Kzz, Hp) == (False, False)
    # ready gate, count and runtime caps
    tr = V.ConvTracker(2, 2, 4, trun_min=10.0, count_min=5, count_max=7, runtime=100.0)
    tr.longdy, tr.longdydt = 0.0, 0.0
    assert tr.terminate(10.0, 6, Kzz, Hp) == (False, False)     # t not > trun_min
    assert tr.terminate(11.0, 5, Kzz, Hp) == (False, False)     # count not > count_min
    assert tr.terminate(11.0, 6, Kzz, Hp) == (True, False)
    assert tr.terminate(11.0, 8, Kzz, Hp)[1] and not tr.terminate(11.0, 7, Kzz, Hp)[1]
    assert tr.terminate(101.0, 6, Kzz, Hp)[1]


def test_stall_counter_needs_a_five_percent_drop():
    tr = _tracker(conv_stall_window=2, atol=0.0)
    n0 = np.ones(2)
    mix = np.full((2, 2), 0.5)
    seq = [0.09, 0.088, 0.087, 0.0869]     # only the first is a new minimum
    for k, ld in enumerate(seq, start=1):
        # a state whose longdy against the empty slot equals ld: y = ld * ymix * n0
        y = np.full((2, 2), 0.5 * ld)
        tr.ring_y[:] = 0.0
        tr.ring_t[:] = 0.0
        tr.update(y, mix, float(k), k, True, n0)
        assert tr.longdy == pytest.approx(ld, rel=1e-15)
    assert tr.seen_min == pytest.approx(0.09) and tr.since == 3
    Kzz, Hp = np.full(1, 1e5), np.full(2, 1e5)
    tr.longdydt = 1.0                                 # neither branch converges
    assert tr.terminate(1.0, 5, Kzz, Hp)[0]          # stalled: since 3 > window 2
    tr.ring_y[:] = 0.0
    tr.ring_t[:] = 0.0
    tr.update(np.full((2, 2), 0.5 * 0.08), mix, 9.0, 5, True, n0)   # 0.08 < 0.95 * 0.09
    assert tr.longdy == pytest.approx(0.08, rel=1e-15)
    assert tr.since == 0 and not tr.terminate(9.0, 5, Kzz, Hp)[0]


def test_ring_history_is_in_time_order_after_wrapping():
    tr = V.ConvTracker(1, 1, 3)
    for k in range(1, 6):
        tr.update(np.full((1, 1), float(k)), np.ones((1, 1)), float(k), k, True, np.ones(1))
    ys, ts = tr.history(5)
    assert list(ts) == [3.0, 4.0, 5.0] and list(ys[:, 0, 0]) == [3.0, 4.0, 5.0]
    ys, ts = V.ConvTracker(1, 1, 8).history(0)
    assert len(ts) == 0
