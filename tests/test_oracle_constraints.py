"""Pins of the oracle's accepted-state constraints
(ros2_np.balance_and_constrain, outer_loop.py:847-869).

The ion clamp follows the reference's own known-answer test
(test_outer_loop_ion.py:55-100): nz = 4, ni = 5, densities drawn with
seed 0 in [1e8, 1e12], charges (+1, +2, -1 as stored, 0 at e), e = -sum
charge * y matches the direct sum, other columns untouched, and the net
electron count with +1 for e itself is zero.  The bottom pins set the
bottom layer exactly; the balance leaves non-gas species clipped.
"""
from __future__ import annotations

import numpy as np

from oracle import ros2_np as O


def test_ion_clamp_known_answer():
    nz, ni, e_idx = 4, 5, 4
    rng = np.random.default_rng(0)
    y = rng.uniform(1e8, 1e12, size=(nz, ni))
    y[:, e_idx] = 7.0
    charge = np.array([0.0, +1.0, +2.0, -1.0, 0.0])
    n0 = y.sum(axis=1)
    out = O.balance_and_constrain(y, y / n0[:, None], n0, charge=charge, e_idx=e_idx)
    expected = -(charge[None, :] * out).sum(axis=1)
    assert np.max(np.abs(out[:, e_idx] - expected) / np.abs(expected)) <= 1e-14
    bal = n0[:, None] * (y / n0[:, None])
    for j in range(ni - 1):
        assert np.array_equal(out[:, j], bal[:, j])
    full = charge.copy()
    full[e_idx] = 1.0
    assert np.max(np.abs((full[None, :] * out).sum(axis=1))) < 1e-3


def test_bottom_pins_and_partial_balance():
    rng = np.random.default_rng(3)
    nz, ni = 6, 4
    sol = rng.uniform(1.0, 10.0, (nz, ni))
    gas = np.array([True, True, False, True])
    ys = np.where(gas[None], sol, 0.0).sum(axis=1, keepdims=True)
    n0 = rng.uniform(50.0, 60.0, nz)
    bottom = np.array([1.0, 2.0, 3.0, 4.0])
    out = O.balance_and_constrain(sol, sol / ys, n0, gas_mask=gas, bottom_n=bottom)
    assert np.array_equal(out[0], bottom)
    assert np.array_equal(out[1:, 2], sol[1:, 2])                 # non-gas: clipped value kept
    assert np.allclose(out[1:, gas].sum(axis=1), n0[1:], rtol=1e-14)
    out2 = O.balance_and_constrain(sol, sol / ys, n0, gas_mask=gas, fix_sp_bot={1: 0.25})
    assert out2[0, 1] == 0.25 * n0[0]
    assert np.array_equal(out2[0, [0, 3]], (n0[:, None] * (sol / ys))[0, [0, 3]])
