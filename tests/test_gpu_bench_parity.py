"""GPU parity at the benchmarked configuration (nz=150, ni=93, S=2461, dt=1e-10).

The columns are the first four of the throughput benchmark's batch
(``synth.columns_same_network(1002, B, 150, 93, 1200, 60, dt=1e-10)``), so
the kernels under test are the benchmarked ones: the SL = 6 chemistry
kernels and the SL = 6 tile block factor and solves (and the
cyclic-reduction factor of the latency mode).

Bars:
* chem_rhs: per element within 1e-14 of the cell's gross term sum
  (production + loss), the scale of its cancellation; chem_jac: 1e-14 per
  element;
* Ros2 step (jax_step.py:249-315), with the default refinement of both
  stages' block solves (also with the stage-1 solve refined alone, and on
  the cyclic-reduction factor): 1e-10 elementwise on every cell that
  the three reference solves (per-block LAPACK LU, the C oracle, a banded
  LAPACK solve of the whole column) agree on to 1e-12, and no masked cell
  further from the LAPACK oracle than 10x the references' own worst
  spread.  delta_arr = |sol - yk2| is a difference and is only reported.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available
from parity_util import StepRefs, elem_err

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

NZ, NI, NREV, NPHOTO, DT = 150, 93, 1200, 60, 1e-10
FACTOR_TILE, FACTOR_CR = 0, 6
VK_TUNE_REFINE = 2


@pytest.fixture(scope="module")
def bench_cols():
    from vulcan_amd import synth
    return synth.columns_same_network(1002, 4, NZ, NI, NREV, NPHOTO, dt=DT, name="bench")


@pytest.fixture(scope="module")
def refs(bench_cols):
    return [StepRefs(c.y, c.k, c.dt, c.atm, c.net) for c in bench_cols]


def test_chemistry_at_bench_shape(bench_cols):
    from oracle import ros2_np as O
    from vulcan_amd import chem
    cols = bench_cols
    y = np.stack([c.y for c in cols])
    M = np.stack([c.atm.M for c in cols])
    k = np.stack([c.k for c in cols])
    net = cols[0].net
    f = chem.chem_rhs(y, M, k, net).cpu().numpy()
    J = chem.chem_jac_analytical(y, M, k, net).cpu().numpy()
    for b, c in enumerate(cols):
        fo = O.chem_rhs(c.y, c.atm.M, c.k, net)
        g = O.chem_term_scale(c.y, c.atm.M, c.k, net)
        ef = float(np.max(np.abs(f[b] - fo) / np.maximum(g, 1e-300)))
        Jo = O.chem_jac(c.y, c.atm.M, c.k, net)
        eJ = float(np.max(np.abs(J[b] - Jo) / np.maximum(np.abs(Jo), 1e-300)))
        print(f"column {b}: chem_rhs {ef:.2e} of the term sum, chem_jac {eJ:.2e} per element")
        assert ef < 1e-14 and eJ < 1e-14


def _step(cols, factor, refine):
    from vulcan_amd import _lib, jax_step
    lib = _lib.load()
    sl = (NI + 15) // 16
    _lib.check(lib.vk_tune_set(0, sl, factor), "vk_tune_set")
    _lib.check(lib.vk_tune_set(VK_TUNE_REFINE, 0, refine), "vk_tune_set")
    try:
        sol, delta = jax_step.jax_ros2_step(np.stack([c.y for c in cols]),
                                            np.stack([c.k for c in cols]), DT,
                                            [c.atm for c in cols], cols[0].net)
        return sol.cpu().numpy(), delta.cpu().numpy()
    finally:
        lib.vk_tune_set(0, sl, FACTOR_TILE)
        lib.vk_tune_set(VK_TUNE_REFINE, 0, 2)


@pytest.mark.parametrize("factor,refine", [(FACTOR_TILE, 2), (FACTOR_TILE, 1), (FACTOR_CR, 2)],
                         ids=["tile", "tile-refine-stage1", "cr"])
def test_ros2_step_at_bench_shape(bench_cols, refs, factor, refine):
    sol, delta = _step(bench_cols, factor, refine)
    assert np.all(np.isfinite(sol)) and np.all(np.isfinite(delta))
    for b, r in enumerate(refs):
        e_det, e_all, spread, line = r.judge(sol[b], f"column {b}")
        ed = elem_err(delta[b], r.delta, r.determined & (r.delta > 0.0))
        print(line + f"; delta_arr on determined cells {ed.max():.2e}")
        assert r.determined.sum() >= r.mask.sum() // 2
        assert e_det < 1e-10
        assert e_all <= 10.0 * spread + 1e-10


def test_ros2_step_unrefined_block_solves(bench_cols, refs):
    """Without the refinement (VK_TUNE_REFINE 0) the explicit-inverse block
    solves leave up to ~1e-9 on determined cells; reported, and held to the
    per-layer bar of the references' spread."""
    from parity_util import layer_err
    sol, _ = _step(bench_cols, FACTOR_TILE, 0)
    for b, r in enumerate(refs):
        print(r.judge(sol[b], f"column {b}, unrefined")[3])
        assert layer_err(sol[b], r.np).max() <= 10.0 * r.layer_spread.max() + 1e-10


def test_rate_table_layouts_bitwise_equal(bench_cols):
    """VK_TUNE_KDIRECT: the chemistry reading k in its (S, nz) layout gives
    the same step, bit for bit, as reading the layer-major copy."""
    from vulcan_amd import _lib
    lib = _lib.load()
    VK_TUNE_KDIRECT = 4
    base = _step(bench_cols, FACTOR_TILE, 2)
    _lib.check(lib.vk_tune_set(VK_TUNE_KDIRECT, 0, 1), "vk_tune_set")
    try:
        direct = _step(bench_cols, FACTOR_TILE, 2)
    finally:
        lib.vk_tune_set(VK_TUNE_KDIRECT, 0, 0)
    assert np.array_equal(base[0], direct[0]) and np.array_equal(base[1], direct[1])


def test_held_inverse_sweeps_bitwise_equal(bench_cols):
    """VK_TUNE_HOLDX: the refined sweeps keeping X_j in registers for both
    of its products (two block reads per layer) give the same step, bit for
    bit, as the sweeps reading X_j again for the correction."""
    from vulcan_amd import _lib
    lib = _lib.load()
    VK_TUNE_HOLDX = 5
    assert lib.vk_tune_get(VK_TUNE_HOLDX, 0) == 1
    held = _step(bench_cols, FACTOR_TILE, 2)
    _lib.check(lib.vk_tune_set(VK_TUNE_HOLDX, 0, 0), "vk_tune_set")
    try:
        reread = _step(bench_cols, FACTOR_TILE, 2)
    finally:
        lib.vk_tune_set(VK_TUNE_HOLDX, 0, 1)
    assert np.array_equal(held[0], reread[0]) and np.array_equal(held[1], reread[1])
