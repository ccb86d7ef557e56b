"""GPU steady-state runs vs the CPU oracle driver, graph replay and batch invariance.

Columns use detailed-balance equilibrium constants (synth consistent=True)
so that a steady state exists; the bar for converged mixing ratios is
1e-6 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _small(seed, nz=20, ni=12, n_rev=40):
    from vulcan_amd import synth
    return synth.make_column(seed, nz=nz, ni=ni, n_rev=n_rev, dt=1e-10, name=f"ss{seed}",
                             p_bot=1e4, p_top=1e-2, consistent=True)


def _mix_err(a, b, floor=1e-12):
    """max relative error of mixing ratios above floor."""
    m = b > floor
    return float(np.max(np.abs(a[m] - b[m]) / b[m]))


def test_steady_state_matches_oracle():
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = _small(21)
    tr = []
    yo, info = O.steady_state_run(c.y, c.k, c.atm, c.net, c.dt, trace=tr)
    r = SteadyStateRunner([c], RunConfig())
    r.run(max_attempts=20000, check_every=100)
    res = r.results()
    assert res["status"][0] & 4, res["status"]
    so = info["state"]
    print(f"oracle: {info['attempts']} attempts, {so.accept_count} accepted; "
          f"gpu: {res['attempts'][0]} attempts, {res['accept'][0]} accepted")
    mix_o = so.ymix
    err = _mix_err(res["ymix"][0], mix_o)
    print(f"steady-state mixing ratio error {err:.3e}")
    assert err < 1e-6


def test_attempt_trajectory_matches_oracle():
    """Same accept/reject decisions and dt sequence for the first attempts."""
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = _small(22)
    tr = []
    O.steady_state_run(c.y, c.k, c.atm, c.net, c.dt, trace=tr, max_attempts=150)
    r = SteadyStateRunner([c], RunConfig())
    dts = []
    for _ in range(150):
        dts.append(float(r.dt.item()))
        r.attempt()
    dto = np.array([x["dt"] for x in tr])
    assert np.max(np.abs(np.array(dts) - dto) / dto) < 1e-8


def test_graph_replay_equals_eager():
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = _small(23)
    eager = SteadyStateRunner([c], RunConfig())
    eager.run(max_attempts=400, check_every=40)
    graph = SteadyStateRunner([c], RunConfig())
    graph.run(max_attempts=400, check_every=40, use_graph=True)
    a, b = eager.results(), graph.results()
    assert np.array_equal(a["y"][0], b["y"][0])
    assert np.array_equal(a["attempts"], b["attempts"])


def test_uneven_batch_is_bitwise_alone():
    """Columns of different nz, ni and networks in one batch give exactly
    the results each gives alone (no cross-column coupling, fixed orders)."""
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols = [_small(31, nz=18, ni=12, n_rev=40), _small(32, nz=25, ni=20, n_rev=60),
            _small(33, nz=12, ni=37, n_rev=90)]
    cfg = RunConfig(count_max=300)
    mix = SteadyStateRunner(cols, cfg)
    mix.run(max_attempts=2000, check_every=100)
    rm = mix.results()
    for i, c in enumerate(cols):
        alone = SteadyStateRunner([c], cfg)
        alone.run(max_attempts=2000, check_every=100)
        ra = alone.results()
        assert np.array_equal(ra["y"][0], rm["y"][i]), i
        assert ra["attempts"][0] == rm["attempts"][i]


def test_uneven_batch_graph_replay_equals_eager():
    """Three size buckets run on forked streams inside each attempt; captured
    in one HIP graph (fork / join through events) and replayed, the batch
    gives exactly the eager results."""
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols = [_small(41, nz=18, ni=12, n_rev=40), _small(42, nz=25, ni=20, n_rev=60),
            _small(43, nz=12, ni=37, n_rev=90), _small(44, nz=20, ni=14, n_rev=40)]
    cfg = RunConfig(count_max=300)
    eager = SteadyStateRunner(cols, cfg)
    eager.run(max_attempts=600, check_every=50)
    graph = SteadyStateRunner(cols, cfg)
    graph.run(max_attempts=600, check_every=50, use_graph=True)
    a, b = eager.results(), graph.results()
    for i in range(len(cols)):
        assert np.array_equal(a["y"][i], b["y"][i]), i
    assert np.array_equal(a["attempts"], b["attempts"])


def test_nonfinite_step_is_rejected_like_the_oracle():
    """A step whose solution holds NaN is rejected, as the reference does:
    all(sol_clip >= 0) is false for NaN and the NaN atom loss fails its test
    (outer_loop.py:741-750).  A NaN rate row makes every attempt's solution
    NaN, so dt halves on each reject, the loss counter counts them, and after
    batch_max_retries, or once dt would fall below dt_min, the step is
    force-accepted with max(y_prev, 0) and dt_min (outer_loop.py:753-768,
    1031-1038): 6 forced steps in 20 attempts.  The same dt
    sequence, accept flags, counters and state as the oracle; the state
    stays finite and the non-finite status bit is set."""
    import dataclasses
    from oracle import ros2_np as O
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner, ST_NONFINITE
    c = _small(24)
    k = c.k.copy()
    k[5, :] = np.nan
    c = dataclasses.replace(c, k=k)
    tr = []
    _, info = O.steady_state_run(c.y, k, c.atm, c.net, c.dt, trace=tr, max_attempts=20)
    so = info["state"]
    r = SteadyStateRunner([c], RunConfig())
    dts = []
    for _ in range(20):
        dts.append(float(r.dt.item()))
        r.attempt()
    res = r.results()
    dto = np.array([x["dt"] for x in tr])
    print(f"dt gpu {dts[:10]} oracle {dto[:10]}; accepted {res['accept'][0]} vs "
          f"{so.accept_count}; counts {res['counts']}")
    assert np.array_equal(np.array(dts), dto)
    assert int(res["accept"][0]) == so.accept_count == 6
    assert int(res["counts"]["loss"][0]) == so.loss_count
    assert int(res["counts"]["nega"][0]) == so.nega_count
    assert int(res["counts"]["delta"][0]) == so.delta_count
    assert float(res["t"][0]) == so.t and float(res["dt"][0]) == so.dt
    assert np.all(np.isfinite(res["y"][0])) and np.array_equal(res["y"][0], so.y)
    assert res["status"][0] & ST_NONFINITE
