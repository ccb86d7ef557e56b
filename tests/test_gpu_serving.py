"""Continuous batching: each request's steady state is the same alone and in
the mix (bitwise), across slot refills, two size buckets on two streams, and
graph-replayed or eager attempts."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _requests():
    from vulcan_amd import synth
    sizes = [(18, 14, 40), (24, 20, 55), (16, 14, 40), (30, 20, 55), (20, 14, 40), (22, 20, 55)]
    return [synth.make_column(500 + i, nz=a, ni=b, n_rev=r, dt=1e-10, name=f"req{i}",
                              p_bot=1e5, p_top=1e-1, consistent=True)
            for i, (a, b, r) in enumerate(sizes)]


def _alone(col, latency=True):
    """The request solved alone, on the block factor the engine's pools use
    (the cyclic-reduction latency mode by default)."""
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    r = SteadyStateRunner([col], RunConfig(count_max=1500), latency_mode=latency)
    r.run(max_attempts=3000, check_every=100)
    res = r.results()
    return res["y"][0], int(res["attempts"][0]), int(res["status"][0])


@pytest.mark.parametrize("use_graph,latency", [(True, True), (False, True), (True, False)])
def test_served_results_equal_alone(use_graph, latency):
    from vulcan_amd.outer_loop import RunConfig
    from vulcan_amd.serving import ServingEngine, latency_summary
    reqs = _requests()
    ref = [_alone(c, latency) for c in reqs]
    # two slots per bucket: later requests wait and refill slots freed by earlier ones
    eng = ServingEngine(hbm_budget_bytes=1e9, max_slots=2, cfg=RunConfig(count_max=1500),
                        chunk=25, use_graph=use_graph, latency_mode=latency)
    arrivals = np.array([0.0, 0.01, 0.02, 0.2, 0.25, 0.3])
    out = eng.run(reqs, arrivals)
    assert len(out) == len(reqs)
    print(latency_summary(out), eng.slot_counts)
    for r in out:
        y0, att0, st0 = ref[r.index]
        print(f"req {r.index}: attempts {r.attempts} (alone {att0}), status {r.status} ({st0})")
        assert r.attempts == att0 and r.status == st0
        assert np.array_equal(r.y, y0)
        assert r.latency >= 0.0 and r.finish >= r.start >= r.arrival - 1e-9


def test_single_slot_budget_queues_in_order():
    """A budget that buys one slot per bucket still serves every request."""
    from vulcan_amd.outer_loop import RunConfig
    from vulcan_amd.serving import ServingEngine
    reqs = _requests()[:4]
    eng = ServingEngine(hbm_budget_bytes=1.0, max_slots=8, cfg=RunConfig(count_max=1500),
                        chunk=50)
    out = eng.run(reqs, realtime=False)
    assert eng.slot_counts == {1: 1, 2: 1}
    assert sorted(r.index for r in out) == [0, 1, 2, 3]
    ref = [_alone(c) for c in reqs]
    for r in out:
        assert np.array_equal(r.y, ref[r.index][0])


def _rate_source_requests():
    """Requests that carry their rate network and NASA-9 fits instead of k,
    and the same requests with k built beforehand by build_rate_array."""
    from types import SimpleNamespace
    from vulcan_amd import rates, synth
    from vulcan_amd.serving import RateSource
    with_src, with_k = [], []
    for i, col in enumerate(_requests()[:4]):
        rng = np.random.default_rng(900 + i)
        n_rev = (col.net.nr) // 2
        rnet = synth.random_rate_network(rng, col.net, n_rev, 0)
        c9 = np.zeros((col.ni, 2, 10))        # g/RT = 2.5 (1 - ln T) + a8 / T - a9
        c9[:, :, 2] = 2.5
        c9[:, :, 8] = rng.uniform(-2000.0, 2000.0, col.ni)[:, None]
        c9[:, :, 9] = rng.uniform(-2.0, 2.0, col.ni)[:, None]
        src = RateSource(rnet, c9, use_lowT_limit_rates=bool(i % 2), remove_list=[3] if i else None)
        cfg = SimpleNamespace(use_lowT_limit_rates=src.use_lowT_limit_rates,
                              remove_list=src.remove_list)
        k = rates.build_rate_array(cfg, rnet, col.atm, c9).cpu().numpy()
        a = synth.Column(**{**col.__dict__, "k": None})
        a.rate_source = src
        with_src.append(a)
        with_k.append(synth.Column(**{**col.__dict__, "k": k}))
    return with_src, with_k


def test_requests_carrying_rate_sources():
    """A request that carries its rate network has its k built on the device
    in its slot from its own T and M; it is served exactly like the same
    request with k built beforehand."""
    from vulcan_amd.outer_loop import RunConfig
    from vulcan_amd.serving import ServingEngine
    with_src, with_k = _rate_source_requests()
    ref = [_alone(c) for c in with_k]
    eng = ServingEngine(hbm_budget_bytes=1e9, max_slots=2, cfg=RunConfig(count_max=1500),
                        chunk=25, use_graph=True)
    out = eng.run(with_src, realtime=False)
    assert sorted(r.index for r in out) == list(range(len(with_src)))
    for r in out:
        y0, att0, st0 = ref[r.index]
        print(f"req {r.index}: attempts {r.attempts} (k given: {att0}), status {r.status} ({st0})")
        assert r.attempts == att0 and r.status == st0
        assert np.array_equal(r.y, y0)


def test_metrics_records_of_runs_and_serving(tmp_path):
    """The JSON Lines records agree with the results they describe."""
    import json
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    from vulcan_amd.serving import ServingEngine
    reqs = _requests()
    path = tmp_path / "run.jsonl"
    r = SteadyStateRunner(reqs[:2], RunConfig(count_max=300))
    r.run(max_attempts=600, check_every=100, metrics=path)
    recs = [json.loads(line) for line in path.read_text().splitlines()]
    res = r.results()
    assert recs and all(x["event"] == "steady_check" for x in recs)
    last = recs[-1]
    assert last["accepted_total"] == int(res["accept"].sum())
    assert last["attempts_total"] == int(res["attempts"].sum())
    assert last["done"] == 2 and last["columns"] == 2
    assert last["rejects"]["delta"] == int(res["counts"]["delta"].sum())
    path2 = tmp_path / "serve.jsonl"
    eng = ServingEngine(hbm_budget_bytes=1e9, max_slots=2, cfg=RunConfig(count_max=300), chunk=25)
    out = eng.run(reqs, metrics=path2)
    recs = [json.loads(line) for line in path2.read_text().splitlines()]
    done = [x for x in recs if x["event"] == "request_done"]
    summ = [x for x in recs if x["event"] == "serving_summary"]
    assert sorted(x["index"] for x in done) == sorted(o.index for o in out)
    assert len(summ) == 1 and summ[0]["n"] == len(out)
    by_idx = {o.index: o for o in out}
    for x in done:
        assert x["attempts"] == by_idx[x["index"]].attempts
