"""GPU: the cyclic-reduction block factor (VK_FACTOR_CR) in the Ros2 step.

The same block-tridiagonal system as the tile factor (solver.py:33-126),
eliminated in log2(nz) levels instead of one sweep; every diagonal block is
still inverted with getrf's partial pivoting, but rows are combined in
another order, so the bar is agreement to the system's conditioning:

* workload 1 (nz = 100, ni = 50) at a well-conditioned dt: 1e-12 per layer
  against the tile factor and 1e-10 against the oracle;
* workload 1 at its own dt: 1e-10 on every layer where the two oracles
  agree to 1e-12, within 10x their spread elsewhere (the tile factor's
  own bar in test_gpu_parity.py);
* an uneven batch (nz 2..61, four SL buckets, odd and even nz, powers of
  two and one above): each column equals the same column stepped alone bit
  for bit, and the tile factor to 1e-12 at a small dt; the wide buckets
  (ni 69, 93, 112: SL 5-7) likewise, and the standalone solver there
  against a dense solve;
* steady-state attempts replayed from a HIP graph equal the eager ones bit
  for bit (the factor and the two solves are about 50 launches per attempt);
* a workspace sized for the tile factor is refused, not overrun;
* the per-batch latency mode gives one runner the factor without touching
  the process-wide tune, and runs with the in-loop atmosphere refresh on
  the tile factor's trajectory at well-conditioned dt.
"""
from __future__ import annotations

from contextlib import contextmanager

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VK_FACTOR_TILE, VK_FACTOR_CR = 0, 6


@contextmanager
def factor(variant):
    from vulcan_amd import _lib
    lib = _lib.load()
    try:
        for sl in range(1, 8):
            _lib.check(lib.vk_tune_set(0, sl, variant), "vk_tune_set")
        yield
    finally:
        for sl in range(1, 8):
            lib.vk_tune_set(0, sl, VK_FACTOR_TILE)


def _np(t):
    return t.detach().cpu().numpy()


def _layer_err(sol, ref):
    return np.max(np.abs(sol - ref), axis=1) / np.max(np.abs(ref), axis=1)


def test_cr_step_well_conditioned(cfg1):
    from oracle import ros2_np as O
    from vulcan_amd import jax_step
    c = cfg1
    dt = 1e-18
    with factor(VK_FACTOR_CR):
        cr, _ = jax_step.jax_ros2_step(c.y, c.k, dt, c.atm, c.net)
    tile, _ = jax_step.jax_ros2_step(c.y, c.k, dt, c.atm, c.net)
    so, _ = O.ros2_step(c.y, c.k, dt, c.atm, c.net)
    e_tile = _layer_err(_np(cr), _np(tile)).max()
    e_or = _layer_err(_np(cr), so).max()
    print(f"cr vs tile {e_tile:.2e}, vs oracle {e_or:.2e}")
    assert e_tile < 1e-12 and e_or < 1e-10


def test_cr_step_cfg1_matches_oracle(cfg1):
    from oracle import ros2_c, ros2_np as O
    from vulcan_amd import core, jax_step
    c = cfg1
    with factor(VK_FACTOR_CR):
        sol, delta = jax_step.jax_ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    so, do = O.ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    rec = core.pack_atm_record(c.atm).numpy()[None]
    sc, _ = ros2_c.ros2_steps(c.y[None], c.k[None], rec, c.dt, c.net)
    e_gpu = _layer_err(_np(sol), so)
    e_floor = _layer_err(sc[0], so)
    well = e_floor < 1e-12
    print(f"cr cfg1: well-determined layers {well.sum()}/{c.nz}, gpu err there "
          f"{e_gpu[well].max():.3e}; worst {e_gpu.max():.3e} vs oracle spread "
          f"{e_floor.max():.3e}")
    assert np.all(np.isfinite(_np(sol)))
    assert e_gpu[well].max() < 1e-10
    assert e_gpu.max() <= 10.0 * e_floor.max() + 1e-10
    from parity_util import StepRefs
    r = StepRefs(c.y, c.k, c.dt, c.atm, c.net)
    e_det, e_all, spread, line = r.judge(_np(sol), "cr cfg1 elementwise")
    print(line)
    assert e_det < 1e-10
    assert e_all <= 10.0 * spread + 1e-10


def _uneven():
    from vulcan_amd import synth
    shapes = ((2, 8), (3, 14), (5, 20), (9, 33), (16, 20), (17, 50), (33, 14), (61, 40))
    return [synth.make_column(900 + i, nz=nz, ni=ni, n_rev=3 * ni, dt=1e-16, p_bot=1e6,
                              p_top=1e0) for i, (nz, ni) in enumerate(shapes)]


def _batch_step(cols):
    """One Ros2 step of the columns as one uneven batch (vk_ros2_step_f64)."""
    import torch
    from vulcan_amd import _lib
    from vulcan_amd.core import F64, Batch, device, device_network
    dev = device()
    nets = [device_network(c.net) for c in cols]
    b = Batch([c.nz for c in cols], nets, list(range(len(cols))))
    y = b.pack_state([c.y for c in cols], dev)
    k = b.pack_k([c.k for c in cols], dev)
    atm = b.pack_atm([c.atm for c in cols], dev)
    dt = torch.tensor([float(c.dt) for c in cols], dtype=F64, device=dev)
    sol, delta = torch.empty_like(y), torch.empty_like(y)
    status = torch.zeros(len(cols), dtype=torch.int32, device=dev)
    ws = b.workspace()
    _lib.check(_lib.load().vk_ros2_step_f64(b.handle, y.data_ptr(), k.data_ptr(),
                                            atm.data_ptr(), dt.data_ptr(), None, sol.data_ptr(),
                                            delta.data_ptr(), ws.data_ptr(), ws.numel(),
                                            status.data_ptr(), _lib.stream_ptr()),
               "vk_ros2_step_f64")
    torch.cuda.synchronize()
    return [_np(t) for t in b.unpack_state(sol)], _np(status)


def test_cr_uneven_batch_is_bitwise_alone_and_matches_tile():
    cols = _uneven()
    with factor(VK_FACTOR_CR):
        batch, st = _batch_step(cols)
        alone = [_batch_step([c])[0][0] for c in cols]
    tile, _ = _batch_step(cols)
    assert not st.any()
    for c, col in enumerate(cols):
        assert np.array_equal(batch[c], alone[c]), (c, col.nz, col.ni)
        assert np.all(np.isfinite(batch[c]))
        e = _layer_err(batch[c], tile[c]).max()
        assert e < 1e-12, (c, col.nz, col.ni, e)


def test_cr_attempt_graph_replay_equals_eager():
    import torch
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    from vulcan_amd import synth
    cols = [synth.make_column(500 + c, nz=nz, ni=ni, n_rev=n_rev, dt=1e-10, p_bot=1e5,
                              p_top=1e-1, consistent=True)
            for c, (nz, ni, n_rev) in enumerate(((18, 14, 40), (24, 20, 55), (16, 14, 40)))]
    with factor(VK_FACTOR_CR):
        r = SteadyStateRunner(cols, RunConfig(count_max=400))
        for _ in range(30):
            r.attempt()
        torch.cuda.synchronize()
        eager = [x.clone() for x in (r.y, r.dt, r.accept, r.status)]
        r.reset_to_initial()
        r.capture(15)
        for _ in range(2):
            r.graph[0].replay()
        torch.cuda.synchronize()
        for a, e in zip((r.y, r.dt, r.accept, r.status), eager):
            assert torch.equal(a, e)
        assert int(r.accept.min().item()) >= 10


def test_cr_refuses_a_workspace_sized_for_the_tile_factor():
    """A runner built under the tile factor holds the smaller workspace; once
    its bucket is tuned to CR the attempt is refused with an error instead of
    writing past it, and it runs again after the tune is restored."""
    import torch
    from vulcan_amd import _lib, synth
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    col = synth.make_column(501, nz=24, ni=20, n_rev=55, dt=1e-10, p_bot=1e5, p_top=1e-1,
                            consistent=True)
    r = SteadyStateRunner([col], RunConfig())
    with factor(VK_FACTOR_CR):
        with pytest.raises(_lib.VulcanError, match="workspace"):
            r.attempt()
    r.attempt()
    torch.cuda.synchronize()
    assert int(r.attempts.item()) == 1


def test_cr_workspace_is_seven_block_stacks_larger():
    from vulcan_amd import _lib
    from vulcan_amd.core import Batch, device_network
    cols = _uneven()[:4]
    b = Batch([c.nz for c in cols], [device_network(c.net) for c in cols], [0, 1, 2, 3])
    lib = _lib.load()
    plain = int(lib.vk_ros2_workspace_bytes(b.handle))
    with factor(VK_FACTOR_CR):
        cr = int(lib.vk_ros2_workspace_bytes(b.handle))
    blk = int(lib.vk_batch_block_doubles(b.handle))
    assert cr - plain >= 7 * blk * 8 and cr - plain < 7 * blk * 8 + 4096


def test_latency_mode_is_per_batch(cfg1):
    """vk_batch_set_latency_mode gives one runner the cyclic-reduction factor
    (bit for bit what the process-wide tune gives) and leaves the tune and
    other runners on the tile factor.  At dt = 1e-10 both factors round
    alike; workload 1 at its own dt (stiff blocks) tells them apart."""
    import torch
    from vulcan_amd import _lib
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    from vulcan_amd import synth
    cols = [synth.make_column(500 + c, nz=nz, ni=ni, n_rev=n_rev, dt=1e-10, p_bot=1e5,
                              p_top=1e-1, consistent=True)
            for c, (nz, ni, n_rev) in enumerate(((18, 14, 40), (24, 20, 55)))]
    lib = _lib.load()
    fast = SteadyStateRunner(cols, RunConfig(), latency_mode=True)
    plain = SteadyStateRunner(cols, RunConfig())
    assert all(lib.vk_tune_get(0, sl) == VK_FACTOR_TILE for sl in range(1, 8))
    assert fast.ws.numel() > plain.ws.numel()
    with factor(VK_FACTOR_CR):
        tuned = SteadyStateRunner(cols, RunConfig())
        for _ in range(25):
            tuned.attempt()
    with factor(VK_FACTOR_TILE):
        tile = SteadyStateRunner(cols, RunConfig())
        for _ in range(25):
            tile.attempt()
    for _ in range(25):
        fast.attempt()
        plain.attempt()
    torch.cuda.synchronize()
    assert torch.equal(fast.y, tuned.y) and torch.equal(fast.accept, tuned.accept)
    assert torch.equal(plain.y, tile.y) and torch.equal(plain.accept, tile.accept)
    stiff = [SteadyStateRunner([cfg1], RunConfig(), latency_mode=m) for m in (True, False)]
    for r in stiff:
        r.attempt()
    torch.cuda.synchronize()
    assert not torch.equal(stiff[0].sol, stiff[1].sol)



def test_latency_mode_with_the_in_loop_refresh():
    """The cyclic-reduction factor inside the full attempt: molecular
    diffusion, top fluxes and the in-loop atmosphere refresh every third
    accepted step.  At these columns' dt the blocks are well conditioned, so
    the trajectory follows the tile factor's: the same accept / reject
    sequence and states within 1e-10 per layer after 30 attempts."""
    import torch
    from test_gpu_refresh import _columns
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols, statics = _columns()
    cfg = RunConfig(update_frq=3, count_max=400)
    runs = [SteadyStateRunner(cols, cfg, refresh=statics, latency_mode=m) for m in (True, False)]
    for _ in range(30):
        for r in runs:
            r.attempt()
    torch.cuda.synchronize()
    assert torch.equal(runs[0].accept, runs[1].accept)
    assert torch.equal(runs[0].attempts, runs[1].attempts)
    assert int(runs[0].accept.min().item()) >= 10
    b = runs[0].batch
    for c, (a, t) in enumerate(zip(b.unpack_state(runs[0].y), b.unpack_state(runs[1].y))):
        a, t = a.cpu().numpy(), t.cpu().numpy()
        err = np.max(np.abs(a - t), axis=1) / np.max(np.abs(t), axis=1)
        assert np.all(np.isfinite(a)) and err.max() < 1e-10, (c, err.max())


@pytest.mark.parametrize("transpose", [False, True])
def test_standalone_cr_solver_matches_tile_and_dense(cfg1, transpose):
    """vk_bt_cr_factor_f64 / vk_bt_cr_solve_f64 on a random diagonally
    weighted block-tridiagonal system of workload 1's shape (nz 100, ni 50),
    plain and transposed: against the tile solver to 1e-12 and a dense
    solve to 1e-10."""
    import torch
    from vulcan_amd import steady_state_grad as G
    from vulcan_amd.core import F64, device
    c = cfg1
    nz, ni = c.nz, c.ni
    rng = np.random.default_rng(5)
    diag = rng.normal(size=(nz, ni, ni)) + 4.0 * np.sqrt(ni) * np.eye(ni)[None]
    sup = rng.normal(size=(nz, ni))
    sub = rng.normal(size=(nz, ni))
    rhs = rng.normal(size=(nz, ni))
    dev = device()
    t = lambda a: torch.tensor(a, dtype=F64, device=dev)
    out = {}
    try:
        for mode in (False, True):
            G.set_latency_mode(mode)
            x, st = G._factor_solve(t(diag), t(sup), t(sub), t(rhs), c.net, nz, transpose)
            torch.cuda.synchronize()
            assert int(st.item()) == 0
            out[mode] = _np(x)
    finally:
        G.set_latency_mode(False)
    A = np.zeros((nz * ni, nz * ni))
    for j in range(nz):
        A[j * ni:(j + 1) * ni, j * ni:(j + 1) * ni] = diag[j]
        if j >= 1:
            A[j * ni:(j + 1) * ni, (j - 1) * ni:j * ni] = np.diag(sub[j - 1])
            A[(j - 1) * ni:j * ni, j * ni:(j + 1) * ni] = np.diag(sup[j - 1])
    xd = np.linalg.solve(A.T if transpose else A, rhs.reshape(-1)).reshape(nz, ni)
    scale = np.abs(xd).max()
    e_tile = np.abs(out[True] - out[False]).max() / scale
    e_dense = np.abs(out[True] - xd).max() / scale
    print(f"transpose={transpose}: cr vs tile {e_tile:.2e}, vs dense {e_dense:.2e}")
    assert e_tile < 1e-12 and e_dense < 1e-10


def test_steady_state_gradient_in_latency_mode():
    """The whole gradient chain of the sensitivity test (Newton polish on the
    device, transposed adjoint solve, parameter VJP) with the solves in
    latency mode equals the block-Thomas one to 1e-9 of the largest
    gradient."""
    from test_gpu_sensitivity import Problem, _gpu_gradients
    from vulcan_amd import steady_state_grad as G
    pb = Problem(kzz=0.0)
    base = _gpu_gradients(pb)
    try:
        G.set_latency_mode(True)
        cr = _gpu_gradients(pb)
    finally:
        G.set_latency_mode(False)
    for key in ("m", "T", "bf"):
        a, b = cr[4][key], base[4][key]
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-9 * scale, (key, np.abs(a - b).max() / scale)
    assert np.abs(cr[0] - base[0]).max() <= 1e-12 * np.abs(base[0]).max()


@pytest.mark.parametrize("ni", [69, 93, 112])
def test_cr_wide_buckets_match_tile(ni):
    """SL 5, 6 and 7 (the widest, whose elimination kernel spills): a batch
    step on the cyclic-reduction factor equals the tile factor to 1e-12 at a
    well-conditioned dt, and the standalone solver matches a dense solve."""
    import torch
    from vulcan_amd import steady_state_grad as G, synth
    from vulcan_amd.core import F64, device
    cols = [synth.make_column(1200 + ni + i, nz=nz, ni=ni, n_rev=3 * ni, dt=1e-16, p_bot=1e6,
                              p_top=1e0) for i, nz in enumerate((7, 20))]
    with factor(VK_FACTOR_CR):
        cr, st = _batch_step(cols)
    tile, _ = _batch_step(cols)
    assert not st.any()
    for c in range(len(cols)):
        assert np.all(np.isfinite(cr[c]))
        assert _layer_err(cr[c], tile[c]).max() < 1e-12, (ni, c)
    nz = 20
    rng = np.random.default_rng(ni)
    diag = rng.normal(size=(nz, ni, ni)) + 4.0 * np.sqrt(ni) * np.eye(ni)[None]
    sup, sub, rhs = (rng.normal(size=(nz, ni)) for _ in range(3))
    dev = device()
    t = lambda a: torch.tensor(a, dtype=F64, device=dev)
    try:
        G.set_latency_mode(True)
        x, s = G._factor_solve(t(diag), t(sup), t(sub), t(rhs), cols[1].net, nz, False)
        torch.cuda.synchronize()
    finally:
        G.set_latency_mode(False)
    A = np.zeros((nz * ni, nz * ni))
    for j in range(nz):
        A[j * ni:(j + 1) * ni, j * ni:(j + 1) * ni] = diag[j]
        if j >= 1:
            A[j * ni:(j + 1) * ni, (j - 1) * ni:j * ni] = np.diag(sub[j - 1])
            A[(j - 1) * ni:j * ni, j * ni:(j + 1) * ni] = np.diag(sup[j - 1])
    xd = np.linalg.solve(A, rhs.reshape(-1)).reshape(nz, ni)
    assert int(s.item()) == 0
    assert np.abs(_np(x) - xd).max() < 1e-10 * np.abs(xd).max()


def test_cr_wide_levels_are_bitwise_narrow():
    """Sixteen nz=150, ni=93 columns: the first cyclic-reduction levels
    launch 16 x 75 workgroups, enough for the four-per-CU elimination build
    and the panel-loading update build; deeper levels, and each column
    alone, take the narrow builds.  The builds perform the same operations,
    so every column's step equals its step alone, bit for bit."""
    from vulcan_amd import synth
    cols = synth.columns_same_network(1002, 16, 150, 93, 1200, 60, dt=1e-10, name="crwide")
    with factor(VK_FACTOR_CR):
        batch, st = _batch_step(cols)
        alone = [_batch_step([c])[0][0] for c in cols[:4]] + [None] * 12
    assert not st.any()
    for c in range(4):
        assert np.array_equal(batch[c], alone[c]), c
    for c in range(16):
        assert np.all(np.isfinite(batch[c]))
