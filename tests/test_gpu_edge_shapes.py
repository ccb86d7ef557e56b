"""GPU: the Ros2 step at the edges of the supported shapes, against the
three reference solves (elementwise on the cells they determine to 1e-12,
bar 1e-10, and no counted cell further than 10x their spread): the fewest
layers (nz = 2), one species, the 16-species bucket boundaries (ni = 16,
17), the largest species count (ni = 112, the 7 x 7 tile) and a tall
column (nz = 300); plus a batch of those shapes together, each column equal
bit for bit to its step alone."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available
from parity_util import StepRefs

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

SHAPES = [(2, 5, 12), (6, 1, 2), (20, 16, 40), (20, 17, 40), (12, 112, 300), (300, 21, 60)]


def _col(nz, ni, n_rev, seed):
    from vulcan_amd import synth
    return synth.make_column(seed, nz=nz, ni=ni, n_rev=n_rev, dt=1e-10, name=f"edge{nz}x{ni}",
                             p_bot=1e7, p_top=1e1, consistent=True)


@pytest.mark.parametrize("nz,ni,n_rev", SHAPES, ids=[f"nz{a}-ni{b}" for a, b, _ in SHAPES])
def test_step_at_edge_shapes(nz, ni, n_rev):
    from vulcan_amd import jax_step
    c = _col(nz, ni, n_rev, 4000 + nz + ni)
    sol, delta = jax_step.jax_ros2_step(c.y, c.k, c.dt, c.atm, c.net)
    s = sol.cpu().numpy()
    assert np.all(np.isfinite(s)) and np.all(np.isfinite(delta.cpu().numpy()))
    r = StepRefs(c.y, c.k, c.dt, c.atm, c.net)
    e_det, e_all, spread, line = r.judge(s, f"nz={nz} ni={ni}")
    print(line)
    assert e_det < 1e-10
    assert e_all <= 10.0 * spread + 1e-10


def test_mixed_edge_batch_is_bitwise_alone():
    import torch
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    cols = [_col(nz, ni, n_rev, 5000 + i) for i, (nz, ni, n_rev) in enumerate(SHAPES)]
    r = SteadyStateRunner(cols, RunConfig())
    for _ in range(3):
        r.attempt()
    torch.cuda.synchronize()
    res = r.results()
    for i, c in enumerate(cols):
        a = SteadyStateRunner([c], RunConfig())
        for _ in range(3):
            a.attempt()
        ra = a.results()
        assert np.array_equal(ra["y"][0], res["y"][i]), i
        assert ra["accept"][0] == res["accept"][i]


def test_held_inverse_sweeps_bitwise_equal_every_bucket():
    """VK_TUNE_HOLDX 1 (X_j kept in registers across a refined layer) and 0
    (X_j read again) give the same attempts, bit for bit, in every tile
    size bucket of the edge shapes (SL 1, 2 and 7)."""
    import torch
    from vulcan_amd import _lib
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    lib = _lib.load()
    VK_TUNE_HOLDX = 5
    cols = [_col(nz, ni, n_rev, 5000 + i) for i, (nz, ni, n_rev) in enumerate(SHAPES)]
    out = []
    for hold in (1, 0):
        _lib.check(lib.vk_tune_set(VK_TUNE_HOLDX, 0, hold), "vk_tune_set")
        try:
            r = SteadyStateRunner(cols, RunConfig())
            for _ in range(3):
                r.attempt()
            torch.cuda.synchronize()
            out.append(r.results())
        finally:
            lib.vk_tune_set(VK_TUNE_HOLDX, 0, 1)
    for key in ("y", "dt", "accept"):
        for i in range(len(cols)):
            assert np.array_equal(np.asarray(out[0][key][i]), np.asarray(out[1][key][i])), (key, i)
