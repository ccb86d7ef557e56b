"""GPU: the remaining names of the reference's step API.

* solver.block_thomas (solver.py:129-176), the dense-off-diagonal block
  Thomas: the reference's seeded system (test_block_thomas.py:22-50, seed
  42, nz=8, ni=5) against a dense LAPACK solve at its bar 1e-9, and larger
  systems (ni 21 .. 100, one needing pivoting) against the oracle's LU block
  Thomas and LAPACK; a batch equals its columns alone bit for bit; the
  diagonal-off-diagonal system through the dense path matches
  block_thomas_diag_offdiag.
* jax_step.compute_diff_grav (jax_step.py:67-147) against the oracle's
  restatement, with molecular diffusion, thermal diffusion, molecular
  advection (use_vm_mol) and settling switched on and off.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _np(t):
    return t.detach().cpu().numpy()


def _rel(x, ref, floor=1e-300):
    return float(np.max(np.abs(x - ref) / np.maximum(np.abs(ref), floor)))


def test_block_thomas_dense_reference_seed42():
    from oracle import ros2_np as O
    from vulcan_amd import solver
    rng = np.random.default_rng(42)
    nz, ni = 8, 5
    diag = rng.standard_normal((nz, ni, ni)) + 10.0 * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni, ni))
    sub = rng.standard_normal((nz - 1, ni, ni))
    rhs = rng.standard_normal((nz, ni))
    x = _np(solver.block_thomas(diag, sup, sub, rhs))
    ref = np.linalg.solve(O.assemble_dense(diag, sup, sub), rhs.reshape(-1)).reshape(nz, ni)
    e = float(np.max(np.abs(x - ref) / np.maximum(np.abs(ref), 1e-12)))
    print(f"seed 42, nz=8 ni=5: {e:.2e} vs dense")
    assert e < 1e-9


@pytest.mark.parametrize("nz,ni,scale", [(30, 21, 4.0), (40, 69, 4.0), (24, 93, 4.0), (20, 37, 0.0),
                                         (12, 100, 4.0)])
def test_block_thomas_dense_vs_oracle_and_lapack(nz, ni, scale):
    """scale 0: no diagonal shift, so the blocks need row exchanges."""
    from oracle import ros2_np as O
    from vulcan_amd import solver
    rng = np.random.default_rng(1000 + ni)
    diag = rng.standard_normal((nz, ni, ni)) + scale * np.sqrt(ni) * np.eye(ni)
    sup = rng.standard_normal((nz - 1, ni, ni)) / np.sqrt(ni)
    sub = rng.standard_normal((nz - 1, ni, ni)) / np.sqrt(ni)
    rhs = rng.standard_normal((nz, ni))
    x = _np(solver.block_thomas(diag, sup, sub, rhs))
    xo = O.block_thomas_dense(diag, sup, sub, rhs)
    A = O.assemble_dense(diag, sup, sub)
    ref = np.linalg.solve(A, rhs.reshape(-1)).reshape(nz, ni)
    # normwise, as the errors of two elimination orders compare
    e_ref = np.abs(x - ref).max() / np.abs(ref).max()
    e_o = np.abs(xo - ref).max() / np.abs(ref).max()
    bw = np.abs(A @ x.reshape(-1) - rhs.reshape(-1)).max() / (
        np.abs(A).sum(axis=1).max() * np.abs(x).max())
    print(f"nz={nz} ni={ni} shift={scale}: GPU {e_ref:.2e} vs LAPACK (oracle LU "
          f"block Thomas {e_o:.2e}); backward error {bw:.2e}")
    assert e_ref <= 10.0 * e_o + 1e-13
    assert bw < 1e-14


def test_block_thomas_dense_batch_and_diag_form():
    from vulcan_amd import solver
    rng = np.random.default_rng(5)
    B, nz, ni = 3, 16, 29
    diag = rng.standard_normal((B, nz, ni, ni)) + 6.0 * np.eye(ni)
    sup = rng.standard_normal((B, nz - 1, ni, ni)) * 0.3
    sub = rng.standard_normal((B, nz - 1, ni, ni)) * 0.3
    rhs = rng.standard_normal((B, nz, ni))
    xb = _np(solver.block_thomas(diag, sup, sub, rhs))
    for c in range(B):
        xc = _np(solver.block_thomas(diag[c], sup[c], sub[c], rhs[c]))
        assert np.array_equal(xb[c], xc)
    # diagonal off-diagonal blocks through the dense path
    sd = rng.standard_normal((nz - 1, ni))
    bd = rng.standard_normal((nz - 1, ni))
    x1 = _np(solver.block_thomas(diag[0], np.stack([np.diag(v) for v in sd]),
                                 np.stack([np.diag(v) for v in bd]), rhs[0]))
    x2 = _np(solver.block_thomas_diag_offdiag(diag[0], sd, bd, rhs[0]))
    e = np.abs(x1 - x2).max() / np.abs(x2).max()
    print(f"dense path vs diagonal path on a diagonal-off-diagonal system: {e:.2e}")
    assert e < 1e-13


@pytest.mark.parametrize("use_vm,use_set", [(False, False), (True, False), (False, True),
                                            (True, True)])
def test_compute_diff_grav_matches_oracle(use_vm, use_set):
    from oracle import ros2_np as O
    from vulcan_amd import jax_step, synth
    col = synth.make_column(31, nz=40, ni=21, n_rev=40)
    rng = np.random.default_rng(9)
    nz, ni = col.nz, col.ni
    atm = col.atm._replace(
        Dzz=10.0 ** rng.uniform(3.0, 7.0, size=(nz - 1, ni)),
        ms=rng.uniform(1.0, 80.0, size=ni), alpha=rng.uniform(-0.4, 0.3, size=ni),
        vm=rng.normal(size=(nz, ni)) * 1e2, vs=rng.normal(size=(nz - 1, ni)) * 1e1,
        use_vm_mol=use_vm, use_settling=use_set)
    g = jax_step.compute_diff_grav(atm)
    go = O.compute_diff_grav(atm)
    worst = 0.0
    for f, a, b in zip(jax_step.DiffGrav._fields, g, go):   # same field order
        a, b = _np(a), np.asarray(b)
        assert a.shape == b.shape, f
        scale = np.abs(b).max()
        e = float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a).max())
        worst = max(worst, e)
        assert e < 1e-14, (f, e)
    print(f"use_vm_mol={use_vm} use_settling={use_set}: worst field {worst:.2e}")
