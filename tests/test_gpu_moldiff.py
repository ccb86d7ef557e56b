"""GPU: the Ros2 step with molecular diffusion, vm, settling and boundary
fluxes switched on (jax_step.py:67-246).

The synthetic workloads of the other parity tests carry no molecular
diffusion, so these columns exercise the molecular terms:
* the step's per-step precompute of the y-independent factors must give
  the same step, bit for bit, as forming them inside the chemistry
  launches (VK_TUNE_MOLPRE 1 vs 0);
* the step must agree with the oracle at a dt where diffusion moves the
  state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VK_TUNE_MOLPRE = 1


def _column(seed, nz, ni, flags):
    from vulcan_amd import synth
    c = synth.make_column(seed, nz=nz, ni=ni, n_rev=3 * ni, dt=1e-6, p_bot=1e5, p_top=1e1)
    rng = np.random.default_rng(seed + 7)
    a = c.atm
    sgn = lambda shape: rng.choice([-1.0, 1.0], size=shape)
    atm = a._replace(
        Dzz=a.Kzz[:, None] * rng.uniform(0.05, 0.5, (nz - 1, ni)),
        vz=sgn(nz - 1) * rng.uniform(1.0, 10.0, nz - 1),
        vm=sgn((nz, ni)) * rng.uniform(0.5, 5.0, (nz, ni)),
        vs=sgn((nz - 1, ni)) * rng.uniform(0.5, 5.0, (nz - 1, ni)),
        alpha=rng.uniform(-0.4, 0.4, ni),
        ms=rng.uniform(2.0, 60.0, ni),
        top_flux=rng.uniform(-1e8, 1e8, ni), bot_flux=rng.uniform(1e8, 1e10, ni),
        bot_vdep=rng.uniform(0.01, 1.0, ni),
        gas_indx_mask=np.array([True] * (ni - 2) + [False, False]),
        use_vm_mol=flags[0], use_settling=flags[1], use_topflux=True, use_botflux=True)
    return c, atm


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("flags", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("ni", [21, 93])
def test_molecular_factor_precompute_is_bitwise(flags, ni):
    """VK_TUNE_MOLPRE 1 (k_mol_static once per step) and 0 (the factors
    formed in each chemistry launch) give identical steps."""
    import torch
    from vulcan_amd import _lib, jax_step
    lib = _lib.load()
    c, atm = _column(41 + ni, 40, ni, flags)
    out = {}
    try:
        for v in (1, 0):
            _lib.check(lib.vk_tune_set(VK_TUNE_MOLPRE, 0, v), "vk_tune_set")
            assert lib.vk_tune_get(VK_TUNE_MOLPRE, 0) == v
            sol, delta = jax_step.jax_ros2_step(c.y, c.k, 1e-3, atm, c.net)
            out[v] = (sol.clone(), delta.clone())
    finally:
        lib.vk_tune_set(VK_TUNE_MOLPRE, 0, 1)
    assert torch.isfinite(out[1][0]).all()
    assert torch.equal(out[1][0], out[0][0])
    assert torch.equal(out[1][1], out[0][1])


@pytest.mark.parametrize("flags", [(False, False), (True, True)])
def test_ros2_step_molecular_diffusion_matches_oracle(flags):
    """Per-step parity with the molecular terms on, at dt where diffusion
    changes the state (the step moves y by well over 1e-6 relative).  As in
    the workload-1 parity test, two correct LU codes (the NumPy and C
    oracles) set the floor on ill-conditioned blocks: 1e-10 of each layer's
    scale where they agree to 1e-12, and within 10x their spread elsewhere."""
    from oracle import ros2_c, ros2_np as O
    from vulcan_amd import core, jax_step
    c, atm = _column(977, 30, 21, flags)
    dt = 1e-4
    sol, _ = jax_step.jax_ros2_step(c.y, c.k, dt, atm, c.net)
    so, _ = O.ros2_step(c.y, c.k, dt, atm, c.net)
    rec = core.pack_atm_record(atm).numpy()[None]
    sc, _ = ros2_c.ros2_steps(c.y[None], c.k[None], rec, dt, c.net)
    sol = _np(sol)
    scale = np.max(np.abs(so), axis=1)
    moved = np.max(np.abs(so - c.y), axis=1) / scale
    err = np.max(np.abs(sol - so), axis=1) / scale
    floor = np.max(np.abs(sc[0] - so), axis=1) / scale
    well = floor < 1e-12
    print(f"flags {flags}: step moves y by up to {moved.max():.2e}; well-determined layers "
          f"{well.sum()}/{c.nz}, gpu err there {err[well].max():.2e}; worst {err.max():.2e} "
          f"vs oracle spread {floor.max():.2e}")
    assert moved.max() > 1e-6
    assert well.sum() >= c.nz // 2
    assert err[well].max() < 1e-10
    assert err.max() <= 10.0 * floor.max() + 1e-10
    from parity_util import StepRefs
    r = StepRefs(c.y, c.k, dt, atm, c.net)
    e_det, e_all, spread, line = r.judge(sol, f"flags {flags} elementwise")
    print(line)
    assert e_det < 1e-10
    assert e_all <= 10.0 * spread + 1e-10
