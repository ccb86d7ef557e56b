"""CPU: host-side checks of the run configuration and the workload generators (no GPU calls)."""
from __future__ import annotations

import pytest


def test_conv_mode_is_validated():
    from vulcan_amd.outer_loop import RunConfig
    p = RunConfig(conv_mode="longdy", conv_step=40, yconv_cri=0.02).c_params(1)
    assert p.conv_mode == 1 and p.conv_step == 40 and p.yconv_cri == 0.02
    assert p.st_factor == 0.5 and p.conv_stall_window == 200 and p.runtime == 1e22
    assert RunConfig().c_params(1).conv_mode == 0
    with pytest.raises(ValueError):
        RunConfig(conv_mode="window").c_params(1)


def test_serving_refuses_the_longdy_rule():
    from vulcan_amd.outer_loop import RunConfig
    from vulcan_amd.serving import ServingEngine
    with pytest.raises(ValueError, match="relmax"):
        ServingEngine(cfg=RunConfig(conv_mode="longdy"))


def test_vul_end_case_codes():
    """The .vul end_case follows the reference (outer_loop.py:1938-1942,
    2326-2339): 3 when the step cap was exceeded, 2 when the runtime was,
    1 otherwise; never the device status word."""
    from vulcan_amd.output import end_case
    from vulcan_amd.outer_loop import RunConfig, ST_CONVERGED, ST_MAXSTEPS
    ld = RunConfig(conv_mode="longdy", count_max=100, runtime=1e5)
    assert end_case(101, 1.0, ST_MAXSTEPS, ld) == 3
    assert end_case(100, 1.0, 0, ld) == 1
    assert end_case(50, 2e5, ST_MAXSTEPS, ld) == 2
    assert end_case(50, 1.0, ST_CONVERGED, ld) == 1
    rm = RunConfig(count_max=100)
    assert end_case(100, 1.0, ST_MAXSTEPS, rm) == 3
    assert end_case(100, 1.0, ST_MAXSTEPS | ST_CONVERGED, rm) == 1
    assert end_case(100, 1.0, ST_CONVERGED | 2, rm) == 1
    assert end_case(40, 1.0, 0, rm) == 1


def test_config4_rate_tangent_matches_leaf_differences():
    """The exact d log k / d log T of the differentiation workload's leaf
    generator against central differences of the generator itself."""
    import numpy as np
    from vulcan_amd import synth
    pb = synth.config4()
    s = pb.dlogk_dlogT(pb.T0)
    for z in (0, 57, 149):
        h = 1e-6
        Tp, Tm = pb.T0.copy(), pb.T0.copy()
        Tp[z] *= np.exp(h)
        Tm[z] *= np.exp(-h)
        kp, _ = pb.leaves(Tp)
        km, _ = pb.leaves(Tm)
        k0, _ = pb.leaves(pb.T0)
        fd = (kp[:, z] - km[:, z]) / (2 * h)
        exact = np.concatenate([[0.0], np.repeat(s[:, z], 2)]) * k0[:, z]
        np.testing.assert_allclose(fd, exact, rtol=1e-6, atol=1e-300)
        # other layers do not move
        assert np.array_equal(np.delete(kp, z, axis=1), np.delete(k0, z, axis=1))


def test_make_atm_static_follows_the_configuration():
    """jax_step.make_atm_static (jax_step.py:318-351): flags from the run
    configuration, molecular-diffusion leaves zeroed where it switches them
    off, the gas mask from gas_indx."""
    import types
    import numpy as np
    from vulcan_amd import jax_step
    rng = np.random.default_rng(2)
    nz, ni = 6, 4
    legacy = types.SimpleNamespace(
        Kzz=rng.random(nz - 1), Dzz=rng.random((nz - 1, ni)), dzi=rng.random(nz - 1),
        vz=rng.random(nz - 1), Hpi=rng.random(nz - 1), Ti=rng.random(nz - 1),
        Tco=rng.random(nz), g=rng.random(nz), ms=rng.random(ni), alpha=rng.random(ni),
        M=rng.random(nz), vm=rng.random((nz, ni)), vs=rng.random((nz - 1, ni)),
        top_flux=rng.random(ni), bot_flux=rng.random(ni), bot_vdep=rng.random(ni),
        gas_indx=[0, 2, 3])
    a = jax_step.make_atm_static(legacy, ni, nz, cfg=types.SimpleNamespace())
    assert not (a.use_vm_mol or a.use_settling or a.use_topflux or a.use_botflux)
    assert np.array_equal(a.Dzz, legacy.Dzz) and not a.vm.any() and not a.vs.any()
    assert a.gas_indx_mask.tolist() == [True, False, True, True]
    assert a.vm.shape == (nz, ni) and a.vs.shape == (nz - 1, ni)
    cfg = types.SimpleNamespace(use_vm_mol=True, use_settling=True, use_topflux=True,
                                use_botflux=True)
    a = jax_step.make_atm_static(legacy, ni, nz, cfg=cfg)
    assert a.use_vm_mol and a.use_settling and a.use_topflux and a.use_botflux
    assert np.array_equal(a.vm, legacy.vm) and np.array_equal(a.vs, legacy.vs)
    # molecular diffusion off: Dzz, vm and vs zeroed, vm / settling flags off
    cfg.use_moldiff = False
    a = jax_step.make_atm_static(legacy, ni, nz, cfg=cfg)
    assert not (a.use_vm_mol or a.use_settling) and a.use_topflux and a.use_botflux
    assert not (a.Dzz.any() or a.vm.any() or a.vs.any())
    for f in ("Kzz", "dzi", "Tco", "M", "bot_vdep"):
        assert np.array_equal(getattr(a, f), getattr(legacy, f))


def test_photo_data_from_static_selects_rows_and_indices():
    """photo.photo_data_from_static (photo.py:35-81): absorbers with a
    temperature-dependent cross section leave the plain table, the T table
    passes through, indices follow species_list."""
    import types
    import numpy as np
    from vulcan_amd import photo
    species = ["H2", "H2O", "CH4", "CO", "CO2", "NH3", "N2", "He"]
    nbin, nz = 7, 3
    rng = np.random.default_rng(1)
    static = types.SimpleNamespace(
        nbin=nbin, absp_sp=["H2O", "CH4", "CO2", "NH3"], absp_cross=rng.random((4, nbin)),
        absp_T_sp=["CO2"], absp_T_cross=rng.random((1, nz, nbin)),
        scat_sp=["H2", "He"], scat_cross=rng.random((2, nbin)))
    pd = photo.photo_data_from_static(static, species)
    assert pd.absp_idx.tolist() == [1, 2, 5]
    assert np.array_equal(pd.absp_cross, static.absp_cross[[0, 1, 3]])
    assert pd.absp_T_idx.tolist() == [4]
    assert np.array_equal(pd.absp_T_cross, static.absp_T_cross)
    assert pd.scat_idx.tolist() == [0, 7]
    assert np.array_equal(pd.scat_cross, static.scat_cross)
    empty = types.SimpleNamespace(nbin=nbin, absp_sp=[], absp_cross=np.zeros((0, nbin)),
                                  absp_T_sp=[], absp_T_cross=np.zeros((0, nz, nbin)),
                                  scat_sp=[], scat_cross=np.zeros((0, nbin)))
    pd = photo.photo_data_from_static(empty, species)
    assert pd.absp_idx.size == 0 and pd.absp_cross.shape == (0, nbin)
    assert pd.absp_T_idx.size == 0 and pd.absp_T_cross.shape == (0, nz, nbin)
    assert pd.scat_idx.size == 0 and pd.scat_cross.shape == (0, nbin)
