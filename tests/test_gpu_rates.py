"""GPU: rate construction k(T, M) against the oracle.

The device evaluates the oracle's expressions in the same order (fp
contraction off) except that a T^n exp(-E/T) is formed as
a exp(n ln T - E/T), one exponential with ln T shared by the layer; the
differences are that rewrite and the device's pow, exp and log against the
host C library.  Forward rates and caps agree to rtol 1e-13 (the rewrite
costs at most ~3e-14 for |E/T| <= 700); the Gibbs energies to 1e-12 of
their scale; K_eq from the same Gibbs input to 1e-13; the reverse fill from
the same k and K bit for bit.  In a full build the reverse rows inherit the
Gibbs differences through exp(sum of g), hence rtol 1e-11.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "NASA9"


def _np(t):
    return t.detach().cpu().numpy()


def _close(gpu, ref, rtol):
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    assert gpu.shape == ref.shape
    zero = ref == 0.0
    assert np.array_equal(gpu[zero], ref[zero]), "zero pattern differs"
    fin = np.isfinite(ref) & ~zero
    assert np.array_equal(np.isfinite(gpu), np.isfinite(ref))
    err = np.max(np.abs(gpu[fin] - ref[fin]) / np.abs(ref[fin]), initial=0.0)
    assert err <= rtol, f"max relative error {err:.3e} > {rtol:.1e}"
    return err


def _setup(seed=11, ni=93, n_rev=1200, n_photo=60, nz=150):
    from vulcan_amd import synth
    rng = np.random.default_rng(seed)
    base, _, _ = synth.random_network(rng, ni, n_rev, n_photo)
    net = synth.random_rate_network(rng, base, n_rev, n_photo)
    coeffs = synth.random_nasa9(rng, ni)
    T = np.concatenate([np.linspace(120.0, 320.0, nz // 3),
                        np.linspace(320.0, 2800.0, nz - nz // 3)])
    T[5] = 277.5
    T[7] = 1000.0
    M = np.logspace(19.5, 7.0, nz)
    return base, net, coeffs, T, M


def test_forward_and_caps_match_oracle():
    from oracle import rates_np as O
    from vulcan_amd import rates
    _, net, _, T, M = _setup()
    kf = rates.compute_forward_k(net, T, M)
    ref = O.compute_forward_k(net, T, M)
    e1 = _close(_np(kf), ref, 1e-13)
    capped = rates.apply_lowT_caps(net, kf, T, M)
    ref_c = O.apply_lowT_caps(net, _np(kf), T, M)
    e2 = _close(_np(capped), ref_c, 1e-13)
    assert not np.array_equal(ref_c, _np(kf))     # the caps fired
    print(f"forward {e1:.2e}, caps {e2:.2e}")


def test_gibbs_keq_fill_match_oracle():
    from oracle import rates_np as O
    from vulcan_amd import gibbs
    _, net, coeffs, T, M = _setup()
    g = _np(gibbs.gibbs_sp_vector(coeffs, T))
    g_ref = O.gibbs_sp_vector(coeffs, T)
    scale = 1.0 + np.abs(g_ref)
    assert np.max(np.abs(g - g_ref) / scale) < 1e-12
    with np.errstate(over="ignore"):
        K_ref = O.K_eq_array(net, g_ref, T)
    K = _np(gibbs.K_eq_array(net, g_ref, T))
    _close(K, K_ref, 1e-13)
    k_f = O.compute_forward_k(net, T, M)
    for remove in (None, [1, 6, 40, net.stop_rev_indx + 2]):
        with np.errstate(divide="ignore", invalid="ignore"):
            ref = O.fill_reverse_k(net, k_f, K_ref, remove_list=remove)
        out = _np(gibbs.fill_reverse_k(net, k_f, K_ref, remove_list=remove))
        assert np.array_equal(out, ref)


def test_gibbs_real_species_match_oracle():
    """The reference's NASA-9 data for ten species, both ranges."""
    from oracle import rates_np as O
    from vulcan_amd import gibbs
    sp = ("H", "H2", "O", "O2", "OH", "H2O", "CO", "CO2", "CH3", "CH4")
    c, present = gibbs.load_nasa9(sp, GOLD.parent)
    assert present.all()
    assert np.array_equal(c, O.load_nasa9([GOLD / f"{s}.txt" for s in sp]))
    T = np.linspace(100.0, 3000.0, 59)
    g = _np(gibbs.gibbs_sp_vector(c, T))
    g_ref = O.gibbs_sp_vector(c, T)
    assert np.max(np.abs(g - g_ref) / (1.0 + np.abs(g_ref))) < 1e-12


@pytest.mark.parametrize("lowt,remove", [(False, None), (True, None), (True, [1, 4, 9, 33])])
def test_build_rate_array_matches_oracle(lowt, remove):
    from types import SimpleNamespace
    from oracle import rates_np as O
    from vulcan_amd import rates
    _, net, coeffs, T, M = _setup()
    cfg = SimpleNamespace(use_lowT_limit_rates=lowt, remove_list=remove)
    atm = SimpleNamespace(Tco=T, M=M)
    k = _np(rates.build_rate_array(cfg, net, atm, coeffs))
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        ref = O.build_rate_array(net, T, M, coeffs, use_lowT_limit_rates=lowt, remove_list=remove)
    err = _close(k, ref, 1e-11)
    if remove:
        for i in remove:
            assert np.all(k[i] == 0.0)
        assert np.any(k[10] != 0.0)        # row 9 removed, its reverse kept
    print(f"build lowt={lowt} remove={remove}: {err:.2e}")


def test_compute_all_k_paired_removal():
    from oracle import rates_np as O
    from vulcan_amd import gibbs
    _, net, coeffs, T, M = _setup(seed=12, ni=50, n_rev=300, n_photo=10, nz=60)
    remove = [3, 10, 77]
    k = _np(gibbs.compute_all_k(net, T, M, coeffs, remove_list=remove))
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        ref = O.compute_all_k(net, T, M, coeffs, remove_list=remove)
    _close(k, ref, 1e-11)
    for i in (3, 4, 9, 10, 77, 78):
        assert np.all(k[i] == 0.0)


def test_batch_build_matches_single_builds():
    """An uneven batch (three networks, nz from 60 to 200) built in one
    launch equals each column's own build_rate_array bit for bit, and the
    oracle to rtol 1e-11."""
    import torch
    from types import SimpleNamespace
    from oracle import rates_np as O
    from vulcan_amd import core, rates, synth
    rng = np.random.default_rng(3)
    nets, rnets, c9 = [], [], []
    for ni, n_rev in ((50, 300), (69, 500), (93, 1200)):
        base, _, _ = synth.random_network(rng, ni, n_rev, 0)
        nets.append(core.device_network(base))
        rnets.append(synth.random_rate_network(rng, base, n_rev, 0))
        c9.append(synth.random_nasa9(rng, ni))
    ncols = 24
    nzs = rng.integers(60, 201, size=ncols).tolist()
    ids = rng.integers(0, 3, size=ncols).tolist()
    batch = core.Batch(nzs, nets, ids)
    atms = []
    for c in range(ncols):
        col = synth.make_column(500 + c, nz=nzs[c], ni=nets[ids[c]].ni, n_rev=3)
        atm = col.atm
        if c % 2 == 0:      # cold columns: the low-T caps fire
            atm = atm._replace(Tco=np.linspace(130.0, 1300.0, nzs[c]))
        atms.append(atm)
    atm_packed = batch.pack_atm(atms)
    remove = [[5], None, [2, 7]]
    for lowt in (False, True):
        k = rates.build_rate_batch(batch, rnets, c9, atm_packed, lowt, remove)
        torch.cuda.synchronize()
        for c in range(ncols):
            j = ids[c]
            S = rnets[j].nr + 1
            got = _np(k[batch.off_k[c]:batch.off_k[c] + S * nzs[c]]).reshape(S, nzs[c])
            cfg = SimpleNamespace(use_lowT_limit_rates=lowt, remove_list=remove[j])
            one = _np(rates.build_rate_array(cfg, rnets[j], atms[c], c9[j]))
            assert np.array_equal(got, one), f"column {c}"
            if c % 6 == 0:
                with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
                    ref = O.build_rate_array(rnets[j], atms[c].Tco, atms[c].M, c9[j],
                                             use_lowT_limit_rates=lowt, remove_list=remove[j])
                _close(got, ref, 1e-11)


def test_rates_errors():
    from vulcan_amd import _lib, gibbs, rates
    _, net, coeffs, T, M = _setup(ni=20, n_rev=30, n_photo=0, nz=10)
    with pytest.raises(ValueError):
        rates.compute_forward_k(net, T, M[:-1])
    with pytest.raises(ValueError):
        gibbs.K_eq_array(net, np.zeros((19, 10)), T)
    d = rates.DeviceRates(net)               # no NASA-9 table
    with pytest.raises(_lib.VulcanError):
        d.eval(rates.RATES_BUILD, T, M)


def _cap_free_grid(nz, lo, hi, seed):
    """Layer temperatures at least 0.5 K away from the cap thresholds and
    the NASA-9 range switch, where k(T) has a jump."""
    rng = np.random.default_rng(seed)
    T = np.sort(rng.uniform(lo, hi, nz))
    for edge in (200.0, 277.5, 300.0, 1000.0):
        near = np.abs(T - edge) < 0.5
        T[near] += 1.0
    return T


@pytest.mark.parametrize("lowt", [False, True])
def test_rates_vjp_matches_oracle_central_differences(lowt):
    """T_bar and M_bar of the device VJP against five-point central
    differences of the oracle's build_rate_array in T and M (h = 1e-5 of T,
    1e-4 of M: truncation O(h^4), rounding ~1e-11 of each term), with k_bar
    weighted by 1 / |k| so every row counts: the VJP is then
    sum_i w_i d ln k_i / dT.  Agreement to 1e-6 of the sum of |terms|."""
    from oracle import rates_np as O
    from vulcan_amd import rates
    _, net, coeffs, _, _ = _setup(seed=21, ni=50, n_rev=300, n_photo=10, nz=40)
    nz = 40
    T = _cap_free_grid(nz, 130.0, 2600.0, 5)
    M = np.logspace(19.0, 9.0, nz)
    with np.errstate(all="ignore"):
        k0 = O.build_rate_array(net, T, M, coeffs, use_lowT_limit_rates=lowt)
    rng = np.random.default_rng(6)
    ok = np.isfinite(k0) & (np.abs(k0) > 1e-290) & (np.abs(k0) < 1e290)
    w = np.where(ok, rng.uniform(-1.0, 1.0, k0.shape) / np.where(ok, np.abs(k0), 1.0), 0.0)
    d = rates.device_rates(net, coeffs, None)
    Tb, Mb = d.vjp(T, M, w, lowt)
    Tb, Mb = _np(Tb), _np(Mb)
    for name, x, g in (("T", T, Tb), ("M", M, Mb)):
        worst = 0.0
        for z in range(nz):
            h = (1e-5 if name == "T" else 1e-4) * x[z]
            ks = []
            for step in (2.0, 1.0, -1.0, -2.0):
                xs = x.copy()
                xs[z] += step * h
                args = (xs, M) if name == "T" else (T, xs)
                with np.errstate(all="ignore"):
                    ks.append(O.build_rate_array(net, *args, coeffs,
                                                 use_lowT_limit_rates=lowt)[:, z])
            d = (-ks[0] + 8.0 * ks[1] - 8.0 * ks[2] + ks[3]) / (12.0 * h)
            terms = w[:, z] * d
            fd = terms.sum()
            scale = np.abs(terms).sum() + 1e-300
            worst = max(worst, abs(g[z] - fd) / scale)
        print(f"lowt={lowt} {name}_bar: worst |vjp - fd| / sum|terms| = {worst:.2e}")
        assert worst < 1e-6


def test_differentiable_build_autograd():
    """torch autograd through build_rate_array_differentiable returns the
    VJP of the device build, and the forward equals build_rate_array."""
    import torch
    from types import SimpleNamespace
    from vulcan_amd import rates
    _, net, coeffs, _, _ = _setup(seed=22, ni=30, n_rev=80, n_photo=0, nz=16)
    T = torch.tensor(_cap_free_grid(16, 400.0, 1800.0, 8), dtype=torch.float64,
                     device="cuda", requires_grad=True)
    M = torch.tensor(np.logspace(18.0, 12.0, 16), dtype=torch.float64, device="cuda",
                     requires_grad=True)
    k = rates.build_rate_array_differentiable(net, T, M, coeffs)
    ref = rates.build_rate_array(SimpleNamespace(use_lowT_limit_rates=False, remove_list=None),
                                 net, SimpleNamespace(Tco=T.detach(), M=M.detach()), coeffs)
    assert torch.equal(k.detach(), ref)
    w = torch.where(k.detach() != 0, 1.0 / k.detach().abs().clamp_min(1e-300),
                    torch.zeros_like(k))
    (k * w).sum().backward()
    Tb, Mb = rates.device_rates(net, coeffs, None).vjp(T.detach(), M.detach(), w)
    assert torch.equal(T.grad, Tb) and torch.equal(M.grad, Mb)
    assert torch.isfinite(T.grad).all() and T.grad.abs().max() > 0
