"""CPU oracle for the Ros2 column step — TEST INFRASTRUCTURE ONLY.

A plain NumPy/SciPy restatement of VULCAN-JAX's FP64 hot path, written from
the reference source read as text.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only
to check or time against it; the product path never calls it.

Each function cites the reference file:line it restates.  Pins that make the
restatement trustworthy (closed forms, known-answer tests, dense solves,
invariants) live in ``tests/test_oracle_*.py``.  Summation order: the
reference's ``jax.ops.segment_sum`` has no documented order; this oracle
accumulates with ``np.add.at`` (sequential, ascending flat index = reaction
row, then slot), which is the order the device kernels use too.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np
import scipy.linalg as sla

KB = 1.38064852e-16   # phy_const.py:3
NAVO = 6.02214086e23  # phy_const.py:4
GAMMA = 1.0 + 1.0 / np.sqrt(2.0)  # jax_step.py:257


# ---------------------------------------------------------------------------
# chemistry (chem.py)
# ---------------------------------------------------------------------------

def _rates(y, M, k, net):
    """rate[z, r] = k * prod_slots y**nu (* M if three-body); chem.py:92-99."""
    nz, ni = y.shape
    yp = np.concatenate([y, np.ones((nz, 1))], axis=1)
    y_r = yp[:, net.reactant_idx]                       # (nz, S, T)
    st = net.reactant_stoich[None]
    factor = np.where(st > 0, y_r ** np.where(st > 0, st, 1.0), 1.0)
    prod_r = np.ones(factor.shape[:2])
    for s in range(factor.shape[2]):
        prod_r = prod_r * factor[:, :, s]
    rate = k.T * prod_r
    rate = np.where(net.is_three_body[None, :], rate * M[:, None], rate)
    return rate, y_r, factor


def chem_rhs(y, M, k, net):
    """dy/dt from chemistry = prod - loss, summed separately; chem.py:83-131."""
    nz, ni = y.shape
    rate, _, _ = _rates(y, M, k, net)
    S, T = net.reactant_idx.shape
    out = np.zeros((nz, ni))
    for z in range(nz):
        rr = np.repeat(rate[z], T)
        loss = np.zeros(ni + 1)
        prod = np.zeros(ni + 1)
        np.add.at(loss, net.reactant_idx.reshape(-1), net.reactant_stoich.reshape(-1) * rr)
        np.add.at(prod, net.product_idx.reshape(-1), net.product_stoich.reshape(-1) * rr)
        out[z] = prod[:ni] - loss[:ni]
    return out


def chem_rhs_loop(y, M, k, net):
    """Explicit per-reaction loop of chem.chem_rhs_numpy (chem.py:217-243)."""
    nz, ni = y.shape
    dydt = np.zeros_like(y)
    for i in range(1, net.nr + 1):
        rate = np.array(k[i], dtype=np.float64)
        for s in range(net.reactant_idx.shape[1]):
            sp, st = net.reactant_idx[i, s], net.reactant_stoich[i, s]
            if st == 0:
                continue
            rate = rate * y[:, sp] ** st
        if net.is_three_body[i]:
            rate = rate * M
        for s in range(net.product_idx.shape[1]):
            sp, st = net.product_idx[i, s], net.product_stoich[i, s]
            if st == 0 or sp >= ni:
                continue
            dydt[:, sp] += st * rate
        for s in range(net.reactant_idx.shape[1]):
            sp, st = net.reactant_idx[i, s], net.reactant_stoich[i, s]
            if st == 0 or sp >= ni:
                continue
            dydt[:, sp] -= st * rate
    return dydt


def chem_jac(y, M, k, net):
    """Analytical stoichiometric Jacobian (nz, ni, ni); chem.py:144-214."""
    nz, ni = y.shape
    S, T = net.reactant_idx.shape
    _, y_r, factor = _rates(y, M, k, net)
    st = net.reactant_stoich[None]
    leave = np.ones((nz, S, T))
    for j in range(T):
        p = np.ones((nz, S))
        for l in range(T):
            if l != j:
                p = p * factor[:, :, l]
        leave[:, :, j] = p
    pm1 = np.where(st > 0, y_r ** np.where(st > 0, st - 1.0, 0.0), 0.0)
    drate = st * pm1 * leave * k.T[:, :, None]
    drate = np.where(net.is_three_body[None, :, None], drate * M[:, None, None], drate)
    out_idx = np.concatenate([net.reactant_idx, net.product_idx], axis=1)       # (S, 2T)
    out_sign = np.concatenate([-net.reactant_stoich, net.product_stoich], axis=1)
    keys = (out_idx[:, :, None] * (ni + 1) + net.reactant_idx[:, None, :]).reshape(-1)
    J = np.zeros((nz, ni, ni))
    for z in range(nz):
        contrib = (out_sign[:, :, None] * drate[z][:, None, :]).reshape(-1)
        flat = np.zeros((ni + 1) * (ni + 1))
        np.add.at(flat, keys, contrib)
        J[z] = flat.reshape(ni + 1, ni + 1)[:ni, :ni]
    return J


def chem_jac_fd(y, M, k, net, h=1e-7):
    """Central finite-difference Jacobian of chem_rhs (pin for chem_jac)."""
    nz, ni = y.shape
    J = np.zeros((nz, ni, ni))
    for j in range(ni):
        e = np.zeros_like(y)
        step = h * np.maximum(np.abs(y[:, j]), 1e-300)
        e[:, j] = step
        J[:, :, j] = (chem_rhs(y + e, M, k, net) - chem_rhs(y - e, M, k, net)) / (2 * step[:, None])
    return J


# ---------------------------------------------------------------------------
# diffusion (jax_step.py:67-246)
# ---------------------------------------------------------------------------

class DiffGrav(NamedTuple):
    A_grav_int: np.ndarray
    B_grav_int: np.ndarray
    C_grav_int: np.ndarray
    bdry0: np.ndarray        # (2, ni): A, B at layer 0
    bdry_top: np.ndarray     # (2, ni): A, C at layer nz-1
    dz_ave: np.ndarray


def compute_diff_grav(atm) -> DiffGrav:
    """y-independent molecular-diffusion pieces; jax_step.py:67-147."""
    dzi, Hpi, Ti, Tco, g = atm.dzi, atm.Hpi, atm.Ti, atm.Tco, atm.g
    ms, alpha, Dzz, vm, vs = atm.ms, atm.alpha, atm.Dzz, atm.vm, atm.vs
    nz = Tco.shape[0]
    j = np.arange(1, nz - 1)
    dz_ave = 0.5 * (dzi[j - 1] + dzi[j])
    use_vm = float(bool(atm.use_vm_mol))
    use_set = float(bool(atm.use_settling))
    c = lambda a: a[:, None]
    grav_j = -1.0 / c(Hpi[j]) + ms[None] * c(g[j]) / (NAVO * KB * c(Ti[j])) \
        + alpha[None] / c(Ti[j]) * c(Tco[j + 1] - Tco[j]) / c(dzi[j])
    grav_jm = -1.0 / c(Hpi[j - 1]) + ms[None] * c(g[j]) / (NAVO * KB * c(Ti[j - 1])) \
        + alpha[None] / c(Ti[j - 1]) * c(Tco[j] - Tco[j - 1]) / c(dzi[j - 1])
    grav_jp_b = -1.0 / c(Hpi[j]) + ms[None] * c(g[j + 1]) / (NAVO * KB * c(Ti[j])) \
        + alpha[None] / c(Ti[j]) * c(Tco[j + 1] - Tco[j]) / c(dzi[j])
    grav_jm_c = -1.0 / c(Hpi[j - 1]) + ms[None] * c(g[j - 1]) / (NAVO * KB * c(Ti[j - 1])) \
        + alpha[None] / c(Ti[j - 1]) * c(Tco[j] - Tco[j - 1]) / c(dzi[j - 1])
    inv2 = 1.0 / (2.0 * c(dz_ave))
    A_g = inv2 * (Dzz[j] * grav_j - Dzz[j - 1] * grav_jm)
    B_g = inv2 * Dzz[j] * grav_jp_b
    C_g = -inv2 * Dzz[j - 1] * grav_jm_c
    b0 = 1.0 / dzi[0] * Dzz[0] / 2.0 * (-1.0 / Hpi[0] + ms * g[0] / (NAVO * KB * Ti[0])
                                        + alpha / Ti[0] * (Tco[1] - Tco[0]) / dzi[0])
    bt = -1.0 / dzi[-1] * Dzz[-1] / 2.0 * (-1.0 / Hpi[-1] + ms * g[-1] / (NAVO * KB * Ti[-1])
                                          + alpha / Ti[-1] * (Tco[-1] - Tco[-2]) / dzi[-1])
    pos = lambda a: (a > 0) * a
    neg = lambda a: (a < 0) * a
    A_vm = -(pos(vm[j]) - neg(vm[j - 1])) / c(dz_ave)
    B_vm = -neg(vm[j]) / c(dz_ave)
    C_vm = pos(vm[j - 1]) / c(dz_ave)
    b0_vm, b0_vm_B = -pos(vm[0]) / dzi[0], -neg(vm[0]) / dzi[0]
    bt_vm, bt_vm_C = neg(vm[-1]) / dzi[-1], pos(vm[-1]) / dzi[-1]
    A_vs = -(pos(vs[j]) - neg(vs[j - 1])) / c(dz_ave)
    B_vs = -neg(vs[j]) / c(dz_ave)
    C_vs = pos(vs[j - 1]) / c(dz_ave)
    b0_vs, b0_vs_B = -pos(vs[0]) / dzi[0], -neg(vs[0]) / dzi[0]
    bt_vs, bt_vs_C = neg(vs[-1]) / dzi[-1], pos(vs[-1]) / dzi[-1]
    A_base = (1.0 - use_vm) * A_g + use_vm * A_vm
    B_base = (1.0 - use_vm) * B_g + use_vm * B_vm
    C_base = (1.0 - use_vm) * C_g + use_vm * C_vm
    b0A = (1.0 - use_vm) * b0 + use_vm * b0_vm + use_set * b0_vs
    b0B = (1.0 - use_vm) * b0 + use_vm * b0_vm_B + use_set * b0_vs_B
    btA = (1.0 - use_vm) * bt + use_vm * bt_vm + use_set * bt_vs
    btC = (1.0 - use_vm) * bt + use_vm * bt_vm_C + use_set * bt_vs_C
    return DiffGrav(A_base + use_set * A_vs, B_base + use_set * B_vs,
                    C_base + use_set * C_vs, np.stack([b0A, b0B]),
                    np.stack([btA, btC]), dz_ave)


def build_diff_coeffs(y, atm, grav: DiffGrav):
    """(A,B,C)_eddy (nz,), (A,B,C)_mol (nz,ni), ysum; jax_step.py:150-220."""
    Kzz, Dzz, dzi, vz = atm.Kzz, atm.Dzz, atm.dzi, atm.vz
    nz, ni = y.shape
    ysum = np.sum(np.where(atm.gas_indx_mask[None, :], y, 0.0), axis=1)
    j = np.arange(1, nz - 1)
    dz_ave = grav.dz_ave
    pos = lambda a: (a > 0) * a
    neg = lambda a: (a < 0) * a
    A_int = -1.0 / dz_ave * (Kzz[j] / dzi[j] * (ysum[j + 1] + ysum[j]) / 2.0
                             + Kzz[j - 1] / dzi[j - 1] * (ysum[j] + ysum[j - 1]) / 2.0) / ysum[j]
    B_int = 1.0 / dz_ave * Kzz[j] / dzi[j] * (ysum[j + 1] + ysum[j]) / 2.0 / ysum[j + 1]
    C_int = 1.0 / dz_ave * Kzz[j - 1] / dzi[j - 1] * (ysum[j] + ysum[j - 1]) / 2.0 / ysum[j - 1]
    A_int = A_int - (pos(vz[j]) - neg(vz[j - 1])) / dz_ave
    B_int = B_int - neg(vz[j]) / dz_ave
    C_int = C_int + pos(vz[j - 1]) / dz_ave
    A0 = -1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2.0 / ysum[0] - pos(vz[0]) / dzi[0]
    B0 = 1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2.0 / ysum[1] - neg(vz[0]) / dzi[0]
    At = -1.0 / dzi[-1] * (Kzz[-1] / dzi[-1]) * (ysum[-1] + ysum[-2]) / 2.0 / ysum[-1] + neg(vz[-1]) / dzi[-1]
    Ct = 1.0 / dzi[-1] * (Kzz[-1] / dzi[-1]) * (ysum[-1] + ysum[-2]) / 2.0 / ysum[-2] + pos(vz[-1]) / dzi[-1]
    A = np.concatenate([[A0], A_int, [At]])
    B = np.concatenate([[B0], B_int, [0.0]])
    C = np.concatenate([[0.0], C_int, [Ct]])
    c = lambda a: a[:, None]
    Ai_int = -1.0 / c(dz_ave) * (Dzz[j] / c(dzi[j]) * c(ysum[j + 1] + ysum[j]) / 2.0
                                 + Dzz[j - 1] / c(dzi[j - 1]) * c(ysum[j] + ysum[j - 1]) / 2.0) / c(ysum[j])
    Ai_int = Ai_int + grav.A_grav_int
    Bi_int = 1.0 / c(dz_ave) * Dzz[j] / c(dzi[j]) * c(ysum[j + 1] + ysum[j]) / 2.0 / c(ysum[j + 1]) + grav.B_grav_int
    Ci_int = 1.0 / c(dz_ave) * Dzz[j - 1] / c(dzi[j - 1]) * c(ysum[j] + ysum[j - 1]) / 2.0 / c(ysum[j - 1]) + grav.C_grav_int
    Ai0 = -1.0 / dzi[0] * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2.0 / ysum[0] + grav.bdry0[0]
    Bi0 = 1.0 / dzi[0] * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2.0 / ysum[1] + grav.bdry0[1]
    Ait = -1.0 / dzi[-1] * (Dzz[-1] / dzi[-1]) * (ysum[-1] + ysum[-2]) / 2.0 / ysum[-1] + grav.bdry_top[0]
    Cit = 1.0 / dzi[-1] * (Dzz[-1] / dzi[-1]) * (ysum[-1] + ysum[-2]) / 2.0 / ysum[-2] + grav.bdry_top[1]
    Ai = np.concatenate([Ai0[None], Ai_int, Ait[None]])
    Bi = np.concatenate([Bi0[None], Bi_int, np.zeros((1, ni))])
    Ci = np.concatenate([np.zeros((1, ni)), Ci_int, Cit[None]])
    return A, B, C, Ai, Bi, Ci, ysum


def apply_diffusion(y, A, B, C, Ai, Bi, Ci, atm):
    """Three-point stencil plus top/bottom flux BCs; jax_step.py:223-246."""
    At = A[:, None] + Ai
    Bt = B[:, None] + Bi
    Ct = C[:, None] + Ci
    d0 = At[0] * y[0] + Bt[0] * y[1]
    dt_ = At[-1] * y[-1] + Ct[-1] * y[-2]
    di = At[1:-1] * y[1:-1] + Bt[1:-1] * y[2:] + Ct[1:-1] * y[:-2]
    diff = np.concatenate([d0[None], di, dt_[None]])
    if atm.use_topflux:
        diff[-1] = diff[-1] + atm.top_flux / atm.dzi[-1]
    if atm.use_botflux:
        diff[0] = diff[0] + (atm.bot_flux - y[0] * atm.bot_vdep) / atm.dzi[0]
    return diff


def total_rhs(y, k, atm, net, grav=None):
    """f(y) = chem_rhs + diffusion; steady_state_grad.py:116-150."""
    if grav is None:
        grav = compute_diff_grav(atm)
    A, B, C, Ai, Bi, Ci, _ = build_diff_coeffs(y, atm, grav)
    return chem_rhs(y, atm.M, k, net) + apply_diffusion(y, A, B, C, Ai, Bi, Ci, atm)


# ---------------------------------------------------------------------------
# block-tridiagonal Thomas (solver.py)
# ---------------------------------------------------------------------------

class BTFactors(NamedTuple):
    lu: np.ndarray    # (nz, ni, ni)
    piv: np.ndarray   # (nz, ni)
    sup_d: np.ndarray
    sub_d: np.ndarray


def bt_factor(diag, sup_d, sub_d) -> BTFactors:
    """A'_0 = LU(D_0); A'_j = LU(D_j - (c_j b_{j-1}^T) o inv(A'_{j-1})); solver.py:33-64."""
    nz, ni, _ = diag.shape
    lu = np.zeros_like(diag)
    piv = np.zeros((nz, ni), dtype=np.int32)
    lu[0], piv[0] = sla.lu_factor(diag[0], check_finite=False)
    eye = np.eye(ni)
    for j in range(1, nz):
        inv = sla.lu_solve((lu[j - 1], piv[j - 1]), eye, check_finite=False)
        A_new = diag[j] - (sub_d[j - 1][:, None] * sup_d[j - 1][None, :]) * inv
        lu[j], piv[j] = sla.lu_factor(A_new, check_finite=False)
    return BTFactors(lu, piv, sup_d, sub_d)


def bt_solve(f: BTFactors, rhs):
    """Forward r' sweep, then back substitution; solver.py:67-112."""
    nz = rhs.shape[0]
    rmod = np.zeros_like(rhs)
    rmod[0] = rhs[0]
    for j in range(1, nz):
        inv_r = sla.lu_solve((f.lu[j - 1], f.piv[j - 1]), rmod[j - 1], check_finite=False)
        rmod[j] = rhs[j] - f.sub_d[j - 1] * inv_r
    x = np.zeros_like(rhs)
    x[-1] = sla.lu_solve((f.lu[-1], f.piv[-1]), rmod[-1], check_finite=False)
    for j in range(nz - 2, -1, -1):
        x[j] = sla.lu_solve((f.lu[j], f.piv[j]), rmod[j] - f.sup_d[j] * x[j + 1],
                            check_finite=False)
    return x


def block_thomas_diag_offdiag(diag, sup_d, sub_d, rhs):
    """Factor + solve; solver.py:115-126."""
    return bt_solve(bt_factor(diag, sup_d, sub_d), rhs)


def assemble_dense(diag, sup, sub):
    """Dense (nz*ni)^2 matrix; sup/sub may be (nz-1, ni) diagonals or dense blocks."""
    nz, ni, _ = diag.shape
    N = nz * ni
    Mx = np.zeros((N, N))
    for j in range(nz):
        Mx[j * ni:(j + 1) * ni, j * ni:(j + 1) * ni] = diag[j]
    for j in range(nz - 1):
        bs = np.diag(sup[j]) if sup.ndim == 2 else sup[j]
        bc = np.diag(sub[j]) if sub.ndim == 2 else sub[j]
        Mx[j * ni:(j + 1) * ni, (j + 1) * ni:(j + 2) * ni] = bs
        Mx[(j + 1) * ni:(j + 2) * ni, j * ni:(j + 1) * ni] = bc
    return Mx


def block_thomas_dense(diag, sup, sub, rhs):
    """Generic dense-off-block Thomas; solver.py:129-178."""
    nz = diag.shape[0]
    lus = [sla.lu_factor(diag[0], check_finite=False)]
    rmod = [rhs[0]]
    for j in range(1, nz):
        invB = sla.lu_solve(lus[-1], sup[j - 1], check_finite=False)
        invr = sla.lu_solve(lus[-1], rmod[-1], check_finite=False)
        lus.append(sla.lu_factor(diag[j] - sub[j - 1] @ invB, check_finite=False))
        rmod.append(rhs[j] - sub[j - 1] @ invr)
    x = [sla.lu_solve(lus[-1], rmod[-1], check_finite=False)]
    for j in range(nz - 2, -1, -1):
        x.append(sla.lu_solve(lus[j], rmod[j] - sup[j] @ x[-1], check_finite=False))
    return np.stack(x[::-1])


def banded_solve(diag, sup, sub, rhs):
    """The block-tridiagonal system (diagonal off-diagonal blocks) solved as
    one banded matrix by LAPACK gbsv: partial pivoting over the whole column
    rather than within a block.  A third elimination order next to the
    per-block LU of bt_factor and the C oracle, used to tell which cells of
    an ill-conditioned step FP64 determines (not a restatement of a
    reference function)."""
    nz, ni, _ = diag.shape
    N, bw = nz * ni, 2 * ni - 1
    ab = np.zeros((2 * bw + 1, N))          # A[i, j] at ab[bw + i - j, j]
    cols = np.arange(ni)
    for z in range(nz):
        r0 = z * ni
        for a in range(ni):
            ab[bw + a - cols, r0 + cols] = diag[z, a]
        if z + 1 < nz:
            i = r0 + cols
            ab[bw - ni, i + ni] = sup[z]       # A[i, i + ni]
            ab[bw + ni, i] = sub[z]            # A[i + ni, i]
    return sla.solve_banded((bw, bw), ab, rhs.reshape(-1), check_finite=False).reshape(nz, ni)


def _two_sum(a, b):
    """s + e = a + b exactly (Knuth's TwoSum)."""
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _two_prod(a, b):
    """p + e = a * b exactly (Dekker's product with Veltkamp's split)."""
    p = a * b
    def split(v):
        c = 134217729.0 * v           # 2^27 + 1
        h = c - (c - v)
        return h, v - h
    ah, al = split(a)
    bh, bl = split(b)
    return p, al * bl - (((p - ah * bh) - al * bh) - ah * bl)


def _residual_dd(diag, sup, sub, rhs, xh, xl):
    """rhs - A (xh + xl) as accurate as if formed in twice FP64's precision
    and rounded once: the compensated dot product Dot2 of Ogita, Rump and
    Oishi (SIAM J. Sci. Comput. 26(6), 2005, Algorithm 5.3), the terms of
    xl taken in plain FP64 (they are second order)."""
    s, c = rhs.copy(), np.zeros_like(rhs)
    def add(t_h, t_l):
        nonlocal s, c
        p, pe = _two_prod(t_h, t_l[0])
        s, se = _two_sum(s, -p)
        c = c + (se - pe) - t_h * t_l[1]
    ni = rhs.shape[1]
    for j in range(ni):
        add(diag[:, :, j], (np.broadcast_to(xh[:, j:j + 1], rhs.shape),
                            np.broadcast_to(xl[:, j:j + 1], rhs.shape)))
    if rhs.shape[0] > 1:
        z = np.zeros((1, ni))
        add(np.concatenate([sup, z]), (np.concatenate([xh[1:], z]), np.concatenate([xl[1:], z])))
        add(np.concatenate([z, sub]), (np.concatenate([z, xh[:-1]]), np.concatenate([z, xl[:-1]])))
    return s + c


def bt_solve_exact(diag, sup, sub, rhs, fac: BTFactors | None = None, max_iter: int = 40):
    """The block-tridiagonal system (diagonal off-diagonal blocks) solved to
    the exact solution of the FP64 system, rounded to FP64: the per-block LU
    solve of bt_solve, then iterative refinement with the solution carried
    as an unevaluated sum of two doubles and the residual formed with the
    compensated dot product (_residual_dd), until a correction no longer
    changes it.  Each step shrinks the error by ~cond * 2^-53, so this
    converges while cond(A) < ~1e15; the limiting accuracy is
    ~cond * 2^-106.  The yardstick for how close an FP64 solve (the
    reference's or another) comes to the exact solution; not a restatement
    of a reference function."""
    if fac is None:
        fac = bt_factor(diag, sup, sub)
    xh = bt_solve(fac, rhs)
    xl = np.zeros_like(xh)
    for _ in range(max_iter):
        d = bt_solve(fac, _residual_dd(diag, sup, sub, rhs, xh, xl))
        s, e = _two_sum(xh, d)
        xh, xl = _two_sum(s, e + xl)
        if not np.all(np.isfinite(xh)):
            raise FloatingPointError("bt_solve_exact: the refinement diverged")
        if np.all(np.abs(d) <= 1e-20 * np.abs(xh)):
            break
    return xh + xl


def chem_term_scale(y, M, k, net):
    """Per (layer, species): the sum of |terms| of chem_rhs (gross
    production + gross loss), the natural scale of its cancellation error."""
    nz, ni = y.shape
    rate, _, _ = _rates(y, M, k, net)
    S, T = net.reactant_idx.shape
    out = np.zeros((nz, ni))
    for z in range(nz):
        rr = np.repeat(np.abs(rate[z]), T)
        acc = np.zeros(ni + 1)
        np.add.at(acc, net.reactant_idx.reshape(-1), net.reactant_stoich.reshape(-1) * rr)
        np.add.at(acc, net.product_idx.reshape(-1), net.product_stoich.reshape(-1) * rr)
        out[z] = acc[:ni]
    return out


# ---------------------------------------------------------------------------
# Ros2 step (jax_step.py:249-315) and fixed-dt driver (integrate.py:20-36)
# ---------------------------------------------------------------------------

def ros2_lhs(y, k, dt, atm, net, grav=None, fix_mask=None):
    """(diag, sup_neg, sub_neg, rhs_y) of the Ros2 system; jax_step.py:257-298."""
    c0 = 1.0 / (GAMMA * dt)
    nz, ni = y.shape
    if grav is None:
        grav = compute_diff_grav(atm)
    A, B, C, Ai, Bi, Ci, _ = build_diff_coeffs(y, atm, grav)
    rhs_y = chem_rhs(y, atm.M, k, net) + apply_diffusion(y, A, B, C, Ai, Bi, Ci, atm)
    J = chem_jac(y, atm.M, k, net)
    diag_d = A[:, None] + Ai
    sup_d = B[:-1, None] + Bi[:-1]
    sub_d = C[1:, None] + Ci[1:]
    if atm.use_botflux:
        diag_d[0] = diag_d[0] + (-atm.bot_vdep / atm.dzi[0])
    di = np.arange(ni)
    diag = c0 * np.eye(ni)[None] - J
    diag[:, di, di] += -diag_d
    sup_neg, sub_neg = -sup_d, -sub_d
    if fix_mask is not None:
        diag = np.where(fix_mask[:, :, None], 0.0, diag)
        dd = diag[:, di, di]
        diag[:, di, di] = np.where(fix_mask, c0, dd)
        rhs_y = np.where(fix_mask, 0.0, rhs_y)
        sup_neg = np.where(fix_mask[:-1], 0.0, sup_neg)
        sub_neg = np.where(fix_mask[1:], 0.0, sub_neg)
    return diag, sup_neg, sub_neg, rhs_y


def ros2_step(y, k, dt, atm, net, fix_mask=None, return_stages=False):
    """One Ros2 step -> (sol, delta_arr); jax_step.py:249-315."""
    r = GAMMA
    grav = compute_diff_grav(atm)
    diag, sup_neg, sub_neg, rhs_y = ros2_lhs(y, k, dt, atm, net, grav, fix_mask)
    fac = bt_factor(diag, sup_neg, sub_neg)
    k1 = bt_solve(fac, rhs_y)
    yk2 = y + k1 / r
    rhs_yk2 = total_rhs(yk2, k, atm, net, grav)
    if fix_mask is not None:
        rhs_yk2 = np.where(fix_mask, 0.0, rhs_yk2)
    rhs2 = rhs_yk2 - (2.0 / (r * dt)) * k1
    k2 = bt_solve(fac, rhs2)
    sol = y + (3.0 / (2.0 * r)) * k1 + (1.0 / (2.0 * r)) * k2
    delta = np.abs(sol - yk2)
    if return_stages:
        return sol, delta, dict(k1=k1, k2=k2, yk2=yk2, rhs_y=rhs_y, rhs2=rhs2)
    return sol, delta


def ros2_step_banded(y, k, dt, atm, net):
    """ros2_step with both stage solves done by banded_solve (global partial
    pivoting) instead of the block Thomas factor; same formulas otherwise
    (jax_step.py:257-314)."""
    r = GAMMA
    grav = compute_diff_grav(atm)
    diag, sup_neg, sub_neg, rhs_y = ros2_lhs(y, k, dt, atm, net, grav)
    k1 = banded_solve(diag, sup_neg, sub_neg, rhs_y)
    yk2 = y + k1 / r
    rhs2 = total_rhs(yk2, k, atm, net, grav) - (2.0 / (r * dt)) * k1
    k2 = banded_solve(diag, sup_neg, sub_neg, rhs2)
    sol = y + (3.0 / (2.0 * r)) * k1 + (1.0 / (2.0 * r)) * k2
    return sol, np.abs(sol - yk2)


def ros2_step_exact(y, k, dt, atm, net):
    """ros2_step with both stage systems solved exactly (bt_solve_exact);
    the RHS evaluations stay FP64 as in ros2_step (jax_step.py:257-314).
    What the step would be with error-free linear solves: the yardstick for
    the solve error of FP64 steps on ill-conditioned columns, where the
    reference's own per-block LU is off from it by far more than FP64
    rounding."""
    r = GAMMA
    grav = compute_diff_grav(atm)
    diag, sup_neg, sub_neg, rhs_y = ros2_lhs(y, k, dt, atm, net, grav)
    fac = bt_factor(diag, sup_neg, sub_neg)
    k1 = bt_solve_exact(diag, sup_neg, sub_neg, rhs_y, fac)
    yk2 = y + k1 / r
    rhs2 = total_rhs(yk2, k, atm, net, grav) - (2.0 / (r * dt)) * k1
    k2 = bt_solve_exact(diag, sup_neg, sub_neg, rhs2, fac)
    sol = y + (3.0 / (2.0 * r)) * k1 + (1.0 / (2.0 * r)) * k2
    return sol, np.abs(sol - yk2)


def integrate_fixed_dt(y0, k, dt, n_steps, atm, net):
    """n fixed-dt steps; per-step max(delta/max(|sol|,1e-30)); integrate.py:20-36."""
    y = y0
    deltas = np.zeros(n_steps)
    for n in range(n_steps):
        sol, d = ros2_step(y, k, dt, atm, net)
        deltas[n] = np.max(d / np.maximum(np.abs(sol), 1e-30))
        y = sol
    return y, deltas


# ---------------------------------------------------------------------------
# step control (outer_loop.py)
# ---------------------------------------------------------------------------

class StepParams(NamedTuple):
    """Solver tolerances with the defaults of vulcan_cfg.py:125-152."""
    rtol: float = 0.2
    dt_var_min: float = 0.5
    dt_var_max: float = 2.0
    dt_min: float = 1e-14
    dt_max: float = 1e22 * 1e-5
    atol: float = 0.1
    mtol: float = 1e-22
    pos_cut: float = 0.0
    nega_cut: float = -1.0
    loss_eps: float = 0.1
    batch_max_retries: int = 8


def step_size(dt, delta, rtol, dt_var_min, dt_var_max, dt_min, dt_max):
    """outer_loop.py:181-193."""
    delta_eff = 0.01 * rtol if delta == 0.0 else delta
    h = 0.9 * (rtol / delta_eff) ** 0.5
    h = min(max(h, dt_var_min), dt_var_max)
    return min(max(dt * h, dt_min), dt_max)


def clip(y_in, ymix_old, p: StepParams, gas_mask=None):
    """outer_loop.py:196-224 (gas-only ysum when gas_mask is given)."""
    small = np.sum(np.where((y_in < p.pos_cut) & (y_in >= 0), np.abs(y_in), 0.0))
    nega = np.sum(np.where((y_in > p.nega_cut) & (y_in <= 0), np.abs(y_in), 0.0))
    yc = np.where((y_in < p.pos_cut) & (y_in >= p.nega_cut), 0.0, y_in)
    yc = np.where((ymix_old < p.mtol) & (yc < 0), 0.0, yc)
    if gas_mask is None:
        ysum = np.sum(yc, axis=1, keepdims=True)
    else:
        ysum = np.sum(np.where(gas_mask[None], yc, 0.0), axis=1, keepdims=True)
    return yc, yc / ysum, small, nega


def aggregate_delta(sol, delta_arr, ymix_old, mtol, atol, zero_bot_row=False):
    """outer_loop.py:227-250."""
    m = np.where(ymix_old < mtol, 0.0, delta_arr)
    m = np.where(sol < atol, 0.0, m)
    if zero_bot_row:
        m = m.copy()
        m[0] = 0.0
    ratio = np.where(sol > 0, m / np.maximum(np.abs(sol), 1e-300), 0.0)
    return float(np.max(ratio))


def atom_loss(y, compo, atom_ini):
    """outer_loop.py:171-178."""
    return (np.einsum("zi,ia->a", y, compo) - atom_ini) / atom_ini


def rel_change(y, y_prev, ymix, mtol, atol):
    """max |y - y_prev|/y over masked cells; outer_loop.py:2163-2184."""
    dy = np.abs(y - y_prev)
    dy = np.where(ymix < mtol, 0.0, dy)
    dy = np.where(y < atol, 0.0, dy)
    pos = y > 0
    return float(np.max(dy[pos] / y[pos])) if np.any(pos) else 0.0


def h2he_pin_step(pinned, mixes, t_old, ymix_old, h2_idx, he_idx, pin_time):
    """The H2/He pin state after one attempt (outer_loop.py:877-882): the
    snapshot trips on any attempt with t > pin_time, from the bottom layer's
    pre-step mixing ratios.  Returns (pinned, (h2_mix, he_mix))."""
    trip = (not pinned) and t_old > pin_time
    if trip:
        mixes = (float(ymix_old[0, h2_idx]), float(ymix_old[0, he_idx]))
    return pinned or trip, mixes


def balance_and_constrain(sol_clip, ymix_new, n0, gas_mask=None, charge=None, e_idx=None,
                          bottom_n=None, fix_sp_bot=None, h2he=None):
    """The accepted state: hydrostatic balance n0 * ymix_new of the gas
    species, the others kept clipped (outer_loop.py:847-853); then the ion
    charge balance e = -sum_i charge_i y_i (charge 0 at e_idx,
    outer_loop.py:855-860), the fixed bottom layer (:864-865), pinned
    bottom mixing ratios times n0[0] (:866-869) and, once pinned, the H2/He
    bottom mixing ratios (:884-892)."""
    y = n0[:, None] * ymix_new
    if gas_mask is not None:
        y = np.where(gas_mask[None, :], y, sol_clip)
    y = np.array(y, np.float64)
    if charge is not None:
        y[:, e_idx] = -(y * charge[None, :]).sum(axis=1)
    if bottom_n is not None:
        y[0] = bottom_n
    if fix_sp_bot:
        for i, m in fix_sp_bot.items():
            y[0, i] = m * n0[0]
    if h2he is not None:            # (h2_idx, he_idx, h2_mix, he_mix) once pinned
        h2, he, mh2, mhe = h2he
        y[0, h2] = mh2 * n0[0]
        y[0, he] = mhe * n0[0]
    return y


class ControlState(NamedTuple):
    y: np.ndarray
    y_prev: np.ndarray
    ymix: np.ndarray
    dt: float
    t: float
    atom_loss: np.ndarray
    atom_loss_prev: np.ndarray
    accept_count: int
    retry_count: int
    delta: float
    nega_count: int
    loss_count: int
    delta_count: int


def init_control(y, dt, compo):
    ymix = y / y.sum(axis=1, keepdims=True)
    ini = np.einsum("zi,ia->a", y, compo)
    al = atom_loss(y, compo, ini)
    return ControlState(y, y.copy(), ymix, float(dt), 0.0, al, al.copy(),
                        0, 0, 0.0, 0, 0, 0), ini


def control_update(s: ControlState, sol, delta_arr, n0, compo, atom_ini,
                   p: StepParams):
    """Accept / reject / force-accept and dt update of one attempt.

    Restates body_fn's decision logic (outer_loop.py:741-768, 847-861,
    900-917, 1031-1038) with photo, condensation, refresh, ion, fix-species
    and adaptive-rtol branches off (their defaults in vulcan_cfg.py).
    """
    sol_clip, ymix_new, _, _ = clip(sol, s.ymix, p)
    al_new = atom_loss(sol_clip, compo, atom_ini)
    delta = aggregate_delta(sol_clip, delta_arr, s.ymix, p.mtol, p.atol)
    all_nonneg = bool(np.all(sol_clip >= 0))
    loss_diff = float(np.max(np.abs(al_new - s.atom_loss_prev)))
    accept = all_nonneg and (loss_diff < p.loss_eps) and (delta <= p.rtol)
    next_dt_rej = s.dt * p.dt_var_min
    force = ((next_dt_rej < p.dt_min) or (s.retry_count >= p.batch_max_retries)) and not accept
    do_accept = accept or force
    reject = not do_accept
    d_inc = int(reject and delta > p.rtol)
    n_inc = int(reject and not delta > p.rtol and bool(np.any(sol_clip < 0)))
    l_inc = int(reject and not delta > p.rtol and not bool(np.any(sol_clip < 0)))
    dt_used = p.dt_min if force else s.dt
    t_next = s.t + dt_used if do_accept else s.t
    sol_bal = n0[:, None] * ymix_new
    ymix_bal = sol_bal / sol_bal.sum(axis=1, keepdims=True)
    al_bal = atom_loss(sol_bal, compo, atom_ini)
    y_prev_c = np.where(s.y_prev < 0, 0.0, s.y_prev)
    if force:
        y_next, ymix_next, al_next = y_prev_c, s.ymix, s.atom_loss_prev
    elif accept:
        y_next, ymix_next, al_next = sol_bal, ymix_bal, al_bal
    else:
        y_next, ymix_next, al_next = s.y_prev, s.ymix, s.atom_loss
    y_prev_next = y_next if do_accept else s.y_prev
    alp_next = al_next if do_accept else s.atom_loss_prev
    if force:
        dt_next = step_size(p.dt_min, delta, p.rtol, p.dt_var_min, p.dt_var_max, p.dt_min, p.dt_max)
    elif accept:
        dt_next = step_size(s.dt, delta, p.rtol, p.dt_var_min, p.dt_var_max, p.dt_min, p.dt_max)
    else:
        dt_next = next_dt_rej
    return ControlState(
        y_next, y_prev_next, ymix_next, dt_next, t_next, al_next, alp_next,
        s.accept_count + int(do_accept), 0 if do_accept else s.retry_count + 1,
        delta, s.nega_count + n_inc, s.loss_count + l_inc, s.delta_count + d_inc,
    ), dict(accept=accept, force=force, do_accept=do_accept)


# ---------------------------------------------------------------------------
# photolysis (photo.py)
# ---------------------------------------------------------------------------

def _t_abs(v, absp_T_idx, absp_T_cross):
    """sum over T-dependent absorbers s of v[z, idx_s] * cross_T[s, z, :]
    (photo.py:107-111, 156-160); zeros when there are none."""
    nz = v.shape[0]
    if absp_T_idx is None or np.asarray(absp_T_idx).size == 0:
        return None
    cT = np.asarray(absp_T_cross)                       # (n_T, nz, nbin)
    return np.sum(v[:, absp_T_idx][:, :, None] * np.transpose(cT, (1, 0, 2)), axis=1)


def compute_tau(y, dz, absp_idx, absp_cross, scat_idx, scat_cross, absp_T_idx=None,
                absp_T_cross=None):
    """tau (nz+1, nbin), top-down cumulative; photo.py:84-123, with the
    temperature-dependent absorbers (n_T, nz, nbin) of photo.py:107-111."""
    nz = y.shape[0]
    nbin = absp_cross.shape[1]
    per = np.zeros((nz, nbin))
    if absp_idx.size:
        per = per + (y[:, absp_idx] @ absp_cross) * dz[:, None]
    tT = _t_abs(y, absp_T_idx, absp_T_cross)
    if tT is not None:
        per = per + tT * dz[:, None]
    if scat_idx.size:
        per = per + (y[:, scat_idx] @ scat_cross) * dz[:, None]
    tau = np.zeros((nz + 1, nbin))
    tau[:nz] = np.flip(np.cumsum(np.flip(per, axis=0), axis=0), axis=0)
    return tau


def compute_flux(tau, sflux_top, ymix, absp_idx, absp_cross, scat_idx, scat_cross,
                 bins, mu_zenith, edd, ag0, hc, dflux_u_prev, absp_T_idx=None,
                 absp_T_cross=None):
    """Two-stream Eddington flux; photo.py:126-245 (both ag0 branches; the
    T-dependent absorbers enter w0, photo.py:156-160)."""
    mu_ang = -1.0 * mu_zenith
    nz = tau.shape[0] - 1
    nbin = tau.shape[1]
    dtau = tau[:-1] - tau[1:]
    tot_abs = ymix[:, absp_idx] @ absp_cross if absp_idx.size else np.zeros_like(dtau)
    tT = _t_abs(ymix, absp_T_idx, absp_T_cross)
    if tT is not None:
        tot_abs = tot_abs + tT
    tot_scat = ymix[:, scat_idx] @ scat_cross if scat_idx.size else np.zeros_like(dtau)
    with np.errstate(invalid="ignore", divide="ignore"):
        w0 = tot_scat / np.maximum(tot_abs + tot_scat, 1e-300)
    w0 = np.where(np.isnan(w0), 0.0, w0)
    w0 = np.minimum(w0, 1.0 - 1e-8)
    sflux = sflux_top[None, :] * np.exp(tau / mu_ang)
    dir_flux = sflux * (-mu_ang)
    if ag0 == 0.0:
        tran = np.exp(-1.0 / edd * (1.0 - w0) ** 0.5 * dtau)
        zeta_p = 0.5 * (1.0 + (1.0 - w0) ** 0.5)
        zeta_m = 0.5 * (1.0 - (1.0 - w0) ** 0.5)
        ll = -1.0 * w0 / (1.0 / mu_ang ** 2 - 1.0 / edd ** 2 * (1.0 - w0))
        g_p = 0.5 * (ll * (1.0 / edd + 1.0 / mu_ang))
        g_m = 0.5 * (ll * (1.0 / edd - 1.0 / mu_ang))
    else:  # photo.py:183-197
        wg = 1.0 - w0 * ag0
        tran = np.exp(-1.0 / edd * ((1.0 - w0 * ag0) * (1.0 - w0)) ** 0.5 * dtau)
        zeta_p = 0.5 * (1.0 + ((1.0 - w0) / wg) ** 0.5)
        zeta_m = 0.5 * (1.0 - ((1.0 - w0) / wg) ** 0.5)
        ll = ((1.0 - w0) * wg - 1.0) / (1.0 / mu_ang ** 2 - 1.0 / edd ** 2 * (1.0 - w0) * wg)
        g_p = 0.5 * (ll * (1.0 / edd + 1.0 / (mu_ang * wg)) + w0 * ag0 * mu_ang / wg)
        g_m = 0.5 * (ll * (1.0 / edd - 1.0 / (mu_ang * wg)) - w0 * ag0 * mu_ang / wg)
    ll = np.clip(ll, -1e10, 1e10)  # clipped after g_p/g_m, as photo.py:199
    chi = zeta_m ** 2 * tran ** 2 - zeta_p ** 2
    xi = zeta_p * zeta_m * (1.0 - tran ** 2)
    phi = (zeta_m ** 2 - zeta_p ** 2) * tran
    i_u = phi * g_p * dir_flux[:-1] - (xi * g_m + chi * g_p) * dir_flux[1:]
    i_d = phi * g_m * dir_flux[1:] - (chi * g_m + xi * g_p) * dir_flux[:-1]
    dflux_d = np.zeros((nz + 1, nbin))
    for j in range(nz - 1, -1, -1):
        dflux_d[j] = (1.0 / chi[j]) * (phi[j] * dflux_d[j + 1] - xi[j] * dflux_u_prev[j]
                                       + i_d[j] / mu_ang)
    dflux_u = np.zeros((nz + 1, nbin))
    for j in range(1, nz + 1):
        dflux_u[j] = (1.0 / chi[j - 1]) * (phi[j - 1] * dflux_u[j - 1] - xi[j - 1] * dflux_d[j]
                                           + i_u[j - 1] / mu_ang)
    ave_dir = 0.5 * (sflux[:-1] + sflux[1:])
    tot = ave_dir + 0.5 * (dflux_u[:-1] + dflux_u[1:] + dflux_d[1:] + dflux_d[:-1]) / edd
    aflux = tot / (hc / bins[None, :])
    return aflux, sflux, dflux_d, dflux_u


def compute_J(aflux, cross_J, din12, dbin1, dbin2):
    """Trapezoid J per branch, split at din12; photo.py:309-351."""
    f1, f2 = aflux[:, :din12], aflux[:, din12:]
    c1, c2 = cross_J[:, :din12], cross_J[:, din12:]
    J1 = (c1 @ f1.T) * dbin1 - 0.5 * (f1[:, 0][None] * c1[:, 0][:, None]
                                      + f1[:, -1][None] * c1[:, -1][:, None]) * dbin1
    J2 = (c2 @ f2.T) * dbin2 - 0.5 * (f2[:, 0][None] * c2[:, 0][:, None]
                                      + f2[:, -1][None] * c2[:, -1][:, None]) * dbin2
    return J1 + J2


def compute_J_T(aflux, cross_J_T, din12, dbin1, dbin2):
    """J of the temperature-dependent branches (n_brT, nz): per layer z the
    same split trapezoid against cross_J_T[:, z, :]; photo.py:334-347."""
    cT = np.asarray(cross_J_T)
    f1, f2 = aflux[:, :din12], aflux[:, din12:]
    c1, c2 = cT[:, :, :din12], cT[:, :, din12:]
    J1 = np.sum(f1[None] * c1, axis=2) * dbin1 - 0.5 * (
        f1[:, 0][None] * c1[:, :, 0] + f1[:, -1][None] * c1[:, :, -1]) * dbin1
    J2 = np.sum(f2[None] * c2, axis=2) * dbin2 - 0.5 * (
        f2[:, 0][None] * c2[:, :, 0] + f2[:, -1][None] * c2[:, :, -1]) * dbin2
    return J1 + J2


def update_k_with_J(k, J, re_idx, active, f_diurnal):
    """k[re_idx] = J * f_diurnal for active branches; photo.py:455-478."""
    k = k.copy()
    for b in range(J.shape[0]):
        if active[b]:
            k[re_idx[b]] = J[b] * f_diurnal
    return k


# ---------------------------------------------------------------------------
# steady-state adjoint (steady_state_grad.py)
# ---------------------------------------------------------------------------

def jacobian_blocks(y, k, atm, net):
    """J_y blocks without c0; steady_state_grad.py:153-175."""
    grav = compute_diff_grav(atm)
    A, B, C, Ai, Bi, Ci, _ = build_diff_coeffs(y, atm, grav)
    J = chem_jac(y, atm.M, k, net)
    ni = y.shape[1]
    di = np.arange(ni)
    diag = J.copy()
    diag[:, di, di] += A[:, None] + Ai
    if atm.use_botflux:
        diag[0, di, di] += -atm.bot_vdep / atm.dzi[0]
    return diag, B[:-1, None] + Bi[:-1], C[1:, None] + Ci[1:]


def validate_steady_state(y_star, k, atm, net, residual_rtol=1e-6, residual_atol=0.0):
    """Residual gate; steady_state_grad.py:178-196."""
    res = total_rhs(y_star, k, atm, net)
    r = float(np.max(np.abs(res)))
    tol = max(residual_atol, residual_rtol * max(float(np.max(np.abs(y_star))), 1.0))
    if r > tol:
        raise ValueError(f"steady-state residual {r:.3e} exceeds {tol:.3e}")
    return r


def adjoint_lambda(y_star, k, atm, net, v):
    """lambda = J_y^{-T} v via the transposed block Thomas; steady_state_grad.py:314-317."""
    diag, sup_d, sub_d = jacobian_blocks(y_star, k, atm, net)
    return block_thomas_diag_offdiag(np.transpose(diag, (0, 2, 1)), sub_d, sup_d, v)


def chem_vjp_k(y, M, k, net, lam):
    """(d chem_rhs / d k)^T lam, shape (S, nz); steady_state_grad.py:319-323."""
    nz, ni = y.shape
    S, T = net.reactant_idx.shape
    rate, _, _ = _rates(y, M, np.ones_like(k), net)     # rate per unit k
    lp = np.concatenate([lam, np.zeros((nz, 1))], axis=1)
    g = np.zeros((nz, S))
    for s in range(T):
        g = g + net.product_stoich[None, :, s] * lp[:, net.product_idx[:, s]]
        g = g - net.reactant_stoich[None, :, s] * lp[:, net.reactant_idx[:, s]]
    return (g * rate).T


def adjoint_grad_k(y_star, k, atm, net, v):
    """d L/d k = -(df/dk)^T J_y^{-T} v; steady_state_grad.py:306-327."""
    lam = adjoint_lambda(y_star, k, atm, net, v)
    return -chem_vjp_k(y_star, atm.M, k, net, lam), lam


def newton_steady_state(y0, k, atm, net, tol=1e-9, max_iter=100):
    """Newton on f(y)=0 with the same Jacobian blocks (test helper)."""
    y = y0.copy()
    for it in range(max_iter):
        res = total_rhs(y, k, atm, net)
        nrm = float(np.max(np.abs(res)))
        if nrm < tol:
            return y, it, nrm
        diag, sup_d, sub_d = jacobian_blocks(y, k, atm, net)
        y = y + block_thomas_diag_offdiag(diag, sup_d, sub_d, -res)
    raise RuntimeError(f"Newton did not converge: {nrm:.3e}")


# ---------------------------------------------------------------------------
# steady-state driver (outer_loop.py body_fn / cond_fn subset)
# ---------------------------------------------------------------------------

class RunParams(NamedTuple):
    """Stopping rule of the adaptive run: accepted-step relative change below
    conv_tol once ready (t > trun_min and count > count_min, outer_loop.py:
    706-711) and, with photolysis, aflux_change < flux_cri (:688-689); or
    count_max accepted steps (outer_loop.py:673-675)."""
    conv_tol: float = 1e-4
    count_min: int = 120
    trun_min: float = 1e2
    count_max: int = 5000
    photo_frq: int = 5
    flux_cri: float = 0.1
    flux_atol: float = 1.0
    photo_frq_final: int = 0          # 0: no ini -> final switch
    photo_sw_longdy: float = 1.0      # yconv_min * 10 (outer_loop.py:1327-1330)
    photo_sw_longdydt: float = 1e-6


class PhotoState(NamedTuple):
    ph: object          # synth.PhotoSetup
    dz: np.ndarray
    dflux_u: np.ndarray
    aflux: np.ndarray
    aflux_change: float = 0.0
    tcross: dict | None = None   # the column's absp_T_cross / cross_J_T


def photo_state(ph, dz, nz, tcross=None):
    nbin = ph.bins.shape[0]
    return PhotoState(ph, dz, np.zeros((nz + 1, nbin)), np.zeros((nz, nbin)), 0.0, tcross)


def photo_update(y, ymix, k, pst: PhotoState, flux_atol=1.0):
    """tau -> flux -> J -> k rows, aflux_change (outer_loop.py:253-352)."""
    ph = pst.ph
    tc = pst.tcross or {}
    tidx = np.asarray(getattr(ph, "absp_T_idx", np.zeros(0, np.int64)), dtype=np.int64)
    tkw = dict(absp_T_idx=tidx, absp_T_cross=tc.get("absp_T_cross")) if tidx.size else {}
    tau = compute_tau(y, pst.dz, ph.absp_idx, ph.absp_cross, ph.scat_idx, ph.scat_cross, **tkw)
    af, sf, dd, du = compute_flux(tau, ph.sflux_top, ymix, ph.absp_idx, ph.absp_cross,
                                  ph.scat_idx, ph.scat_cross, ph.bins, ph.mu_zenith, ph.edd,
                                  ph.ag0, ph.hc, pst.dflux_u, **tkw)
    J = compute_J(af, ph.cross_J, ph.din12_indx, ph.dbin1, ph.dbin2)
    k = update_k_with_J(k, J, ph.branch_re_idx, ph.branch_active, ph.f_diurnal)
    bT = np.asarray(getattr(ph, "branch_T_re_idx", np.zeros(0, np.int64)), dtype=np.int64)
    if bT.size:
        JT = compute_J_T(af, tc["cross_J_T"], ph.din12_indx, ph.dbin1, ph.dbin2)
        k = update_k_with_J(k, JT, bT, np.asarray(ph.branch_T_active, bool), ph.f_diurnal)
    mask = af > flux_atol
    ratio = np.where(mask, np.abs(af - pst.aflux) / np.maximum(np.abs(af), 1e-300), 0.0)
    change = float(np.max(ratio)) if np.any(mask) else 0.0
    return k, pst._replace(dflux_u=du, aflux=af, aflux_change=change), dict(tau=tau, aflux=af, J=J)


def steady_state_run(y0, k, atm, net, dt0, p: StepParams = StepParams(),
                     rp: RunParams = RunParams(), compo=None, photo: PhotoState | None = None,
                     max_attempts=None, trace=None):
    """Adaptive Ros2 run of one column until the stopping rule; returns
    (y, info).  Photolysis, when given, refreshes k before an attempt with
    retry_count == 0 and accept_count % photo_frq == 0 (outer_loop.py:716-724)."""
    nz, ni = y0.shape
    if compo is None:
        compo = np.ones((ni, 1))
    s, ini = init_control(y0, dt0, compo)
    n0 = atm.M
    attempts = 0
    rel = np.inf
    status = "max_steps"
    frq = rp.photo_frq
    rel_carried = 1.0          # longdy's initial value (state.py:706)
    n_photo = 0
    while True:
        if photo is not None and s.retry_count == 0 and s.accept_count % frq == 0:
            k, photo, _ = photo_update(s.y, s.ymix, k, photo, rp.flux_atol)
            n_photo += 1
        sol, delta = ros2_step(s.y, k, s.dt, atm, net)
        s_new, d = control_update(s, sol, delta, n0, compo, ini, p)
        attempts += 1
        if d["do_accept"]:
            rel = rel_change(s_new.y, s.y_prev, s_new.ymix, p.mtol, p.atol)
            rel_carried = rel
        # ini -> final photolysis cadence (outer_loop.py:1040-1049) on the
        # carried relative change standing in for longdy, over dt for longdydt
        if (photo is not None and rp.photo_frq_final > 0 and frq != rp.photo_frq_final
                and rel_carried < rp.photo_sw_longdy
                and rel_carried / s.dt < rp.photo_sw_longdydt):
            frq = rp.photo_frq_final
        if trace is not None:
            trace.append(dict(dt=s.dt, accept=d["do_accept"], delta=s_new.delta, rel=rel))
        s = s_new
        ready = (s.t > rp.trun_min) and (s.accept_count > rp.count_min)
        flux_ok = photo is None or photo.aflux_change < rp.flux_cri
        if d["do_accept"] and ready and rel < rp.conv_tol and flux_ok:
            status = "converged"
            break
        if s.accept_count >= rp.count_max:
            break
        if max_attempts is not None and attempts >= max_attempts:
            status = "max_attempts"
            break
    return s.y, dict(state=s, attempts=attempts, rel=rel, status=status, k=k,
                     photo_updates=n_photo, photo_frq=frq)
