"""CPU: the cyclic-reduction factor's level layout and elimination rules.

``cr_level`` restates the offsets the kernels compute (cr_level in
csrc/vk_bt.hip).  The layout test checks that the regions of every level
are disjoint and inside the 7 nz N^2 doubles per column the workspace
reserves, for nz 1..399 and N 16..112.  ``cr_solve`` follows the kernels'
index rules (P_j = U_{j-1} inv(D_j), Q_j = L_{j+1} inv(D_j), the even-row
update, the top solve and the back substitution) with dense NumPy blocks,
and must reproduce a dense solve of random block-tridiagonal systems
(nz 1..60: odd and even, powers of two and one above) to 1e-10.
"""
from __future__ import annotations

import numpy as np
import pytest


def cr_level(nz, blk, N, level):
    """Offsets (doubles) of one level; past the top, the top level."""
    n, off, l = nz, 0, 0
    while True:
        v = dict(n=n, h=n // 2, D=-1, L=-1, U=-1, P=-1, Q=-1)
        if l >= 1:
            v["D"] = off
            off += n * blk
            if n > 1:
                v["L"] = off
                off += n * blk
                v["U"] = off
                off += n * blk
        v["Dinv"] = off
        if n > 1:
            off += v["h"] * blk
            v["P"] = off
            off += v["h"] * blk
            v["Q"] = off
            off += v["h"] * blk
        else:
            off += blk
        v["r"] = off
        off += n * N
        if l == level or n == 1:
            return v
        n = (n + 1) // 2
        l += 1


def top_level(nz):
    l, n = 0, nz
    while n > 1:
        n = (n + 1) // 2
        l += 1
    return l


def spans(nz, N):
    blk = N * N
    out = []
    for l in range(top_level(nz) + 1):
        v = cr_level(nz, blk, N, l)
        n, h = v["n"], v["h"]
        if l >= 1:
            out.append((v["D"], v["D"] + n * blk))
            if n > 1:
                out += [(v["L"], v["L"] + n * blk), (v["U"], v["U"] + n * blk)]
        if n > 1:
            out += [(v["Dinv"], v["Dinv"] + h * blk), (v["P"], v["P"] + h * blk),
                    (v["Q"], v["Q"] + h * blk)]
        else:
            out.append((v["Dinv"], v["Dinv"] + blk))
        out.append((v["r"], v["r"] + n * N))
    return sorted(out)


@pytest.mark.parametrize("N", [16, 32, 96, 112])
def test_levels_are_disjoint_and_fit_the_workspace(N):
    worst = 0.0
    for nz in range(1, 400):
        sp = spans(nz, N)
        for a, b in zip(sp, sp[1:]):
            assert a[1] <= b[0], (nz, a, b)
        assert sp[0][0] >= 0 and sp[-1][1] <= 7 * nz * N * N, nz
        worst = max(worst, sp[-1][1] / (nz * N * N))
    assert worst <= 6.44


def cr_solve(D0, sub, sup, r0):
    """Cyclic reduction with the kernels' index rules.  D0 (nz, n, n);
    L_j = diag(sub[j-1]), U_j = diag(sup[j])."""
    nz = D0.shape[0]
    levels = []
    D = [D0[j] for j in range(nz)]
    L = [None] + [np.diag(sub[j - 1]) for j in range(1, nz)]
    U = [np.diag(sup[j]) for j in range(nz - 1)] + [None]
    while len(D) > 1:
        m = len(D)
        Dinv, P, Q = [], [], []
        for w in range(m // 2):
            j = 2 * w + 1
            Di = np.linalg.inv(D[j])
            Dinv.append(Di)
            P.append(U[j - 1] @ Di)
            Q.append(L[j + 1] @ Di if j + 1 < m else None)
        Dn, Ln, Un = [], [], []
        for w in range((m + 1) // 2):
            i = 2 * w
            d, ln, un = D[i].copy(), None, None
            if i >= 1:
                d = d - Q[w - 1] @ U[i - 1]
                if i >= 2:
                    ln = -(Q[w - 1] @ L[i - 1])
            if i + 1 < m:
                d = d - P[w] @ L[i + 1]
                if i + 2 < m:
                    un = -(P[w] @ U[i + 1])
            Dn.append(d)
            Ln.append(ln)
            Un.append(un)
        levels.append(dict(L=L, U=U, Dinv=Dinv, P=P, Q=Q))
        D, L, U = Dn, Ln, Un
    top = np.linalg.inv(D[0])
    rs = [list(r0)]
    for lv in levels:
        r, rn = rs[-1], []
        for w in range((len(r) + 1) // 2):
            i = 2 * w
            o = r[i].copy()
            if i >= 1:
                o = o - lv["Q"][w - 1] @ r[i - 1]
            if i + 1 < len(r):
                o = o - lv["P"][w] @ r[i + 1]
            rn.append(o)
        rs.append(rn)
    x = [top @ rs[-1][0]]
    for l in range(len(levels) - 1, -1, -1):
        lv, r = levels[l], rs[l]
        xl = []
        for j in range(len(r)):
            if j % 2 == 0:
                xl.append(x[j // 2])
            else:
                t = r[j] - lv["L"][j] @ x[(j - 1) // 2]
                if j + 1 < len(r):
                    t = t - lv["U"][j] @ x[(j + 1) // 2]
                xl.append(lv["Dinv"][(j - 1) // 2] @ t)
        x = xl
    return np.array(x)


@pytest.mark.parametrize("nz", [1, 2, 3, 4, 5, 7, 8, 16, 17, 33, 60])
def test_reduction_rules_solve_the_system(nz):
    rng = np.random.default_rng(nz)
    n = 5
    D0 = rng.normal(size=(nz, n, n)) + 8 * np.eye(n)
    sub = rng.normal(size=(max(nz - 1, 0), n))
    sup = rng.normal(size=(max(nz - 1, 0), n))
    r0 = rng.normal(size=(nz, n))
    A = np.zeros((nz * n, nz * n))
    for j in range(nz):
        A[j * n:(j + 1) * n, j * n:(j + 1) * n] = D0[j]
        if j >= 1:
            A[j * n:(j + 1) * n, (j - 1) * n:j * n] = np.diag(sub[j - 1])
        if j + 1 < nz:
            A[j * n:(j + 1) * n, (j + 1) * n:(j + 2) * n] = np.diag(sup[j])
    xr = np.linalg.solve(A, r0.reshape(-1)).reshape(nz, n)
    x = cr_solve(D0, sub, sup, r0)
    assert np.max(np.abs(x - xr)) < 1e-10 * np.max(np.abs(xr))
