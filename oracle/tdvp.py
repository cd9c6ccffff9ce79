"""Two-site TDVP and the Krylov exponential (oracle).  TEST INFRASTRUCTURE ONLY.

Follows algs/time.jl:151-166 (twositeOps), :169-178 (joinedtwosite_update),
:182-202 (krylovTimeEvol), :288-362 (tdvp_twosite), with the reference's own
step conventions: step = prefactor/(Ns-1)/2 per bond, the single-site update
of the next centre runs forward with the same step (time.jl:333,346), and
one call is one half-sweep.

TensorPACK's lanczos (called at time.jl:184) is not vendored; the restated
Krylov process is: v_1 = x/|x|, w = H v_j, alpha_j = <v_j, w>, two passes of
classical Gram-Schmidt of w against all previous vectors, beta_j = |w|,
stop when beta_j <= 1e-10; exp(prefactor T) e_1 of the tridiagonal
T = tridiag(beta, alpha, beta) gives the coefficients (time.jl:186-197).
"""
import numpy as np

from . import dmrg as od


def krylov_expm(apply, x, r, pf):
    """Returns (exp(pf*H) x / |x| in the r-dimensional Krylov space, lowest Ritz value)."""
    v = [x / np.linalg.norm(x)]
    alpha, beta = [], []
    for j in range(r):
        w = apply(v[j])
        alpha.append(float(np.vdot(v[j], w).real))
        if j == r - 1:
            break
        for i in range(j + 1):
            for _ in range(2):
                w = w - np.vdot(v[i], w) * v[i]
        b = float(np.linalg.norm(w))
        if b <= 1e-10:
            break
        beta.append(b)
        v.append(w / b)
    k = len(v)
    T = np.diag(alpha[:k]) + np.diag(beta[:k - 1], 1) + np.diag(beta[:k - 1], -1)
    lam, Q = np.linalg.eigh(T)
    c = Q @ (np.exp(pf * lam) * Q[0, :])
    out = sum(cj * vj for cj, vj in zip(c, v))
    return out, float(lam.min())


def tdvp_twosite(psi, mpo, m=0, minm=2, cutoff=0.0, r=2, prefactor=1.0, Lenv=None, Renv=None):
    """One half-sweep of two-site TDVP (algs/time.jl:288-362).  Modifies psi
    (and the environments) in place; returns the last energy."""
    Ns = len(psi)
    timestep = prefactor / (Ns - 1) / 2
    if Lenv is None:
        Lenv, Renv = od.makeEnv(psi, mpo)
    j = 1 if psi.oc == 1 else (-1 if psi.oc == Ns else 1)
    En = 0.0
    for _ in range(Ns - 1):
        i = psi.oc
        iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
        AA = np.tensordot(psi[iL], psi[iR], axes=(2, 0))
        WL, WR = mpo[iL - 1], mpo[iR - 1]

        def h2(t):
            return od.heff2(Lenv[iL], Renv[iR], t, WL, WR)
        newpsi, En = krylov_expm(h2, AA, r, timestep)
        newpsi = newpsi / np.linalg.norm(newpsi)
        if psi.qn:
            U, s, V, _, qnew = od.fsvd(newpsi, m, minm, cutoff, psi.qbond[iL - 1], psi.qphys, psi.qbond[iR])
            psi.qbond[iL] = qnew
        else:
            U, s, V, _, _ = od.fsvd(newpsi, m, minm, cutoff)
        if j > 0:
            psi[iL] = U
            DV = s[:, None, None] * V
            Lenv[iR] = od.Lupdate(Lenv[iL], psi[iL], WL)
            DV = DV / np.linalg.norm(DV)

            def h1(t):
                return od.singlesite_update(Lenv[iR], Renv[iR], t, WR)
            psi[iR], En = krylov_expm(h1, DV, r, timestep)
        else:
            psi[iR] = V
            UD = U * s[None, None, :]
            UD = UD / np.linalg.norm(UD)
            Renv[iL] = od.Rupdate(Renv[iR], psi[iR], WR)

            def h1(t):
                return od.singlesite_update(Lenv[iL], Renv[iL], t, WL)
            psi[iL], En = krylov_expm(h1, UD, r, timestep)
        psi.oc += j
        psi[psi.oc] = psi[psi.oc] / np.linalg.norm(psi[psi.oc])
        if psi.oc == Ns and j > 0:
            j = -1
        elif psi.oc == 1 and j < 0:
            j = 1
    return En, Lenv, Renv


def local_expect(psi, op, site):
    """<psi|O_site|psi> / <psi|psi> by a full contraction (any gauge)."""
    E = np.ones((1, 1))
    for i in range(1, len(psi) + 1):
        A = psi[i]
        if i == site:
            A2 = np.einsum('kj,ajb->akb', op, A)
        else:
            A2 = A
        E = np.einsum('xa,xsB,asb->Bb', E, A.conj(), A2)
    N = np.ones((1, 1))
    for i in range(1, len(psi) + 1):
        A = psi[i]
        N = np.einsum('xa,xsB,asb->Bb', N, A.conj(), A)
    return float((E[0, 0] / N[0, 0]).real)
