"""Single-site DMRG with subspace expansion, "3S" (oracle).  TEST INFRASTRUCTURE ONLY.

Function by function:
  lanczos_ritz   TensorPACK lanczos(psi, ops, r, m=1, updatefct=singlesite_update)
                 as called at algs/DMRG.jl:328 (not vendored; the restated
                 Krylov process is the one of oracle/tdvp.py: v_1 = x/|x|,
                 two classical Gram-Schmidt passes, stop at beta <= 1e-10;
                 the result is the lowest Ritz pair of the r x r tridiagonal,
                 its vector signed so that its overlap with x is >= 0)
  Lexpand        algs/DMRG.jl:273-278  [A | alpha P] along the right bond,
                 P[x, k, (u, b)] = sum L[x,w,a] A[a,j,b] W[w,j,k,u], u fastest
  Rexpand        algs/DMRG.jl:280-286  [A ; alpha P] along the left bond,
                 P[(a, w), k, B] = sum W[w,j,k,u] A[a,j,b] R[b,u,B], a fastest
  move_right     lib/MPS/move.jl:175-190 (moveR! with condition=true: D V
                 restricted to the first size(Rpsi, 1) columns)
  move_left      lib/MPS/move.jl:124-139 (moveL! with condition=true)
  alpha_update   algs/DMRG.jl:288-303
  step3S         algs/DMRG.jl:322-360
  dmrg3S         algs/DMRG.jl:655-671 -> optmps algs/optimizeMPS.jl:759-849
                 with nsites = 1 (range 1: the same bond pairs as two-site)

Charges of the enlarged bond: a new column (u, b) carries q_b + delta_u (right
link of the site's MPO), a new row (a, w) carries q_a + delta_w (left link);
this is the block rule the reference's joinindex! applies to its quantum-number
tensors.  The enlarged tensor is scaled to unit norm before the truncated SVD,
which then follows the documented fsvd rule of oracle/dmrg.py (cutoff on the
discarded weight, at most m and at least minm values, kept values renormalised);
the reference's fast QR branch at cutoff 0 keeps every vector, as that rule
does at cutoff 0, and differs from it only by a gauge of the new bond.

Pinned in tests/test_oracle_3s.py: the expansion contracted with the other
environment gives H_1site A (an identity independent of the restatement), the
step with no truncation leaves the two-site state unchanged (expansion changes
the basis only), the Ritz pair against a brute-force Rayleigh-Ritz on
span{x, Hx} and, for a full Krylov space, against the dense eigensolver, the
branches of alpha_update, and whole runs against exact diagonalisation and
between dense and quantum-number runs (test/dmrg_test.jl:7,82-85).
"""
import numpy as np

from . import dmrg as od
from . import mps as omps

NOISE = 1.0        # algvars defaults, algs/optimizeMPS.jl:528-532
NOISE_GOAL = 0.3
NOISE_INCR = 0.01
ALPHA_MAX = 1.0    # algs/DMRG.jl:287


def lanczos_ritz(apply, x, r):
    """Lowest Ritz pair of the r-dimensional Krylov space of x."""
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
    c = Q[:, 0] * (1.0 if Q[0, 0] >= 0 else -1.0)
    return float(lam[0]), sum(cj * vj for cj, vj in zip(c, v))


def Lexpand(A, W, L, alpha):
    """[x, k, (u, b)] expansion block (u fastest), already times alpha."""
    T = np.tensordot(L, A, axes=(2, 0))                    # [x, w, j, b]
    T = np.tensordot(T, W, axes=([1, 2], [0, 1]))          # [x, b, k, u]
    T = T.transpose(0, 2, 3, 1)                            # [x, k, u, b]
    x, k, u, b = T.shape
    return alpha * T.reshape(x, k, u * b, order="F")


def Rexpand(A, W, R, alpha):
    """[(a, w), k, B] expansion block (a fastest), already times alpha."""
    T = np.tensordot(A, R, axes=(2, 0))                    # [a, j, u, B]
    T = np.tensordot(W, T, axes=([1, 3], [1, 2]))          # [w, k, a, B]
    T = T.transpose(2, 0, 1, 3)                            # [a, w, k, B]
    a, w, k, B = T.shape
    return alpha * T.reshape(a * w, k, B, order="F")


def expand_charges_right(qb, delta):
    """Charges of [b ; (u, b)] with u fastest: q_b + delta_u."""
    new = (qb[:, None, :] + delta[None, :, :]).reshape(-1, qb.shape[1])   # b outer, u inner
    return np.concatenate([qb, new], axis=0)


def expand_charges_left(qa, delta):
    """Charges of [a ; (a, w)] with a fastest: q_a + delta_w."""
    new = (qa[None, :, :] + delta[:, None, :]).reshape(-1, qa.shape[1])   # w outer, a inner
    return np.concatenate([qa, new], axis=0)


def move_right(Aexp, Rpsi, m, minm, cutoff, qL=None, qP=None, qRexp=None):
    """moveR! with condition: returns (Ltens, Rtens, s, truncerr, qnew)."""
    Aexp = Aexp / np.linalg.norm(Aexp)
    a, d, bx = Aexp.shape
    z = None if qP is None else np.zeros((1, qP.shape[1]), dtype=np.int64)
    U, s, V, truncerr, qnew = od.fsvd(Aexp.reshape(a, d, 1, bx), m, minm, cutoff, qL, qP, qRexp, qP2=z)
    nb = Rpsi.shape[0]
    DV = s[:, None] * V[:, 0, :nb]
    return U, np.tensordot(DV, Rpsi, axes=(1, 0)), s, truncerr, qnew


def move_left(Lpsi, Aexp, m, minm, cutoff, qLexp=None, qP=None, qR=None):
    """moveL! with condition: returns (Ltens, Rtens, s, truncerr, qnew)."""
    Aexp = Aexp / np.linalg.norm(Aexp)
    ax, d, b = Aexp.shape
    z = None if qP is None else np.zeros((1, qP.shape[1]), dtype=np.int64)
    U, s, V, truncerr, qnew = od.fsvd(Aexp.reshape(ax, 1, d, b), m, minm, cutoff, qLexp, z, qR, qP2=qP)
    na = Lpsi.shape[2]
    UD = U[:na, 0, :] * s[None, :]
    return np.tensordot(Lpsi, UD, axes=(2, 0)), V, s, truncerr, qnew


def alpha_update(alpha, truncerr, cutoff, lastEnergy, currenergy, noise_goal, dalpha, avgE=None):
    """algs/DMRG.jl:288-303 with IEEE division (x/0 = +-Inf, 0/0 = NaN)."""
    avgE = currenergy if avgE is None else avgE
    trunc_check = truncerr <= max(cutoff, 1e-12)
    errEst = abs(currenergy - lastEnergy) <= max(abs(currenergy * cutoff), 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.float64(lastEnergy) / np.float64(currenergy) - 1.0
        sitecond = bool(np.abs(np.float64(truncerr * avgE) / ratio) > noise_goal)
    if currenergy <= lastEnergy and (sitecond or errEst or trunc_check):
        alpha *= 1.0 - dalpha
    else:
        alpha *= 1.0 + dalpha
    return max(min(alpha, ALPHA_MAX), 1e-42)


class Params3S(od.Params):
    def __init__(self, L, r=2, noise=NOISE, noise_goal=NOISE_GOAL, noise_incr=NOISE_INCR, **kw):
        super().__init__(L, **kw)
        self.r = r
        self.noise = noise
        self.noise_goal = noise_goal
        self.noise_incr = noise_incr


def step3S(j, i, iL, iR, psi, mpo, Lenv, Renv, params, deltas=None):
    """algs/DMRG.jl:322-360.  deltas: MPO link charges (quantum-number runs)."""
    W = mpo[i - 1]
    energy, A = lanczos_ritz(lambda t: od.singlesite_update(Lenv[i], Renv[i], t, W), psi[i], params.r)
    params.energy = energy
    noise = params.noise
    cond = noise > params.noise_incr
    qn = psi.qn
    if j > 0:
        qexp = psi.qbond[iL] if qn else None
        if cond:
            A = np.concatenate([A, Lexpand(A, W, Lenv[iL], noise)], axis=2)
            if qn:
                qexp = expand_charges_right(psi.qbond[iL], deltas[iL])
        Lt, Rt, s, truncerr, qnew = move_right(A, psi[iR], params.maxm, params.minm, params.cutoff,
                                               psi.qbond[iL - 1] if qn else None, psi.qphys, qexp)
    else:
        qexp = psi.qbond[iL] if qn else None
        if cond:
            A = np.concatenate([A, Rexpand(A, W, Renv[iR], noise)], axis=0)
            if qn:
                qexp = expand_charges_left(psi.qbond[iL], deltas[iR - 1])
        Lt, Rt, s, truncerr, qnew = move_left(psi[iL], A, params.maxm, params.minm, params.cutoff,
                                              qexp, psi.qphys, psi.qbond[iR] if qn else None)
    psi[iL], psi[iR] = Lt, Rt
    if qn:
        psi.qbond[iL] = qnew
    params.noise = alpha_update(noise, truncerr, params.cutoff, params.lastenergy, params.energy,
                                params.noise_goal, params.noise_incr)
    params.truncerr = truncerr
    params.biggestm = max(params.biggestm, len(s))
    od.SvNcheck(i, j, s, len(psi), params)


def dmrg3S(psi, mpo, m=0, minm=2, sweeps=1, cutoff=0.0, goal=0.0, r=2, noise=NOISE,
           noise_goal=NOISE_GOAL, noise_incr=NOISE_INCR, SvNbond=None):
    """dmrg(...; method="3S").  Modifies psi in place; returns (energy,
    per-sweep history), each history entry also holding the noise value."""
    Ns = len(psi)
    params = Params3S(Ns, r=r, noise=noise, noise_goal=noise_goal, noise_incr=noise_incr, m=m, minm=minm,
                      sweeps=sweeps, cutoff=cutoff, goal=goal, SvNbond=SvNbond)
    deltas = omps.mpo_deltas(mpo, psi.qphys) if psi.qn else None
    Lenv, Renv = od.makeEnv(psi, mpo)
    H1 = od.singlesite_update(Lenv[psi.oc], Renv[psi.oc], psi[psi.oc], mpo[psi.oc - 1])
    params.energy = float(np.real(np.vdot(psi[psi.oc], H1)))
    params.lastenergy = params.energy
    Nsteps = 2 * (Ns - 1)
    j = 1 if psi.oc == 1 else (-1 if psi.oc == Ns else 1)
    history = []
    for n in range(1, sweeps + 1):
        for _ in range(Nsteps):
            i = psi.oc
            iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
            step3S(j, i, iL, iR, psi, mpo, Lenv, Renv, params, deltas)
            if j > 0:
                Lenv[i + 1] = od.Lupdate(Lenv[i], psi[i], mpo[i - 1])
            else:
                Renv[i - 1] = od.Rupdate(Renv[i], psi[i], mpo[i - 1])
            psi.oc += j
            if (psi.oc == Ns and j > 0) or (psi.oc == 1 and j < 0):
                j = -j
        stop = od.dmrgcvg(n, params, history)
        history[-1]["noise"] = float(params.noise)
        if stop:
            break
    return params.energy, history
