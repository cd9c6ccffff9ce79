"""Two-site DMRG restated in numpy FP64 (oracle).  TEST INFRASTRUCTURE ONLY.

Function by function:
  makeBoundary      lib/Env/makeEnv.jl:79-84   (1x1x1 tensor holding 1.0)
  Lupdate           lib/MPO_MPS/Lupdate.jl:17-26
  Rupdate           lib/MPO_MPS/Rupdate.jl:17-26
  makeEnv           lib/Env/makeEnv.jl:196-222
  expect            lib/MPO_MPS/expect.jl:19-45
  singlesite_update algs/optimizeMPS.jl:503-511
  make2site         algs/DMRG.jl:48-54
  heff2             algs/DMRG.jl:69-73 (second application in simplelanczos)
  twosite_update    algs/DMRG.jl:23-38 (third order of the same contraction)
  simplelanczos     algs/DMRG.jl:57-90
  fsvd              TensorPACK fsvd! as called at algs/DMRG.jl:387
                    (TensorPACK 0.14.1 is not vendored; the truncation rule
                    below is a documented choice: keep at most m values, drop
                    the smallest values whose summed sigma^2 stays below
                    cutoff, keep at least minm, truncerr = discarded sigma^2,
                    kept sigma renormalised to unit 2-norm; with charges one
                    SVD per centre-charge sector and one global sort)
  twostep           algs/DMRG.jl:362-405
  SvNcheck          algs/DMRG.jl:172-210
  dmrgcvg           algs/DMRG.jl:610-652
  optmps / dmrg     algs/optimizeMPS.jl:759-849, algs/DMRG.jl:549-589,675-691

Tensor index conventions (lib/Env/makeEnv.jl:17-26):
  A[alpha, s, beta];  W[w, j(ket), k(bra), u];
  Lenv[bra alpha', w, ket alpha];  Renv[ket beta, u, bra beta'].
"""
import numpy as np

BETATEST = 1e-10   # algs/DMRG.jl:57 (betatest keyword default)


def makeBoundary(dtype=np.float64):
    return np.ones((1, 1, 1), dtype=dtype)


def Lupdate(L, A, W):
    """L'[B,u,b] = sum L[x,w,a] conj(A[x,k,B]) W[w,j,k,u] A[a,j,b]."""
    T = np.tensordot(L, A.conj(), axes=(0, 0))            # [w, a, k, B]
    T = np.tensordot(T, W, axes=([0, 2], [0, 2]))          # [a, B, j, u]
    return np.tensordot(T, A, axes=([0, 2], [0, 1]))       # [B, u, b]


def Rupdate(R, A, W):
    """R'[a,w,x] = sum A[a,j,b] W[w,j,k,u] conj(A[x,k,B]) R[b,u,B]."""
    T = np.tensordot(A.conj(), R, axes=(2, 2))             # [x, k, b, u]
    T = np.tensordot(W, T, axes=([2, 3], [1, 3]))          # [w, j, x, b]
    return np.tensordot(A, T, axes=([1, 2], [1, 3]))       # [a, w, x]


def makeEnv(psi, mpo):
    """Lenv[i] contracts sites 1..i-1, Renv[i] sites i+1..L (1-based lists,
    index 0 unused).  Environments are built up to the orthogonality centre."""
    L = len(psi)
    dt = np.result_type(psi[1].dtype, mpo[0].dtype)
    Lenv = [None] * (L + 1)
    Renv = [None] * (L + 1)
    Lenv[1] = makeBoundary(dt)
    Renv[L] = makeBoundary(dt)
    for i in range(1, psi.oc):
        Lenv[i + 1] = Lupdate(Lenv[i], psi[i], mpo[i - 1])
    for i in range(L, psi.oc, -1):
        Renv[i - 1] = Rupdate(Renv[i], psi[i], mpo[i - 1])
    return Lenv, Renv


def expect(psi, mpo):
    """<psi|H|psi> by a right-to-left sweep of Rupdate (expect.jl:19-45)."""
    R = makeBoundary(np.result_type(psi[1].dtype, mpo[0].dtype))
    for i in range(len(psi), 0, -1):
        R = Rupdate(R, psi[i], mpo[i - 1])
    return R[0, 0, 0]


def singlesite_update(L, R, A, W):
    """H_1site A  [x, k, B]  (optimizeMPS.jl:503-511)."""
    T = np.tensordot(L, A, axes=(2, 0))                     # [x, w, j, b]
    T = np.tensordot(T, W, axes=([1, 2], [0, 1]))           # [x, b, k, u]
    return np.tensordot(T, R, axes=([1, 3], [0, 1]))        # [x, k, B]


def make2site(L, R, AL, AR, WL, WR):
    """H_eff applied to psiL*psiR, contraction order of DMRG.jl:48-54."""
    T = np.tensordot(L, AL, axes=(2, 0))                    # [x, w, j1, g]
    T = np.tensordot(T, WL, axes=([1, 2], [0, 1]))          # [x, g, k1, v]
    T = np.tensordot(T, AR, axes=(1, 0))                    # [x, k1, v, j2, b]
    T = np.tensordot(T, WR, axes=([2, 3], [0, 1]))          # [x, k1, b, k2, u]
    return np.tensordot(T, R, axes=([2, 4], [0, 1]))        # [x, k1, k2, B]


def heff2(L, R, theta, WL, WR):
    """H_eff applied to a fused two-site tensor, order of DMRG.jl:69-73."""
    T = np.tensordot(L, theta, axes=(2, 0))                 # [x, w, j1, j2, b]
    ops = np.tensordot(WL, WR, axes=(3, 0))                 # [w, j1, k1, j2, k2, u]
    T = np.tensordot(T, ops, axes=([1, 2, 3], [0, 1, 3]))   # [x, b, k1, k2, u]
    return np.tensordot(T, R, axes=([1, 4], [0, 1]))        # [x, k1, k2, B]


def twosite_update(L, R, AL, AR, WL, WR):
    """Third contraction order (DMRG.jl:23-38); its result is unused by twostep."""
    AA = np.tensordot(AL, AR, axes=(2, 0))
    ops = np.tensordot(WL, WR, axes=(3, 0)).transpose(0, 2, 4, 5, 1, 3)  # [w,k1,k2,u,j1,j2]
    H = np.tensordot(ops, AA, axes=([4, 5], [1, 2]))        # [w,k1,k2,u,a,b]
    H = np.tensordot(L, H, axes=([1, 2], [0, 4]))           # [x,k1,k2,u,b]
    return np.tensordot(H, R, axes=([4, 3], [0, 1]))        # [x,k1,k2,B]


def eig2(a1, b1, a2):
    """Lowest eigenpair of [[a1, b1], [b1, a2]] via LAPACK (Julia eigen of a
    symmetric matrix, DMRG.jl:80-82)."""
    M = np.array([[a1, b1], [b1, a2]], dtype=np.float64)
    D, U = np.linalg.eigh(M)
    return D[0], U[0, 0], U[1, 0]


def simplelanczos(L, R, AL, AR, WL, WR, betatest=BETATEST):
    """algs/DMRG.jl:57-90.  Returns (energy, AA')."""
    Hpsi = make2site(L, R, AL, AR, WL, WR)
    AA = np.tensordot(AL, AR, axes=(2, 0))
    AA = AA / np.linalg.norm(AA)
    alpha1 = np.real(np.vdot(AA, Hpsi))
    psi2 = Hpsi - alpha1 * AA
    beta1 = np.linalg.norm(psi2)
    if beta1 > betatest:
        psi2 = psi2 / beta1
        Hpsi2 = heff2(L, R, psi2, WL, WR)
        alpha2 = np.real(np.vdot(psi2, Hpsi2))
        energy, u11, u21 = eig2(alpha1, beta1, alpha2)
        outAA = np.conj(u11) * AA + np.conj(u21) * psi2
    else:
        energy = alpha1
        outAA = AA
    return energy, outAA


def truncation_count(s_desc, m, minm, cutoff):
    """Number of singular values kept (documented TensorPACK fsvd! rule)."""
    n = len(s_desc)
    keep = n
    acc = 0.0
    while keep > 0 and acc + s_desc[keep - 1] ** 2 < cutoff:
        acc += s_desc[keep - 1] ** 2
        keep -= 1
    if m > 0:
        keep = min(keep, m)
    keep = max(keep, min(minm, n))
    return keep


def fsvd(AA, m=0, minm=2, cutoff=0.0, qL=None, qP=None, qR=None, qP2=None):
    """Two-site SVD split AA[(a,s1),(s2,b)] = U diag(s) V with truncation.

    Returns U[a,s1,k], s[k], V[k,s2,b], truncerr, qnew (charges of the new
    bond, or None).  With charges the new bond is ordered by sector charge
    (ascending, lexicographic) and by descending sigma inside a sector.
    qP charges s1 and, unless qP2 is given, s2 as well.
    """
    a, d1, d2, b = AA.shape
    M = AA.reshape(a * d1, d2 * b)
    if qL is None:
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        keep = truncation_count(s, m, minm, cutoff)
        truncerr = float(np.sum(s[keep:] ** 2))
        U, s, Vt = U[:, :keep], s[:keep], Vt[:keep]
        s = s / np.linalg.norm(s)
        return U.reshape(a, d1, keep), s, Vt.reshape(keep, d2, b), truncerr, None
    nq = qP.shape[1]
    qP2 = qP if qP2 is None else qP2
    qrow = (qL[:, None, :] + qP[None, :, :]).reshape(a * d1, nq)
    qcol = (qR[None, :, :] - qP2[:, None, :]).reshape(d2 * b, nq)
    keys = sorted(set(map(tuple, qrow)) & set(map(tuple, qcol)))
    sectors = []
    allS, allSec, allLoc = [], [], []
    for si, Q in enumerate(keys):
        r = np.nonzero(np.all(qrow == np.array(Q), axis=1))[0]
        c = np.nonzero(np.all(qcol == np.array(Q), axis=1))[0]
        Ub, sb, Vb = np.linalg.svd(M[np.ix_(r, c)], full_matrices=False)
        sectors.append((Q, r, c, Ub, sb, Vb))
        allS.append(sb)
        allSec.append(np.full(len(sb), si))
        allLoc.append(np.arange(len(sb)))
    S = np.concatenate(allS)
    SEC = np.concatenate(allSec)
    LOC = np.concatenate(allLoc)
    order = np.lexsort((LOC, SEC, -S))
    keep = truncation_count(S[order], m, minm, cutoff)
    truncerr = float(np.sum(S[order][keep:] ** 2))
    kept = order[:keep]
    Ucols, Vrows, svals, qnew = [], [], [], []
    for si, (Q, r, c, Ub, sb, Vb) in enumerate(sectors):
        loc = np.sort(LOC[kept[SEC[kept] == si]])
        for l in loc:
            u = np.zeros(a * d1, dtype=M.dtype)
            u[r] = Ub[:, l]
            v = np.zeros(d2 * b, dtype=M.dtype)
            v[c] = Vb[l]
            Ucols.append(u)
            Vrows.append(v)
            svals.append(sb[l])
            qnew.append(Q)
    U = np.stack(Ucols, axis=1)
    Vt = np.stack(Vrows, axis=0)
    s = np.array(svals)
    s = s / np.linalg.norm(s)
    k = len(s)
    return (U.reshape(a, d1, k), s, Vt.reshape(k, d2, b), truncerr,
            np.array(qnew, dtype=np.int64).reshape(k, nq))


class Params:
    """The algvars fields the two-site path reads or writes
    (algs/optimizeMPS.jl:36-106, defaults :514-577, SvNbond :614)."""

    def __init__(self, L, m=0, minm=2, sweeps=1, cutoff=0.0, goal=0.0, SvNbond=None):
        self.maxm = m
        self.minm = minm
        self.sweeps = sweeps
        self.cutoff = cutoff
        self.goal = goal
        self.SvNbond = (L + 1) // 2 if SvNbond is None else SvNbond
        self.energy = 0.0
        self.lastenergy = 0.0
        self.truncerr = 0.0
        self.maxtrunc = 0.0
        self.biggestm = 0
        self.entropy = 0.0
        self.storeD = [0.0] * 8
        self.maxshowD = 8


def SvNcheck(i, j, s, Ns, params):
    """algs/DMRG.jl:172-210 (allSvNbond=false path)."""
    if params.maxtrunc < params.truncerr:
        params.maxtrunc = params.truncerr
    if params.biggestm < len(s):
        params.biggestm = len(s)
    if i == min(max(2, params.SvNbond + 1), Ns) and j < 0:
        D = np.sort(np.asarray(s, dtype=np.float64))[::-1]
        p = D ** 2
        with np.errstate(divide='ignore', invalid='ignore'):
            params.entropy = float(-np.sum(np.where(p > 0, p * np.log(p), 0.0)))
        params.storeD = list(D[:min(len(D), params.maxshowD)])


def twostep(j, i, iL, iR, psi, mpo, Lenv, Renv, params):
    """algs/DMRG.jl:362-405 (the unused twosite_update result is omitted)."""
    if j > 0:
        psi[iL] = psi[iL] / np.linalg.norm(psi[iL])
    else:
        psi[iR] = psi[iR] / np.linalg.norm(psi[iR])
    energy, AA = simplelanczos(Lenv[iL], Renv[iR], psi[iL], psi[iR], mpo[iL - 1], mpo[iR - 1])
    params.energy = energy
    if psi.qn:
        U, s, V, truncerr, qnew = fsvd(AA, params.maxm, params.minm, params.cutoff,
                                       psi.qbond[iL - 1], psi.qphys, psi.qbond[iR])
        psi.qbond[iL] = qnew
    else:
        U, s, V, truncerr, _ = fsvd(AA, params.maxm, params.minm, params.cutoff)
    if j < 0:
        psi[iL] = U * s[None, None, :]
        psi[iR] = V
    else:
        psi[iL] = U
        psi[iR] = s[:, None, None] * V
    params.truncerr = truncerr
    params.biggestm = max(params.biggestm, len(s))
    SvNcheck(i, j, s, len(psi), params)


def dmrgcvg(n, params, history):
    """algs/DMRG.jl:610-652 (cvgE=true).  Records the per-sweep report."""
    history.append(dict(sweep=n, energy=float(params.energy), maxtrunc=float(params.maxtrunc),
                        biggestm=int(params.biggestm), entropy=float(params.entropy),
                        storeD=[float(x) for x in params.storeD]))
    params.biggestm = 0
    params.truncerr = params.maxtrunc
    params.maxtrunc = 0.0
    checkE = abs(params.energy - params.lastenergy)
    if checkE < params.goal:
        return True
    params.lastenergy = params.energy
    return False


def dmrg(psi, mpo, m=0, minm=2, sweeps=1, cutoff=0.0, goal=0.0, SvNbond=None):
    """Two-site DMRG (algs/DMRG.jl:549-589 -> dmrg_twosite :675-691 ->
    optmps algs/optimizeMPS.jl:759-849).  Modifies psi in place.
    Returns (energy, per-sweep history)."""
    Ns = len(psi)
    params = Params(Ns, m=m, minm=minm, sweeps=sweeps, cutoff=cutoff, goal=goal, SvNbond=SvNbond)
    Lenv, Renv = makeEnv(psi, mpo)
    # startval (optimizeMPS.jl:255-258, :365-370)
    H1 = singlesite_update(Lenv[psi.oc], Renv[psi.oc], psi[psi.oc], mpo[psi.oc - 1])
    params.energy = float(np.real(np.vdot(psi[psi.oc], H1)))
    params.lastenergy = params.energy
    Nsteps = 2 * (Ns - 1)
    if psi.oc == 1:
        j = 1
    elif psi.oc == Ns:
        j = -1
    else:
        j = 1
    history = []
    for n in range(1, sweeps + 1):
        for _ in range(Nsteps):
            i = psi.oc
            iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
            twostep(j, i, iL, iR, psi, mpo, Lenv, Renv, params)
            if j > 0:
                Lenv[i + 1] = Lupdate(Lenv[i], psi[i], mpo[i - 1])
            else:
                Renv[i - 1] = Rupdate(Renv[i], psi[i], mpo[i - 1])
            psi.oc += j
            if (psi.oc == Ns and j > 0) or (psi.oc == 1 and j < 0):
                j = -j
        if dmrgcvg(n, params, history):
            break
    return params.energy, history
