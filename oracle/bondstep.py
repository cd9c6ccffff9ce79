"""One reference two-site bond step on block-sparse (charge-sector) tensors.

TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py (the whole bond step
of the reference, timed on the host cores) and, through the tests, checked
against the dense oracle (oracle.dmrg) on random U(1) / U(1)xU(1) tensors.

It restates algs/DMRG.jl:362-405 (twostep) the way a quantum-number tensor
library executes it (TensorPACK Qtens: one dense GEMM per allowed pair of
charge sectors, nothing stored or computed for forbidden blocks):

  normalise the centre                 algs/DMRG.jl:365-369
  simplelanczos (two H_eff*psi)        algs/DMRG.jl:57-90
  twosite_update (result unused)       algs/DMRG.jl:23-38, call at :374
  fsvd! per centre-charge sector       algs/DMRG.jl:387 (LAPACK gesdd)
  split U | D*V  or  U*D | V           algs/DMRG.jl:389-395
  Lupdate / Rupdate of the moved bond  lib/MPO_MPS/Lupdate.jl:17-26,
                                       lib/MPO_MPS/Rupdate.jl:17-26,
                                       algs/optimizeMPS.jl:810,815

Block storage: a dict {key: ndarray}.  MPS tensor A[a, s, b] -> key (a, s, b)
with sector charges a, b (tuples) and b = a + q(s); two-site tensor
theta[a, j1, j2, b] -> (a, j1, j2, b); left environment Lenv[bra x, w, ket a]
-> (x, w, a) with a = x - delta(w); right environment Renv[ket b, u, bra b']
-> (b, u, b') with b' = b + delta(u).  Bond bases are sorted by charge, so a
sector is a contiguous range of the dense index.
"""
import numpy as np

from oracle import dmrg as od
from oracle.blocks import sectors, _add, _sub

# algorithmic flops of the H_eff*psi and environment-update GEMMs done so far
# (2 m n k per block product, the same count the GPU path reports)
FLOPS = [0.0]


def _mm(a, b):
    FLOPS[0] += 2.0 * a.shape[0] * a.shape[1] * b.shape[1]
    return a @ b


class Bond:
    """Sector ranges of one bond: {charge: slice} in dense-index order."""

    def __init__(self, q):
        q = np.asarray(q).reshape(len(q), -1)
        self.q = q
        self.sec = sectors(q)
        self.n = q.shape[0]

    def dim(self, c):
        s = self.sec[c]
        return s.stop - s.start


# ---------------------------------------------------------------- conversions
def mps_to_blocks(A, bl, qP, br):
    out = {}
    for a, sa in bl.sec.items():
        for s in range(A.shape[1]):
            b = _add(a, qP[s])
            if b in br.sec:
                out[(a, s, b)] = np.ascontiguousarray(A[sa, s, br.sec[b]])
    return out


def mps_to_dense(blk, bl, d, br):
    A = np.zeros((bl.n, d, br.n))
    for (a, s, b), x in blk.items():
        A[bl.sec[a], s, br.sec[b]] = x
    return A


def theta_to_dense(blk, bl, d1, d2, br):
    T = np.zeros((bl.n, d1, d2, br.n))
    for (a, j1, j2, b), x in blk.items():
        T[bl.sec[a], j1, j2, br.sec[b]] = x
    return T


def lenv_to_blocks(L, bl, dW):
    out = {}
    for x, sx in bl.sec.items():
        for w in range(L.shape[1]):
            a = _sub(x, dW[w])
            if a in bl.sec:
                out[(x, w, a)] = np.ascontiguousarray(L[sx, w, bl.sec[a]])
    return out


def renv_to_blocks(R, br, dW):
    out = {}
    for b, sb in br.sec.items():
        for u in range(R.shape[1]):
            bp = _add(b, dW[u])
            if bp in br.sec:
                out[(b, u, bp)] = np.ascontiguousarray(R[sb, u, br.sec[bp]])
    return out


def lenv_to_dense(blk, bl, D):
    L = np.zeros((bl.n, D, bl.n))
    for (x, w, a), v in blk.items():
        L[bl.sec[x], w, bl.sec[a]] = v
    return L


def renv_to_dense(blk, br, D):
    R = np.zeros((br.n, D, br.n))
    for (b, u, bp), v in blk.items():
        R[br.sec[b], u, br.sec[bp]] = v
    return R


# ------------------------------------------------------------ vector algebra
def bnorm(x):
    return float(np.sqrt(sum(float(np.vdot(v, v).real) for v in x.values())))


def bdot(x, y):
    return float(sum(np.vdot(x[k], y[k]).real for k in x if k in y))


def bcomb(a, x, b, y):
    """a*x + b*y on the union of the block keys."""
    out = {}
    for k in set(x) | set(y):
        if k in x and k in y:
            out[k] = a * x[k] + b * y[k]
        elif k in x:
            out[k] = a * x[k]
        else:
            out[k] = b * y[k]
    return out


# -------------------------------------------------------------- contractions
def theta_blocks(AL, AR):
    """contract(psiL, 3, psiR, 1) (algs/DMRG.jl:60) sector by sector."""
    byleft = {}
    for (g, j2, b), x in AR.items():
        byleft.setdefault(g, []).append((j2, b, x))
    out = {}
    for (a, j1, g), x in AL.items():
        for (j2, b, y) in byleft.get(g, ()):
            key = (a, j1, j2, b)
            p = x @ y
            if key in out:
                out[key] += p
            else:
                out[key] = p
    return out


def _w12(W1, W2):
    W12 = np.tensordot(W1, W2, axes=(3, 0))                # [w, j1, k1, j2, k2, u]
    return W12, [tuple(int(i) for i in ix) for ix in np.argwhere(W12 != 0.0)]


def heff2_blocks(Lb, Rb, W1, W2, th, dL, dR):
    """H_eff*theta in the order of algs/DMRG.jl:69-73: Lenv(3)*theta(1), then the
    fused MPO pair over (w, j1, j2), then Renv over (b, u)."""
    W12, nz = _w12(W1, W2)
    Lby = {}
    for (x, w, a), v in Lb.items():
        Lby.setdefault((w, a), []).append((x, v))
    T1 = {}
    for (a, j1, j2, b), blk in th.items():
        for w in range(W12.shape[0]):
            for (x, Lv) in Lby.get((w, a), ()):
                T1.setdefault((w, j1, j2), []).append((x, b, _mm(Lv, blk)))
    T3 = {}
    for (w, j1, k1, j2, k2, u) in nz:
        c = W12[w, j1, k1, j2, k2, u]
        for (x, b, blk) in T1.get((w, j1, j2), ()):
            key = (x, k1, k2, u, b)
            if key in T3:
                T3[key] += c * blk
            else:
                T3[key] = c * blk
    out = {}
    for (x, k1, k2, u, b), blk in T3.items():
        bp = _add(b, dR[u])
        Rv = Rb.get((b, u, bp))
        if Rv is None:
            continue
        key = (x, k1, k2, bp)
        if key in out:
            out[key] += _mm(blk, Rv)
        else:
            out[key] = _mm(blk, Rv)
    return out


def twosite_update_blocks(Lb, Rb, W1, W2, th, dL, dR):
    """twosite_update (algs/DMRG.jl:23-38), whose product twostep discards
    (:374): the fused MPO pair applied to theta first (every (w, u) channel
    pair gets its own copy of the two-site data), then Lenv over (w, a), then
    Renv over (b, u)."""
    W12, nz = _w12(W1, W2)
    byphys = {}
    for (a, j1, j2, b), blk in th.items():
        byphys.setdefault((j1, j2), []).append((a, b, blk))
    H = {}
    for (w, j1, k1, j2, k2, u) in nz:
        c = W12[w, j1, k1, j2, k2, u]
        for (a, b, blk) in byphys.get((j1, j2), ()):
            key = (w, k1, k2, u, a, b)
            if key in H:
                H[key] += c * blk
            else:
                H[key] = c * blk
    Lby = {}
    for (x, w, a), v in Lb.items():
        Lby.setdefault((w, a), []).append((x, v))
    T = {}
    for (w, k1, k2, u, a, b), blk in H.items():
        for (x, Lv) in Lby.get((w, a), ()):
            key = (x, k1, k2, u, b)
            if key in T:
                T[key] += Lv @ blk
            else:
                T[key] = Lv @ blk
    out = {}
    for (x, k1, k2, u, b), blk in T.items():
        bp = _add(b, dR[u])
        Rv = Rb.get((b, u, bp))
        if Rv is None:
            continue
        key = (x, k1, k2, bp)
        if key in out:
            out[key] += blk @ Rv
        else:
            out[key] = blk @ Rv
    return out


def lupdate_blocks(Lb, A, W, dL, dR):
    """Lupdate (lib/MPO_MPS/Lupdate.jl:17-26) by sectors:
    L'[B, u, b] = sum L[x, w, a] conj(A[x, k, B]) W[w, j, k, u] A[a, j, b]."""
    Lby = {}
    for (x, w, a), v in Lb.items():
        Lby.setdefault((w, a), []).append((x, v))
    T1 = {}
    for (a, j, b), blk in A.items():
        for w in range(W.shape[0]):
            for (x, Lv) in Lby.get((w, a), ()):
                T1.setdefault((w, j), []).append((x, b, _mm(Lv, blk)))
    nz = [tuple(int(i) for i in ix) for ix in np.argwhere(W != 0.0)]
    T2 = {}
    for (w, j, k, u) in nz:
        c = W[w, j, k, u]
        for (x, b, blk) in T1.get((w, j), ()):
            key = (x, k, u, b)
            if key in T2:
                T2[key] += c * blk
            else:
                T2[key] = c * blk
    Aby = {}
    for (x, k, B), blk in A.items():
        Aby[(x, k)] = (B, blk)
    out = {}
    for (x, k, u, b), blk in T2.items():
        hit = Aby.get((x, k))
        if hit is None:
            continue
        B, Ab = hit
        key = (B, u, b)
        if key in out:
            out[key] += _mm(Ab.conj().T, blk)
        else:
            out[key] = _mm(Ab.conj().T, blk)
    return out


def rupdate_blocks(Rb, A, W, dL, dR):
    """Rupdate (lib/MPO_MPS/Rupdate.jl:17-26) by sectors:
    R'[a, w, x] = sum A[a, j, b] W[w, j, k, u] conj(A[x, k, B]) R[b, u, B]."""
    Rby = {}
    for (b, u, B), v in Rb.items():
        Rby.setdefault((b, u), []).append((B, v))
    T1 = {}
    for (a, j, b), blk in A.items():
        for u in range(W.shape[3]):
            for (B, Rv) in Rby.get((b, u), ()):
                T1.setdefault((j, u), []).append((a, B, _mm(blk, Rv)))
    nz = [tuple(int(i) for i in ix) for ix in np.argwhere(W != 0.0)]
    T2 = {}
    for (w, j, k, u) in nz:
        c = W[w, j, k, u]
        for (a, B, blk) in T1.get((j, u), ()):
            key = (a, w, k, B)
            if key in T2:
                T2[key] += c * blk
            else:
                T2[key] = c * blk
    Aby = {}
    for (x, k, B), blk in A.items():
        Aby[(k, B)] = (x, blk)
    out = {}
    for (a, w, k, B), blk in T2.items():
        hit = Aby.get((k, B))
        if hit is None:
            continue
        x, Ab = hit
        key = (a, w, x)
        if key in out:
            out[key] += _mm(blk, Ab.conj().T)
        else:
            out[key] = _mm(blk, Ab.conj().T)
    return out


def fsvd_blocks(th, bl, qP, br, m, minm, cutoff):
    """fsvd!(AA, [[1,2],[3,4]]; m, minm, cutoff, mag=1.) (algs/DMRG.jl:387)
    with one LAPACK gesdd per centre-charge sector and the truncation rule of
    oracle.dmrg.truncation_count over the globally sorted values.  Returns
    (U blocks, s per new sector, V blocks, truncerr, charges of the new bond
    sorted by sector)."""
    d = len(qP)
    rows, cols = {}, {}
    for a, sa in bl.sec.items():
        for j1 in range(d):
            Q = _add(a, qP[j1])
            rows.setdefault(Q, []).append((a, j1))
    for j2 in range(d):
        for b, sb in br.sec.items():
            Q = _sub(b, qP[j2])
            cols.setdefault(Q, []).append((j2, b))
    secs = []
    for Q in sorted(set(rows) & set(cols)):
        rr, cc = rows[Q], cols[Q]
        R = sum(bl.dim(a) for a, _ in rr)
        C = sum(br.dim(b) for _, b in cc)
        M = np.zeros((R, C))
        ro = 0
        for (a, j1) in rr:
            co = 0
            for (j2, b) in cc:
                blk = th.get((a, j1, j2, b))
                if blk is not None:
                    M[ro:ro + bl.dim(a), co:co + br.dim(b)] = blk
                co += br.dim(b)
            ro += bl.dim(a)
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        secs.append((Q, rr, cc, U, s, Vt))
    S = np.concatenate([x[4] for x in secs]) if secs else np.zeros(0)
    SEC = np.concatenate([np.full(len(x[4]), i) for i, x in enumerate(secs)]) if secs else np.zeros(0, int)
    LOC = np.concatenate([np.arange(len(x[4])) for x in secs]) if secs else np.zeros(0, int)
    order = np.lexsort((LOC, SEC, -S))
    keep = od.truncation_count(S[order], m, minm, cutoff)
    truncerr = float(np.sum(S[order][keep:] ** 2))
    kept = order[:keep]
    nrm = float(np.linalg.norm(S[kept]))
    Ub, Vb, sv, qnew = {}, {}, {}, []
    for si, (Q, rr, cc, U, s, Vt) in enumerate(secs):
        loc = np.sort(LOC[kept[SEC[kept] == si]])
        if len(loc) == 0:
            continue
        sv[Q] = s[loc] / nrm
        qnew.extend([Q] * len(loc))
        ro = 0
        for (a, j1) in rr:
            Ub[(a, j1, Q)] = np.ascontiguousarray(U[ro:ro + bl.dim(a), loc])
            ro += bl.dim(a)
        co = 0
        for (j2, b) in cc:
            Vb[(Q, j2, b)] = np.ascontiguousarray(Vt[loc, co:co + br.dim(b)])
            co += br.dim(b)
    qn = np.array(qnew, dtype=np.int64).reshape(len(qnew), -1) if qnew else np.zeros((0, len(next(iter(bl.sec)))))
    return Ub, sv, Vb, truncerr, qn


def simplelanczos_blocks(Lb, Rb, AL, AR, W1, W2, dL, dR, betatest=od.BETATEST):
    """simplelanczos (algs/DMRG.jl:57-90) on block tensors: two H_eff*psi
    applications, three inner products, the 2x2 Ritz solve."""
    th = theta_blocks(AL, AR)
    Hpsi = heff2_blocks(Lb, Rb, W1, W2, th, dL, dR)
    nrm = bnorm(th)
    AA = {k: v / nrm for k, v in th.items()}
    alpha1 = bdot(AA, Hpsi)
    psi2 = bcomb(1.0, Hpsi, -alpha1, AA)
    beta1 = bnorm(psi2)
    if beta1 > betatest:
        psi2 = {k: v / beta1 for k, v in psi2.items()}
        Hpsi2 = heff2_blocks(Lb, Rb, W1, W2, psi2, dL, dR)
        alpha2 = bdot(psi2, Hpsi2)
        energy, u11, u21 = od.eig2(alpha1, beta1, alpha2)
        return energy, bcomb(u11, AA, u21, psi2)
    return alpha1, AA


def bond_step_right(Lb, Rb, AL, AR, W1, W2, bl, bm, br, qP, dL, dM, dR, m, minm=2, cutoff=0.0,
                    discarded=True):
    """twostep with j = +1 at bond (i, i+1) followed by the Lupdate of
    optmps (algs/optimizeMPS.jl:810): returns (energy, new psiL blocks, new
    psiR blocks, new Lenv[i+1] blocks, new bond charges, truncerr).
    ``discarded`` keeps the reference's unused twosite_update product.
    dL, dM, dR: MPO link charges left of W1, between W1 and W2, right of W2."""
    n = bnorm(AL)
    AL = {k: v / n for k, v in AL.items()}
    energy, AA = simplelanczos_blocks(Lb, Rb, AL, AR, W1, W2, dL, dR)
    if discarded:
        twosite_update_blocks(Lb, Rb, W1, W2, theta_blocks(AL, AR), dL, dR)
    Ub, sv, Vb, truncerr, qnew = fsvd_blocks(AA, bl, qP, br, m, minm, cutoff)
    newR = {(Q, j2, b): sv[Q][:, None] * v for (Q, j2, b), v in Vb.items()}
    newL = Ub
    Lnext = lupdate_blocks(Lb, newL, W1, dL, dM)
    return energy, newL, newR, Lnext, qnew, truncerr
