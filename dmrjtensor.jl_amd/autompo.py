"""Automatic MPO construction from operator terms (host side).

Mirrors lib/MPO/autoMPO.jl:
  mpoterm(val, O1, i1, O2, i2, ..., [trail])   :38-63  one product term
  MPO(term, base)                              :430-451 term as a bond-1 MPO
  MPO(terms)                                   :453-556 sum of terms, compressed

Each term becomes a bond-dimension-1 MPO ("Pauli string"): identities, the
term's operators at their sites (val on the first), and the trail (fermion
string) multiplied onto every site left of each operator, processed from the
last operator to the first as the reference does (for c+_i c_j, i < j:
(c+ F)_i F_{i+1} ... F_{j-1} c_j).  Terms are summed by direct sums of the
link spaces and the sum is compressed exactly: left-to-right and
right-to-left sweeps of SVDs of every link, taken separately inside each
link charge when physical charges are given (so the result conserves them),
keeping singular values above ``tol`` times the largest of the link; the
compression works on W / sqrt(d) so that identity strings keep unit norm.  The
compressed MPO is not the reference's (compressMPO!/deparallelize! keep other
gauges); it represents the same operator, which is what the tests check.

W tensors follow makeMPO's convention W[a, j, k, b] = O[j, k] (ket index j,
bra index k), with one-dimensional edge links.
"""
import numpy as np


class mpoterm:
    """mpoterm(val, O1, i1, O2, i2, ..., [trail]) or mpoterm(O1, i1, ...) with
    1-based sites (lib/MPO/autoMPO.jl:38-63)."""

    def __init__(self, *args):
        args = list(args)
        val = 1.0
        if np.isscalar(args[0]):
            val = args.pop(0)
        trail = None
        if len(args) % 2 == 1:
            trail = np.asarray(args.pop(), dtype=np.float64)
        if not args:
            raise ValueError("mpoterm needs at least one operator and site")
        self.val = float(val)
        self.T = [np.asarray(args[k], dtype=np.float64) for k in range(0, len(args), 2)]
        self.ind = [int(args[k]) for k in range(1, len(args), 2)]
        self.trail = trail

    def __add__(self, other):
        return [self] + (other if isinstance(other, list) else [other])

    def __radd__(self, other):
        return (other if isinstance(other, list) else [other]) + [self]

    def __len__(self):
        return len(self.T)


def _string(term, Ns, d):
    """lib/MPO/autoMPO.jl:430-451: the term's operator on every site."""
    ops = [np.eye(d[i]) for i in range(Ns)]
    for w in range(len(term) - 1, -1, -1):
        x = term.ind[w] - 1
        ops[x] = term.T[w] @ ops[x]
        if w == 0:
            ops[x] = ops[x] * term.val
        if term.trail is not None:
            for z in range(x):
                ops[z] = term.trail @ ops[z]
    return ops


def _link_charges(Ws, qphys):
    from .network import mpo_deltas
    return mpo_deltas(Ws, qphys)


def _compress_link(Wl, Wr, ql, tol, left, qphys=None, qa=None, qc=None):
    """Exact rank reduction of the link between Wl [a, j, k, b] and Wr [b, j, k, c]
    (ql: charge of every channel b, or None).  left=True factors Wl, else Wr.
    With charges (qa, qc: charges of links a and c), every entry the charge
    rule forbids is set to zero in both new tensors, removing the rounding
    noise of the orthogonal factors.  Returns the tensors and the charges of
    the new channels."""
    a, d1, d2, b = Wl.shape
    c = Wr.shape[3]
    groups = [np.arange(b)] if ql is None else [np.flatnonzero((ql == q).all(axis=1)) for q in np.unique(ql, axis=0)]
    newl, newr, newq = [], [], []
    for g in groups:
        if left:
            M = Wl[..., g].reshape(a * d1 * d2, len(g))
            U, S, Vt = np.linalg.svd(M, full_matrices=False)
            keep = S > tol * max(S[0] if S.size else 0.0, 1e-300)
            U, S, Vt = U[:, keep], S[keep], Vt[keep]
            newl.append(U.reshape(a, d1, d2, -1))
            newr.append(np.tensordot(S[:, None] * Vt, Wr[g], axes=(1, 0)))
        else:
            M = Wr[g].reshape(len(g), Wr.shape[1] * Wr.shape[2] * c)
            U, S, Vt = np.linalg.svd(M, full_matrices=False)
            keep = S > tol * max(S[0] if S.size else 0.0, 1e-300)
            U, S, Vt = U[:, keep], S[keep], Vt[keep]
            newr.append(Vt.reshape(-1, Wr.shape[1], Wr.shape[2], c))
            newl.append(np.tensordot(Wl[..., g], U * S[None, :], axes=(3, 0)))
        if ql is not None:
            newq.append(np.repeat(ql[g[:1]], newl[-1].shape[3], axis=0))
    Wl2 = np.concatenate(newl, axis=3)
    Wr2 = np.concatenate(newr, axis=0)
    qb = np.concatenate(newq, axis=0) if ql is not None else None
    if ql is not None:
        q = np.asarray(qphys)
        dq = q[None, :, :] - q[:, None, :]          # [j, k] -> q[k] - q[j]
        okl = ((qa[:, None, None, None, :] + dq[None, :, :, None, :]) == qb[None, None, None, :, :]).all(axis=4)
        okr = ((qb[:, None, None, None, :] + dq[None, :, :, None, :]) == qc[None, None, None, :, :]).all(axis=4)
        Wl2[~okl] = 0.0
        Wr2[~okr] = 0.0
    else:
        for X in (Wl2, Wr2):
            X[np.abs(X) < 1e-14 * max(np.max(np.abs(X)), 1e-300)] = 0.0
    return Wl2, Wr2, qb


def _direct_sum(A, B):
    """Sum of two MPOs by direct sums of their link spaces."""
    Ns = len(A)
    out = []
    for i in range(Ns):
        a, b = A[i], B[i]
        if Ns == 1:
            out.append(a + b)
            continue
        if i == 0:
            out.append(np.concatenate([a, b], axis=3))
        elif i == Ns - 1:
            out.append(np.concatenate([a, b], axis=0))
        else:
            W = np.zeros((a.shape[0] + b.shape[0], a.shape[1], a.shape[2], a.shape[3] + b.shape[3]))
            W[:a.shape[0], :, :, :a.shape[3]] = a
            W[a.shape[0]:, :, :, a.shape[3]:] = b
            out.append(W)
    return out


def _compress(Ws, qphys, tol):
    Ns = len(Ws)
    lq = _link_charges(Ws, qphys) if qphys is not None else None
    for _ in range(2):
        for left in (True, False):
            rng = range(Ns - 1) if left else range(Ns - 2, -1, -1)
            for i in rng:
                if lq is None:
                    Ws[i], Ws[i + 1], _ = _compress_link(Ws[i], Ws[i + 1], None, tol, left)
                else:
                    Ws[i], Ws[i + 1], lq[i + 1] = _compress_link(Ws[i], Ws[i + 1], lq[i + 1], tol, left,
                                                                 qphys, lq[i], lq[i + 2])
    return Ws


def autoMPO(terms, Ns=None, d=None, qphys=None, tol=1e-13, chunk=32):
    """MPO(terms) (lib/MPO/autoMPO.jl:453-556): W tensors of the sum of the
    terms, compressed.  ``d``: physical dimension (int or per site; default
    from the operators); ``qphys``: physical charges [d, nq] to keep the
    compression charge-conserving.  Terms are added ``chunk`` at a time and
    the running sum is recompressed after each chunk (the reference
    recompresses every ``compresscheck`` terms), so link dimensions stay near
    the final ones."""
    if isinstance(terms, mpoterm):
        terms = [terms]
    Ns = Ns or max(max(t.ind) for t in terms)
    if d is None:
        dims = {}
        for t in terms:
            for o, i in zip(t.T, t.ind):
                dims[i] = o.shape[0]
        dd = set(dims.values())
        if len(dd) != 1:
            raise ValueError("autoMPO: give d when sites differ")
        d = dd.pop()
    d = [d] * Ns if np.isscalar(d) else list(d)
    total = None
    for c0 in range(0, len(terms), chunk):
        part = terms[c0:c0 + chunk]
        strings = [_string(t, Ns, d) for t in part]
        T = len(part)
        Ws = []
        for i in range(Ns):
            # scaled by 1/sqrt(d): the identity has unit norm, so the link
            # factors stay O(1) along the chain instead of growing as d^(n/2)
            W = np.zeros((1 if i == 0 else T, d[i], d[i], 1 if i == Ns - 1 else T))
            for t, ops in enumerate(strings):
                W[0 if i == 0 else t, :, :, 0 if i == Ns - 1 else t] = ops[i] / np.sqrt(d[i])
            Ws.append(W)
        total = Ws if total is None else _direct_sum(total, Ws)
        if Ns > 1:
            total = _compress(total, qphys, tol)
    return [W * np.sqrt(d[i]) for i, W in enumerate(total)]
