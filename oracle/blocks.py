"""Block-sparse (charge-sector) H_eff*theta on the CPU (oracle).

TEST INFRASTRUCTURE ONLY: used by the tests (pinned against the dense
``oracle.dmrg.heff2``) and as the CPU baseline of bench.py.  It restates the
two-site effective-Hamiltonian application of algs/DMRG.jl:69-73 the way a
quantum-number tensor library contracts it (TensorPACK Qtens, called at
algs/DMRG.jl:49-53): one dense GEMM per allowed pair of charge sectors, on
numpy/OpenBLAS, with no work on forbidden blocks.

Conventions: bond bases sorted by charge (sectors are contiguous ranges);
Lenv[bra, w, ket], Renv[ket, u, bra]; MPO link charges delta with
delta(u) = delta(w) + q(k) - q(j).
"""
import numpy as np


def sectors(q):
    """Contiguous charge sectors of a sorted charge list: {charge: slice}."""
    q = np.asarray(q)
    out = {}
    start = 0
    n = q.shape[0]
    for i in range(1, n + 1):
        if i == n or not np.array_equal(q[i], q[start]):
            out[tuple(int(x) for x in q[start])] = slice(start, i)
            start = i
    return out


def _add(a, b):
    return tuple(int(x) + int(y) for x, y in zip(a, b))


def _sub(a, b):
    return tuple(int(x) - int(y) for x, y in zip(a, b))


class Heff2Plan:
    """Precomputed sector bookkeeping for y = H_eff theta on one bond."""

    def __init__(self, L, R, W1, W2, qL, qP, qR, dL, dR):
        self.L, self.R = L, R
        self.sL, self.sR = sectors(qL), sectors(qR)
        self.qP = [tuple(int(x) for x in r) for r in np.asarray(qP)]
        self.dL = [tuple(int(x) for x in r) for r in np.asarray(dL)]
        self.dR = [tuple(int(x) for x in r) for r in np.asarray(dR)]
        d = W1.shape[1]
        self.d = d
        W12 = np.tensordot(W1, W2, axes=(3, 0))          # [w, j1, k1, j2, k2, u]
        self.W12 = W12
        self.nzW = [tuple(int(i) for i in ix) for ix in np.argwhere(W12 != 0.0)]
        self.flops = 0.0
        self._count_flops()

    def _blocks(self):
        """Allowed theta blocks (a, j1, j2) -> b as sector keys."""
        for a in self.sL:
            for j1 in range(self.d):
                for j2 in range(self.d):
                    b = _add(_add(a, self.qP[j1]), self.qP[j2])
                    if b in self.sR:
                        yield a, j1, j2, b

    def _count_flops(self):
        f = 0.0
        D1 = self.W12.shape[0]
        D2 = self.W12.shape[5]
        for (a, j1, j2, b) in self._blocks():
            na, nb = self.sL[a].stop - self.sL[a].start, self.sR[b].stop - self.sR[b].start
            for w in range(D1):
                ap = _add(a, self.dL[w])
                if ap in self.sL:
                    nap = self.sL[ap].stop - self.sL[ap].start
                    f += 2.0 * nap * na * nb
        for ap in self.sL:
            for k1 in range(self.d):
                for k2 in range(self.d):
                    bp = _add(_add(ap, self.qP[k1]), self.qP[k2])
                    if bp not in self.sR:
                        continue
                    nap = self.sL[ap].stop - self.sL[ap].start
                    nbp = self.sR[bp].stop - self.sR[bp].start
                    for u in range(D2):
                        b = _sub(bp, self.dR[u])
                        if b in self.sR:
                            nb = self.sR[b].stop - self.sR[b].start
                            f += 2.0 * nap * nb * nbp
        self.flops = f

    def apply(self, theta):
        L, R, d = self.L, self.R, self.d
        D1 = self.W12.shape[0]
        # step A: T1[(w, j1, j2)] = [(a, a', b, L[a', w, a] theta[a, j1, j2, b])]
        T1 = {}
        for (a, j1, j2, b) in self._blocks():
            blk = theta[self.sL[a], j1, j2, self.sR[b]]
            for w in range(D1):
                ap = _add(a, self.dL[w])
                if ap not in self.sL:
                    continue
                T1.setdefault((w, j1, j2), []).append((ap, b, L[self.sL[ap], w, self.sL[a]] @ blk))
        # step B: T3[(a', k1, k2, u, b)] = sum_{w, j1, j2} W12 T1
        T3 = {}
        for (w, j1, k1, j2, k2, u) in self.nzW:
            c = self.W12[w, j1, k1, j2, k2, u]
            for (ap, b, blk) in T1.get((w, j1, j2), ()):
                key = (ap, k1, k2, u, b)
                if key in T3:
                    T3[key] += c * blk
                else:
                    T3[key] = c * blk
        # step C: y[a', k1, k2, b'] = sum_u T3 R[b, u, b']
        y = np.zeros_like(theta)
        for (ap, k1, k2, u, b), blk in T3.items():
            bp = _add(b, self.dR[u])
            if bp not in self.sR:
                continue
            y[self.sL[ap], k1, k2, self.sR[bp]] += blk @ R[self.sR[b], u, self.sR[bp]]
        return y


class Heff1Plan:
    """y = H_1site A (algs/optimizeMPS.jl:503-511) by charge sectors: one dense
    GEMM per allowed pair of sectors (the single-site form that 3S applies)."""

    def __init__(self, L, R, W, qL, qP, qR, dL, dR):
        self.L, self.R, self.W = L, R, W
        self.sL, self.sR = sectors(qL), sectors(qR)
        self.qP = [tuple(int(x) for x in r) for r in np.asarray(qP)]
        self.dL = [tuple(int(x) for x in r) for r in np.asarray(dL)]
        self.dR = [tuple(int(x) for x in r) for r in np.asarray(dR)]
        self.d = W.shape[1]
        self.nzW = [tuple(int(i) for i in ix) for ix in np.argwhere(W != 0.0)]
        f = 0.0
        for (a, j, b) in self._blocks():
            na, nb = self._n(self.sL, a), self._n(self.sR, b)
            for w in range(W.shape[0]):
                ap = _add(a, self.dL[w])
                if ap in self.sL:
                    f += 2.0 * self._n(self.sL, ap) * na * nb
        for ap in self.sL:
            for k in range(self.d):
                bp = _add(ap, self.qP[k])
                if bp not in self.sR:
                    continue
                for u in range(W.shape[3]):
                    b = _sub(bp, self.dR[u])
                    if b in self.sR:
                        f += 2.0 * self._n(self.sL, ap) * self._n(self.sR, b) * self._n(self.sR, bp)
        self.flops = f

    @staticmethod
    def _n(s, key):
        return s[key].stop - s[key].start

    def _blocks(self):
        for a in self.sL:
            for j in range(self.d):
                b = _add(a, self.qP[j])
                if b in self.sR:
                    yield a, j, b

    def apply(self, A):
        L, R, W = self.L, self.R, self.W
        T1 = {}
        for (a, j, b) in self._blocks():
            blk = A[self.sL[a], j, self.sR[b]]
            for w in range(W.shape[0]):
                ap = _add(a, self.dL[w])
                if ap in self.sL:
                    T1.setdefault((w, j), []).append((ap, b, L[self.sL[ap], w, self.sL[a]] @ blk))
        T3 = {}
        for (w, j, k, u) in self.nzW:
            c = W[w, j, k, u]
            for (ap, b, blk) in T1.get((w, j), ()):
                key = (ap, k, u, b)
                if key in T3:
                    T3[key] += c * blk
                else:
                    T3[key] = c * blk
        y = np.zeros_like(A)
        for (ap, k, u, b), blk in T3.items():
            bp = _add(b, self.dR[u])
            if bp in self.sR:
                y[self.sL[ap], k, self.sR[bp]] += blk @ R[self.sR[b], u, self.sR[bp]]
        return y
