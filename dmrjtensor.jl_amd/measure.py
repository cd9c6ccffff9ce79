"""Measurements on a GPU-resident MPS beyond expect.

correlationmatrix   lib/MPO_MPS/correlationmatrix.jl:21-83: the single-state
                    form (psi, Cc, Ca) and the two-state form (dualpsi, psi,
                    Cc, Ca); real or ComplexF64 MPS, real operators
correlation         lib/MPO_MPS/correlation.jl:74-237: n-point correlation
                    tensor <dualpsi| O1_{p1} ... On_{pn} |psi> with trails

rho[i, j] = <psi| (Cc F)_i F_{i+1} ... F_{j-1} (Ca)_j |psi> for i < j (the
trail F only when given), rho[i, i] = <psi| (Cc Ca)_i |psi>, rho[j, i] =
conj(rho[i, j]); operators act on the ket.  Every contraction is an
environment update on the GPU with a one-site operator in place of the MPO
tensor: left environments of the identity from site 1, right environments of
the identity from site L, then for each i one left chain through i .. L that
carries the operator string, closed at every j against the right environment
of site j+1 by the bilinear sum of the two environment tensors (they share
the bond, the operator link and therefore the block layout; no conjugation:
the bra's conjugate is inside the environments, so the sum is complex for a
complex state).  O(L^2)
environment updates, like the reference's loop.  The result is divided by
<psi|psi> (the reference assumes a normalised state).
"""
import ctypes as C

import numpy as np

from .network import MPO, DeviceTensor, makeBoundary, total_charge


def _op_mpo(ops):
    """Operator string as an MPO (one 1 x d x d x 1 tensor per operator,
    W[0, j, k, 0] = <k|O|j>) so that link charges follow from the charges of
    the operators before it."""
    Ws = []
    for o in ops:
        d = o.shape[0]
        W = np.zeros((1, d, d, 1))
        W[0, :, :, 0] = np.asarray(o, dtype=np.float64).T
        Ws.append(W)
    return Ws


class _Ops:
    def __init__(self, psi, Cc, Ca, trail):
        d = psi.d[0]
        if any(x != d for x in psi.d):
            raise ValueError("correlationmatrix: all sites need the same physical dimension")
        eye = np.eye(d)
        F = eye if trail is None else np.asarray(trail, dtype=np.float64)
        FCc = np.asarray(Cc, dtype=np.float64) @ F if trail is not None else np.asarray(Cc, dtype=np.float64)
        Ca = np.asarray(Ca, dtype=np.float64)
        q = psi.qphys

        def mk(ops):
            return MPO(_op_mpo(ops), qphys=q) if q is not None else MPO(_op_mpo(ops))
        # the site whose handle is used is the last one of each string, so its
        # left link carries the charge of the operators before it
        self.ident = (mk([eye]), 1)
        self.onsite = (mk([np.asarray(Cc, dtype=np.float64) @ Ca]), 1)
        self.first = (mk([FCc]), 1)
        self.trail = (mk([FCc, F]), 2)
        self.last = (mk([FCc, Ca]), 2)


def _left(ctx, Lenv, A, op):
    mpo, k = op
    h = C.c_void_p()
    ctx.call("dmrj_env_left", Lenv.h, A.h, mpo.device(ctx, k), C.byref(h))
    return DeviceTensor(ctx, h)


def _right(ctx, Renv, A, op):
    mpo, k = op
    h = C.c_void_p()
    ctx.call("dmrj_env_right", Renv.h, A.h, mpo.device(ctx, k), C.byref(h))
    return DeviceTensor(ctx, h)


def _close(Lx, Rx):
    """Full contraction of a left and a right environment of the same bond."""
    if Lx.info()[1:] != Rx.info()[1:] or Lx.nelem != Rx.nelem:
        raise RuntimeError("correlationmatrix: environment layouts differ")
    if Lx.is_complex or Rx.is_complex:
        return Lx.dotu(Rx)
    return Lx.dot(Rx)


def correlationmatrix(*args, trail=None):
    """correlationmatrix(psi, Cc, Ca; trail) or correlationmatrix(dualpsi, psi,
    Cc, Ca; trail) (lib/MPO_MPS/correlationmatrix.jl:21-83) for real or
    ComplexF64 MPS on the GPU (real operators); returns the L x L matrix on
    the host (complex for a complex state).  The single-state form is divided
    by <psi|psi>; the two-state form is the plain transition element
    <dualpsi| ... |psi> over full overlap environments, which is what the
    reference's loop (centres of both states moved to i, sites i..j
    contracted) gives whenever the parts of the two states outside i..j
    overlap as identities, e.g. for dualpsi = psi as the reference's test runs
    it (test/MPmeas_test.jl:316-318)."""
    from .network import MPS
    if len(args) == 4 and isinstance(args[1], MPS):
        dualpsi, psi, Cc, Ca = args
        if dualpsi is psi:
            return _corrmat_one(psi, Cc, Ca, trail)
        return _corrmat_two(dualpsi, psi, Cc, Ca, trail)
    if len(args) != 3:
        raise TypeError("correlationmatrix(psi, Cc, Ca) or correlationmatrix(dualpsi, psi, Cc, Ca)")
    return _corrmat_one(*args, trail)


def _corrmat_one(psi, Cc, Ca, trail):
    for o in (Cc, Ca) + (() if trail is None else (trail,)):
        if np.iscomplexobj(o):
            raise NotImplementedError("correlationmatrix: real operators only")
    cplx = any(t.is_complex for t in psi.A)
    ctx = psi.ctx
    L = len(psi)
    ops = _Ops(psi, Cc, Ca, trail)
    nq = psi.nq
    Lb = [None] * (L + 2)
    Rb = [None] * (L + 2)
    Lb[1] = makeBoundary(ctx, True, nq, np.zeros(nq, dtype=np.int64) if nq else None)
    Rb[L + 1] = makeBoundary(ctx, False, nq, total_charge(psi) if nq else None)
    for k in range(1, L):
        Lb[k + 1] = _left(ctx, Lb[k], psi[k], ops.ident)
    for k in range(L, 1, -1):
        Rb[k] = _right(ctx, Rb[k + 1], psi[k], ops.ident)
    last_left = _left(ctx, Lb[L], psi[L], ops.ident)
    norm2 = _close(last_left, Rb[L + 1])
    last_left.free()
    rho = np.zeros((L, L), dtype=np.complex128 if cplx else np.float64)
    tmp = []
    for i in range(1, L + 1):
        X = _left(ctx, Lb[i], psi[i], ops.onsite)
        rho[i - 1, i - 1] = _close(X, Rb[i + 1])
        X.free()
    for i in range(1, L):
        X = _left(ctx, Lb[i], psi[i], ops.first)
        for j in range(i + 1, L + 1):
            Y = _left(ctx, X, psi[j], ops.last)
            v = _close(Y, Rb[j + 1])
            Y.free()
            rho[i - 1, j - 1] = v
            rho[j - 1, i - 1] = np.conj(v)
            if j < L:
                X2 = _left(ctx, X, psi[j], ops.trail)
                X.free()
                X = X2
        X.free()
    for e in Lb + Rb + tmp:
        if e is not None:
            e.free()
    return rho / (norm2.real if cplx else norm2)


# ------------------------------------------------------- two-state measurements
class _SiteOps:
    """Device MPO sites (1 x d x d x 1) of one-site operators, created on
    demand for an (operator, left link charge) pair: W[0, j, k, 0] = O[k, j],
    right link charge = left + q(k) - q(j) of the operator's entries."""

    def __init__(self, ctx, d, qphys):
        self.ctx, self.d, self.qphys = ctx, d, qphys
        self.nq = 0 if qphys is None else qphys.shape[1]
        self.cache = {}

    def charge(self, O):
        if not self.nq:
            return ()
        nz = np.argwhere(np.asarray(O) != 0.0)
        dq = {tuple(int(v) for v in self.qphys[k] - self.qphys[j]) for k, j in nz}
        if len(dq) > 1:
            raise ValueError("operator does not carry one definite charge")
        return next(iter(dq)) if dq else (0,) * self.nq

    def site(self, O, dl):
        O = np.asarray(O, dtype=np.float64)
        key = (O.tobytes(), tuple(dl))
        h = self.cache.get(key)
        if h is None:
            from . import _lib
            d = self.d
            W = np.zeros((1, d, d, 1))
            W[0, :, :, 0] = O.T
            buf = np.asfortranarray(W).ravel(order="F")
            nq = self.nq
            qp = _lib.charges_arg(self.qphys, d, nq) if nq else None
            dr = tuple(a + b for a, b in zip(dl, self.charge(O)))
            l = np.array([dl], dtype=np.int32) if nq else None
            r = np.array([dr], dtype=np.int32) if nq else None
            hh = C.c_void_p()
            self.ctx.call("dmrj_mpo_create", nq, 1, d, 1, _lib.dptr(buf), _lib.iptr(qp), _lib.iptr(l),
                          _lib.iptr(r), C.byref(hh))
            h = self.cache[key] = hh
        return h

    def free(self):
        for h in self.cache.values():
            self.ctx.lib.dmrj_mpo_free(self.ctx.h, h)
        self.cache = {}


def _left2(ctx, Lenv, B, A, W):
    h = C.c_void_p()
    ctx.call("dmrj_env_left2", Lenv.h, B.h, A.h, W, C.byref(h))
    return DeviceTensor(ctx, h)


def _right2(ctx, Renv, B, A, W):
    h = C.c_void_p()
    ctx.call("dmrj_env_right2", Renv.h, B.h, A.h, W, C.byref(h))
    return DeviceTensor(ctx, h)


def _two_state_setup(dualpsi, psi):
    if len(dualpsi) != len(psi):
        raise ValueError("the two states have different lengths")
    if dualpsi.nq != psi.nq:
        raise ValueError("the two states carry different charge types")
    d = psi.d[0]
    if any(x != d for x in psi.d) or any(x != d for x in dualpsi.d):
        raise ValueError("all sites need the same physical dimension")
    if psi.nq:
        qa, qb = total_charge(dualpsi), total_charge(psi)
        if not np.array_equal(np.asarray(qa), np.asarray(qb)):
            raise NotImplementedError("two states of different total charge: the boundary environments "
                                      "of the GPU path carry one charge")
    return d


def _bare_envs(ctx, dualpsi, psi, ops):
    """Overlap environments with the identity: Lb[k] contracts sites 1..k-1,
    Rb[k] sites k..L (Lb[1], Rb[L+1] the boundaries)."""
    L = len(psi)
    nq = psi.nq
    eye = np.eye(psi.d[0])
    z = (0,) * nq
    Lb = [None] * (L + 2)
    Rb = [None] * (L + 2)
    Lb[1] = makeBoundary(ctx, True, nq, np.zeros(nq, dtype=np.int64) if nq else None)
    Rb[L + 1] = makeBoundary(ctx, False, nq, total_charge(psi) if nq else None)
    for k in range(L, 0, -1):
        Rb[k] = _right2(ctx, Rb[k + 1], dualpsi[k], psi[k], ops.site(eye, z))
    return Lb, Rb


def _close2(X, Y):
    if X.info()[1:] != Y.info()[1:] or X.nelem != Y.nelem:
        raise RuntimeError("environment layouts differ")
    if X.is_complex or Y.is_complex:
        return X.dotu(Y)
    return X.dot(Y)


def _corrmat_two(dualpsi, psi, Cc, Ca, trail):
    """Two-state correlationmatrix: <dualpsi| (Cc F)_i F.. (Ca)_j |psi> over
    the bra/ket environments of dmrj_env_left2 / dmrj_env_right2."""
    for o in (Cc, Ca) + (() if trail is None else (trail,)):
        if np.iscomplexobj(o):
            raise NotImplementedError("correlationmatrix: real operators only")
    ctx = psi.ctx
    d = _two_state_setup(dualpsi, psi)
    L = len(psi)
    nq = psi.nq
    cplx = any(t.is_complex for t in psi.A) or any(t.is_complex for t in dualpsi.A)
    ops = _SiteOps(ctx, d, psi.qphys)
    eye = np.eye(d)
    F = eye if trail is None else np.asarray(trail, dtype=np.float64)
    Cc = np.asarray(Cc, dtype=np.float64)
    Ca = np.asarray(Ca, dtype=np.float64)
    FCc = Cc @ F if trail is not None else Cc
    z = (0,) * nq
    qc = ops.charge(FCc)
    try:
        Lb, Rb = _bare_envs(ctx, dualpsi, psi, ops)
        for k in range(1, L):
            Lb[k + 1] = _left2(ctx, Lb[k], dualpsi[k], psi[k], ops.site(eye, z))
        rho = np.zeros((L, L), dtype=np.complex128 if cplx else np.float64)
        for i in range(1, L + 1):
            X = _left2(ctx, Lb[i], dualpsi[i], psi[i], ops.site(Cc @ Ca, z))
            rho[i - 1, i - 1] = _close2(X, Rb[i + 1])
            X.free()
        for i in range(1, L):
            X = _left2(ctx, Lb[i], dualpsi[i], psi[i], ops.site(FCc, z))
            for j in range(i + 1, L + 1):
                Y = _left2(ctx, X, dualpsi[j], psi[j], ops.site(Ca, qc))
                v = _close2(Y, Rb[j + 1])
                Y.free()
                rho[i - 1, j - 1] = v
                rho[j - 1, i - 1] = np.conj(v)
                if j < L:
                    X2 = _left2(ctx, X, dualpsi[j], psi[j], ops.site(F, qc))
                    X.free()
                    X = X2
            X.free()
        for e in Lb + Rb:
            if e is not None:
                e.free()
    finally:
        ops.free()
    return rho


def correlation(*args, sites=None, trail=None):
    """correlation(psi, O1, ..., On; sites, trail) or correlation(dualpsi, psi,
    O1, ..., On; sites, trail) (lib/MPO_MPS/correlation.jl:74-237) on the GPU:
    omega[y1, ..., yn] = <dualpsi| X |psi> with operator g at p_g =
    sites[g][y_g] (default: every site), X built as the reference builds its
    modified ket: operators n .. 2 applied in that order, each followed by its
    trail on every site left of it, then operator 1, whose trail covers the
    sites left of p1 that the scan over sites[1] has passed.  trail: None,
    one matrix for every operator, or a tuple cycled over the operators (an
    identity trail counts as none).  Single-site real operators; the total
    charge of O1 ... On must be zero (otherwise omega is identically zero and
    is returned as such).  Every contraction is an environment update on the
    GPU (dmrj_env_left2 / dmrj_env_right2): for each tuple (p2, ..., pn) the
    right environments are recomputed from the rightmost of those positions
    down from the operator-free ones, and operator 1 is scanned from the left
    with one left update per site, as the reference's loop does."""
    from .network import MPS
    if len(args) >= 3 and isinstance(args[1], MPS):
        dualpsi, psi, opsin = args[0], args[1], args[2:]
    else:
        psi, opsin = args[0], args[1:]
        dualpsi = psi
    n = len(opsin)
    if n < 1:
        raise ValueError("correlation needs at least one operator")
    for o in opsin:
        if np.iscomplexobj(o) or np.ndim(o) != 2:
            raise NotImplementedError("correlation: real single-site operators")
    ctx = psi.ctx
    d = _two_state_setup(dualpsi, psi)
    L = len(psi)
    nq = psi.nq
    eye = np.eye(d)
    O = [np.asarray(o, dtype=np.float64) for o in opsin]
    if trail is None or (isinstance(trail, tuple) and len(trail) == 0):
        trails = [None] * n
    elif isinstance(trail, tuple):
        trails = [np.asarray(trail[g % len(trail)], dtype=np.float64) for g in range(n)]
    else:
        trails = [np.asarray(trail, dtype=np.float64)] * n
    trails = [None if (t is None or np.array_equal(t, eye)) else t for t in trails]
    sites = tuple(range(1, L + 1) for _ in range(n)) if sites is None else tuple(tuple(s) for s in sites)
    if len(sites) != n:
        raise ValueError("one site list per operator")
    cplx = any(t.is_complex for t in psi.A) or any(t.is_complex for t in dualpsi.A)
    omega = np.zeros(tuple(len(s) for s in sites), dtype=np.complex128 if cplx else np.float64)
    ops = _SiteOps(ctx, d, psi.qphys)
    for t in trails:
        if t is not None and nq and any(ops.charge(t)):
            raise ValueError("the trail must be neutral")
    z = (0,) * nq
    qtot = tuple(int(sum(ops.charge(o)[k] for o in O)) for k in range(nq))
    if any(qtot):
        ops.free()
        return omega
    try:
        Lb, Rb = _bare_envs(ctx, dualpsi, psi, ops)
        for rest in np.ndindex(*[len(s) for s in sites[1:]]):
            pos = [None] + [sites[g + 1][rest[g]] for g in range(n - 1)]
            siteop = {}
            for g in range(n - 1, 0, -1):
                p = pos[g]
                siteop[p] = O[g] @ siteop.get(p, eye)
                if trails[g] is not None:
                    for s_ in range(1, p):
                        siteop[s_] = trails[g] @ siteop.get(s_, eye)
            maxpos = max(pos[1:]) if n > 1 else 0
            # charge of the link right of every site (counted from the left
            # end, as every MPO link charge is): with a zero total it is minus
            # the charge of the operators further right, whatever p1 is
            qright = {}
            acc = np.zeros(nq, dtype=np.int64)
            for s_ in range(L, 0, -1):
                qright[s_] = tuple(-int(v) for v in acc)
                if s_ in siteop and nq:
                    acc = acc + np.array(ops.charge(siteop[s_]), dtype=np.int64)
            R = {}
            for s_ in range(L + 1, maxpos, -1):
                R[s_] = Rb[s_]
            tmp = []
            for s_ in range(maxpos, 1, -1):
                M = siteop.get(s_, eye)
                dl = tuple(a - b for a, b in zip(qright[s_], ops.charge(M))) if nq else ()
                R[s_] = _right2(ctx, R[s_ + 1], dualpsi[s_], psi[s_], ops.site(M, dl))
                tmp.append(R[s_])
            # scan operator 1 from the left
            Lcur, lpos, qleft = Lb[1], 1, np.zeros(nq, dtype=np.int64)
            lown = []
            done = set()
            for y, w in enumerate(sites[0]):
                while lpos < w:
                    M = siteop.get(lpos, eye)
                    if trails[0] is not None and lpos in done:
                        M = trails[0] @ M
                    Ln = _left2(ctx, Lcur, dualpsi[lpos], psi[lpos], ops.site(M, tuple(int(v) for v in qleft)))
                    lown.append(Ln)
                    if nq:
                        qleft = qleft + np.array(ops.charge(M), dtype=np.int64)
                    Lcur, lpos = Ln, lpos + 1
                if lpos != w:
                    raise ValueError("correlation: the sites of the first operator must increase")
                M1 = O[0] @ siteop.get(w, eye)
                X = _left2(ctx, Lcur, dualpsi[w], psi[w], ops.site(M1, tuple(int(v) for v in qleft)))
                omega[(y,) + tuple(rest)] = _close2(X, R[w + 1])
                X.free()
                done.add(w)
            for e in lown + tmp:
                e.free()
        for e in Lb + Rb:
            if e is not None:
                e.free()
    finally:
        ops.free()
    return omega
