"""Measurements on a GPU-resident MPS beyond expect.

correlationmatrix   lib/MPO_MPS/correlationmatrix.jl:21-83 (single-state form;
                    real or ComplexF64 MPS, real operators)

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


def correlationmatrix(psi, Cc, Ca, trail=None):
    """correlationmatrix(psi, Cc, Ca; trail) (lib/MPO_MPS/correlationmatrix.jl:21-83)
    for a real or ComplexF64 MPS on the GPU (real operators); returns the
    L x L matrix on the host (complex for a complex state)."""
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
