"""Two-site TDVP on the GPU path (algs/time.jl:288-362).

tdvp_twosite(psi, mpo; m, minm, cutoff, r, prefactor) keeps the reference's
conventions: the step per bond is prefactor/(Ns-1)/2, the single-site update
of the next centre runs forward with the same step (algs/time.jl:333,346) and
one call is one half-sweep.  Each bond: dmrj_theta2 -> dmrj_krylov_expm2
(Krylov exponential of the two-site H_eff) -> dmrj_normalize ->
dmrj_svd_trunc2 -> environment update -> dmrj_krylov_expm1 (one-site) ->
dmrj_normalize of the new centre.  A real prefactor evolves in imaginary
time; -i dt evolves in real time, and every tensor the evolution touches
becomes ComplexF64 (real and imaginary planes on the device).
"""
import ctypes as C

import numpy as np

from . import _lib
from .network import DeviceTensor, Lupdate, Rupdate, makeEnv


def _theta(ctx, a, b):
    h = C.c_void_p()
    ctx.call("dmrj_theta2", a.h, b.h, C.byref(h))
    return DeviceTensor(ctx, h)


def krylovTimeEvol2(Lenv, Renv, AA, mpo, iL, r, prefactor, precision="f64", energy=True):
    """krylovTimeEvol with joinedtwosite_update (algs/time.jl:169-202);
    complex prefactor (-i dt: real time).  energy=False leaves the lowest
    Ritz value on the device (None is returned): the call then has no host
    round trip."""
    ctx = AA.ctx
    h = C.c_void_p()
    e = C.c_double()
    k = C.c_int()
    pf = complex(prefactor)
    ctx.call("dmrj_krylov_expm2", Lenv.h, Renv.h, AA.h, mpo.device(ctx, iL), mpo.device(ctx, iL + 1), int(r),
             pf.real, pf.imag, _lib.PREC[precision], C.byref(h), C.byref(e) if energy else None,
             C.byref(k) if energy else None)
    return DeviceTensor(ctx, h), (e.value if energy else None)


def krylovTimeEvol1(Lenv, Renv, A, mpo, i, r, prefactor, precision="f64", energy=True):
    """krylovTimeEvol with singlesite_update (algs/time.jl:182-202);
    energy=False as for krylovTimeEvol2."""
    ctx = A.ctx
    h = C.c_void_p()
    e = C.c_double()
    k = C.c_int()
    pf = complex(prefactor)
    ctx.call("dmrj_krylov_expm1", Lenv.h, Renv.h, A.h, mpo.device(ctx, i), int(r), pf.real, pf.imag,
             _lib.PREC[precision], C.byref(h), C.byref(e) if energy else None, C.byref(k) if energy else None)
    return DeviceTensor(ctx, h), (e.value if energy else None)


def _normalize(t):
    t.ctx.call("dmrj_normalize", t.h)


def _svd(AA, m, minm, cutoff, direction):
    ctx = AA.ctx
    l, rr = C.c_void_p(), C.c_void_p()
    te = C.c_double()
    k = C.c_int64()
    cap = 8
    sv = np.zeros(cap)
    ctx.call("dmrj_svd_trunc2", AA.h, int(m), int(minm), float(cutoff), int(direction), C.byref(l), C.byref(rr),
             C.byref(te), C.byref(k), _lib.dptr(sv), cap)
    return DeviceTensor(ctx, l), DeviceTensor(ctx, rr), te.value


def tdvp_twosite(psi, mpo, m=0, minm=2, cutoff=0.0, r=2, prefactor=1.0, Lenv=None, Renv=None,
                 precision="f64"):
    """One half-sweep of two-site TDVP.  Returns (energy, Lenv, Renv); the
    environments can be passed back in for the next half-sweep.  A prefactor
    -i dt evolves in real time: the tensors touched become ComplexF64."""
    prefactor = complex(prefactor) if np.iscomplexobj(prefactor) else float(prefactor)
    Ns = len(psi)
    timestep = prefactor / (Ns - 1) / 2
    if Lenv is None or Renv is None:
        Lenv, Renv = makeEnv(psi, mpo)
    j = 1 if psi.oc == 1 else (-1 if psi.oc == Ns else 1)
    En = 0.0
    ctx = psi.ctx
    for step in range(Ns - 1):
        last = step == Ns - 2     # the energy the half-sweep returns is the last one's
        i = psi.oc
        iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
        AA = _theta(ctx, psi[iL], psi[iR])
        newpsi, _ = krylovTimeEvol2(Lenv[iL], Renv[iR], AA, mpo, iL, r, timestep, precision, energy=False)
        AA.free()
        _normalize(newpsi)
        left, right, _ = _svd(newpsi, m, minm, cutoff, 1 if j > 0 else -1)
        newpsi.free()
        psi[iL].free()
        psi[iR].free()
        if j > 0:
            psi[iL] = left
            Lenv[iR] = Lupdate(Lenv[iL], psi[iL], mpo, iL)
            _normalize(right)
            psi[iR], e1 = krylovTimeEvol1(Lenv[iR], Renv[iR], right, mpo, iR, r, timestep, precision, energy=last)
            right.free()
        else:
            psi[iR] = right
            Renv[iL] = Rupdate(Renv[iR], psi[iR], mpo, iR)
            _normalize(left)
            psi[iL], e1 = krylovTimeEvol1(Lenv[iL], Renv[iL], left, mpo, iL, r, timestep, precision, energy=last)
            left.free()
        if last:
            En = e1
        psi.oc += j
        _normalize(psi[psi.oc])
        if psi.oc == Ns and j > 0:
            j = -1
        elif psi.oc == 1 and j < 0:
            j = 1
    ctx.sync()
    return En, Lenv, Renv


def local_expect(psi, op, site):
    """<psi|O_site|psi>/<psi|psi> through a one-site operator MPO and the
    GPU environment contraction (expect, lib/MPO_MPS/expect.jl:19-45)."""
    from .network import MPO, expect
    d = op.shape[0]
    Ws = []
    for i in range(1, len(psi) + 1):
        W = np.zeros((1, d, d, 1))
        o = op if i == site else np.eye(d)
        W[0, :, :, 0] = o.T            # W[., j, k, .] = <k|O|j>
        Ws.append(W)
    qphys = psi.qphys
    mpo = MPO(Ws, qphys=qphys) if qphys is not None else MPO(Ws)
    num = expect(psi, mpo)
    Ws1 = []
    for i in range(1, len(psi) + 1):
        W = np.zeros((1, d, d, 1))
        W[0, :, :, 0] = np.eye(d)
        Ws1.append(W)
    den = expect(psi, MPO(Ws1, qphys=qphys) if qphys is not None else MPO(Ws1))
    return num / den
