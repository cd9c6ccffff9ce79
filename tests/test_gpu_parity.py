"""GPU parity of the hot path against the numpy oracle, through the C ABI.

Tolerance: FP64 path, max |gpu - oracle| <= 1e-12 * max |oracle| for
contractions and energies (the FP64 drop-in bar: <= 1e-12 relative); spectra compared
after sorting, gauge-dependent tensors compared through gauge-invariant
products.
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models as om

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(got, want):
    s = max(np.max(np.abs(want)), 1e-300)
    return float(np.max(np.abs(got - want)) / s)


def rand_mps(dims, d, rng):
    return [rng.standard_normal((dims[i], d, dims[i + 1])) for i in range(len(dims) - 1)]


def rand_qn_mps(secs, qphys, rng):
    """secs[b] = list of (charge tuple, dim) sorted by charge; returns dense
    tensors (zeros outside the allowed blocks) and per-basis bond charges."""
    nq = qphys.shape[1]
    qb = []
    for s in secs:
        q = []
        for (c, n) in s:
            q += [list(c)] * n
        qb.append(np.array(q, dtype=np.int64).reshape(-1, nq))
    A = []
    for i in range(len(secs) - 1):
        ql, qr = qb[i], qb[i + 1]
        T = np.zeros((len(ql), qphys.shape[0], len(qr)))
        allowed = np.all((ql[:, None, None, :] + qphys[None, :, None, :]) == qr[None, None, :, :], axis=-1)
        T[allowed] = rng.standard_normal(int(allowed.sum()))
        A.append(T)
    return A, qb


def upload_chain(pkg, ctx, A, qb=None, qphys=None):
    nq = 0 if qphys is None else qphys.shape[1]
    out = []
    for i, a in enumerate(A):
        if nq:
            out.append(pkg.site_from_dense(ctx, a, qb[i], qphys, qb[i + 1], nq))
        else:
            out.append(pkg.site_from_dense(ctx, a))
    return out


def gpu_envs(pkg, ctx, dev, mpo, upto_left, from_right, nq=0, qtot=None):
    """Left envs for sites 1..upto_left (Lenv[i] on the left bond of site i),
    right envs for sites from_right..L."""
    L = len(dev)
    Lenv = {1: pkg.network.makeBoundary(ctx, True, nq, np.zeros(nq) if nq else None)}
    for i in range(1, upto_left):
        Lenv[i + 1] = pkg.Lupdate(Lenv[i], dev[i - 1], mpo, i)
    Renv = {L: pkg.network.makeBoundary(ctx, False, nq, qtot)}
    for i in range(L, from_right, -1):
        Renv[i - 1] = pkg.Rupdate(Renv[i], dev[i - 1], mpo, i)
    return Lenv, Renv


def oracle_envs(A, mpo_h, upto_left, from_right):
    L = len(A)
    Lenv = {1: np.ones((1, 1, 1))}
    for i in range(1, upto_left):
        Lenv[i + 1] = od.Lupdate(Lenv[i], A[i - 1], mpo_h[i - 1])
    Renv = {L: np.ones((1, 1, 1))}
    for i in range(L, from_right, -1):
        Renv[i - 1] = od.Rupdate(Renv[i], A[i - 1], mpo_h[i - 1])
    return Lenv, Renv


DENSE_CASES = [
    # bond dimensions of a 4-site chain; the tested bond is sites (2,3)
    [1, 2, 3, 2, 1],
    [1, 2, 4, 2, 1],
    [1, 7, 16, 9, 1],
    [1, 33, 64, 40, 1],
    [1, 64, 130, 70, 1],
]


@pytest.mark.parametrize("dims", DENSE_CASES, ids=lambda d: "x".join(map(str, d)))
def test_dense_envs_heff_lanczos(pkg, ctx, dims):
    say("case dense dims", dims)
    rng = np.random.default_rng(sum(dims))
    L = len(dims) - 1
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    A = rand_mps(dims, 2, rng)
    dev = upload_chain(pkg, ctx, A)
    gL, gR = gpu_envs(pkg, ctx, dev, mpo, 2, 3)
    oL, oR = oracle_envs(A, mpo.H, 2, 3)
    for i in gL:
        e = relerr(gL[i].to_dense()[0], oL[i])
        say(" Lenv", i, e)
        assert e <= TOL
    for i in gR:
        e = relerr(gR[i].to_dense()[0], oR[i])
        say(" Renv", i, e)
        assert e <= TOL
    # theta and H_eff theta
    th = ctx_theta(pkg, ctx, dev[1], dev[2])
    oth = np.tensordot(A[1], A[2], axes=(2, 0))
    assert relerr(th.to_dense()[0], oth) <= TOL
    y = heff2(pkg, ctx, gL[2], mpo, gR[3], th)
    want = od.heff2(oL[2], oR[3], oth, mpo.H[1], mpo.H[2])
    e = relerr(y.to_dense()[0], want)
    say(" heff2", e)
    assert e <= TOL
    # single-site H_eff
    y1 = pkg.singlesite_update(gL[2], gR[2] if 2 in gR else pkg.Rupdate(gR[3], dev[2], mpo, 3), dev[1], mpo, 2)
    oR2 = od.Rupdate(oR[3], A[2], mpo.H[2])
    want1 = od.singlesite_update(oL[2], oR2, A[1], mpo.H[1])
    e = relerr(y1.to_dense()[0], want1)
    say(" heff1", e)
    assert e <= TOL
    # Lanczos energy
    E, AA = lanczos2(pkg, ctx, gL[2], gR[3], dev[1], dev[2], mpo)
    oE, oAA = od.simplelanczos(oL[2], oR[3], A[1], A[2], mpo.H[1], mpo.H[2])
    say(" lanczos2", E, oE, abs(E - oE) / abs(oE))
    assert abs(E - oE) <= TOL * abs(oE)
    assert min(relerr(AA.to_dense()[0], oAA), relerr(-AA.to_dense()[0], oAA)) <= 1e-10


def ctx_theta(pkg, ctx, a, b):
    import ctypes as C
    h = C.c_void_p()
    ctx.call("dmrj_theta2", a.h, b.h, C.byref(h))
    return pkg.DeviceTensor(ctx, h)


def heff2(pkg, ctx, L, mpo, R, th, iL=2):
    import ctypes as C
    h = C.c_void_p()
    ctx.call("dmrj_heff2_apply", L.h, mpo.device(ctx, iL), mpo.device(ctx, iL + 1), R.h, th.h, 0,
             C.byref(h))
    return pkg.DeviceTensor(ctx, h)


def lanczos2(pkg, ctx, L, R, a, b, mpo, iL=2):
    import ctypes as C
    h = C.c_void_p()
    e = C.c_double()
    ctx.call("dmrj_lanczos2", L.h, R.h, a.h, b.h, mpo.device(ctx, iL), mpo.device(ctx, iL + 1), 0,
             C.byref(h), C.byref(e))
    return e.value, pkg.DeviceTensor(ctx, h)


def svd_trunc(pkg, ctx, AA, m, minm, cutoff, direction):
    import ctypes as C
    l, r = C.c_void_p(), C.c_void_p()
    te = C.c_double()
    k = C.c_int64()
    sv = np.zeros(4096)
    ctx.call("dmrj_svd_trunc2", AA.h, m, minm, cutoff, direction, C.byref(l), C.byref(r), C.byref(te),
             C.byref(k), pkg._lib.dptr(sv), 4096)
    return pkg.DeviceTensor(ctx, l), pkg.DeviceTensor(ctx, r), te.value, sv[:k.value].copy()


QN_SECS = [
    # U(1) bond sectors (2Sz, dim) of a 4-site chain with total charge 0
    [[((0,), 1)], [((-1,), 1), ((1,), 1)], [((-2,), 1), ((0,), 2), ((2,), 1)], [((-1,), 1), ((1,), 1)],
     [((0,), 1)]],
    [[((0,), 1)], [((-1,), 3), ((1,), 4)], [((-2,), 5), ((0,), 9), ((2,), 6)], [((-1,), 7), ((1,), 5)],
     [((0,), 1)]],
    [[((0,), 1)], [((-3,), 2), ((-1,), 20), ((1,), 17), ((3,), 3)],
     [((-4,), 3), ((-2,), 19), ((0,), 40), ((2,), 21), ((4,), 2)],
     [((-3,), 4), ((-1,), 25), ((1,), 23), ((3,), 5)], [((0,), 1)]],
]


@pytest.mark.parametrize("ci", range(len(QN_SECS)))
def test_qn_envs_heff_lanczos_svd(pkg, ctx, ci):
    secs = QN_SECS[ci]
    say("case U(1) sectors", ci)
    rng = np.random.default_rng(100 + ci)
    qphys = np.array([[1], [-1]])
    A, qb = rand_qn_mps(secs, qphys, rng)
    L = len(A)
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    dev = upload_chain(pkg, ctx, A, qb, qphys)
    gL, gR = gpu_envs(pkg, ctx, dev, mpo, 2, 3, nq=1, qtot=np.array([0]))
    oL, oR = oracle_envs(A, mpo.H, 2, 3)
    for i in gL:
        e = relerr(gL[i].to_dense(1)[0], oL[i])
        say(" Lenv", i, e)
        assert e <= TOL
    for i in gR:
        e = relerr(gR[i].to_dense(1)[0], oR[i])
        say(" Renv", i, e)
        assert e <= TOL
    th = ctx_theta(pkg, ctx, dev[1], dev[2])
    oth = np.tensordot(A[1], A[2], axes=(2, 0))
    assert relerr(th.to_dense(1)[0], oth) <= TOL
    y = heff2(pkg, ctx, gL[2], mpo, gR[3], th)
    want = od.heff2(oL[2], oR[3], oth, mpo.H[1], mpo.H[2])
    e = relerr(y.to_dense(1)[0], want)
    say(" heff2", e)
    assert e <= TOL
    E, AA = lanczos2(pkg, ctx, gL[2], gR[3], dev[1], dev[2], mpo)
    oE, oAA = od.simplelanczos(oL[2], oR[3], A[1], A[2], mpo.H[1], mpo.H[2])
    say(" lanczos2", E, oE)
    assert abs(E - oE) <= TOL * abs(oE)
    # the Ritz vector is fixed up to a global sign; the split is checked on
    # the GPU's own vector
    gAA = AA.to_dense(1)[0]
    assert min(relerr(gAA, oAA), relerr(-gAA, oAA)) <= 1e-10
    for direction in (1, -1):
        for m in (0, 6):
            l, r, te, sv = svd_trunc(pkg, ctx, AA, m, 2, 1e-12, direction)
            U, s, V, ote, _ = od.fsvd(gAA, m=m, minm=2, cutoff=1e-12, qL=qb[1], qP=qphys, qR=qb[3])
            say(" svd dir", direction, "m", m, len(sv), len(s), te, ote)
            assert len(sv) == len(s)
            assert np.allclose(np.sort(sv)[::-1], np.sort(s)[::-1], rtol=1e-10, atol=1e-13)
            assert abs(te - ote) <= 1e-12 * max(1.0, ote) + 1e-15
            # gauge-invariant reconstruction psiL*psiR
            rec = np.tensordot(l.to_dense(1)[0], r.to_dense(1)[0], axes=(2, 0))
            orec = np.tensordot(U, s[:, None, None] * V, axes=(2, 0))
            assert relerr(rec, orec) <= 1e-10
            # the orthonormal factor is orthonormal
            if direction > 0:
                Ul = l.to_dense(1)[0]
                G = np.tensordot(Ul, Ul, axes=([0, 1], [0, 1]))
            else:
                Vr = r.to_dense(1)[0]
                G = np.tensordot(Vr, Vr, axes=([1, 2], [1, 2]))
            assert np.allclose(G, np.eye(G.shape[0]), atol=1e-12)


def test_dense_svd_truncation_and_split(pkg, ctx):
    say("case dense svd")
    rng = np.random.default_rng(5)
    A = rand_mps([1, 20, 24, 18, 1], 2, rng)
    dev = upload_chain(pkg, ctx, A)
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, 4)
    gL, gR = gpu_envs(pkg, ctx, dev, mpo, 2, 3)
    oL, oR = oracle_envs(A, mpo.H, 2, 3)
    E, AA = lanczos2(pkg, ctx, gL[2], gR[3], dev[1], dev[2], mpo)
    oE, oAA = od.simplelanczos(oL[2], oR[3], A[1], A[2], mpo.H[1], mpo.H[2])
    for m, cut in [(0, 0.0), (10, 0.0), (0, 1e-3), (5, 1e-12)]:
        l, r, te, sv = svd_trunc(pkg, ctx, AA, m, 2, cut, 1)
        U, s, V, ote, _ = od.fsvd(oAA, m=m, minm=2, cutoff=cut)
        say(" m", m, "cut", cut, len(sv), len(s))
        assert len(sv) == len(s)
        assert np.allclose(sv, s, rtol=1e-10, atol=1e-14)
        assert abs(te - ote) <= 1e-12 * max(1.0, ote) + 1e-15
