"""ComplexF64 tensors on the GPU path (real-time TDVP, algs/time.jl:288-362)
against the oracle, which runs the same contractions in numpy complex128.

Tolerances: single contractions and the Krylov propagator <= 1e-12 relative
on identical inputs; singular values <= 1e-12 relative; whole real-time
half-sweeps: local <Sz_i> <= 1e-10 absolute and energies <= 1e-10
relative (truncated SVDs in between, gauge-free quantities).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models as om
from oracle import mps as omps
from oracle import tdvp as ot

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_complex_upload_roundtrip_and_dots(pkg, ctx):
    say("case complex roundtrip")
    rng = np.random.default_rng(1)
    A = crand(rng, (3, 2, 5))
    B = crand(rng, (3, 2, 5))
    a = pkg.site_from_dense(ctx, A)
    b = pkg.site_from_dense(ctx, B)
    assert a.is_complex
    assert np.array_equal(a.to_dense()[0], A)
    assert abs(a.norm() - np.linalg.norm(A)) <= 1e-13 * np.linalg.norm(A)
    z = a.dotc(b)
    want = np.vdot(A, B)
    assert abs(z - want) <= 1e-13 * abs(want)
    assert abs(a.dot(b) - want.real) <= 1e-13 * abs(want)
    r = pkg.site_from_dense(ctx, A.real)
    assert not r.is_complex
    assert abs(r.dotc(b) - np.vdot(A.real, B)) <= 1e-13 * abs(want)
    assert abs(b.dotc(r) - np.vdot(B, A.real)) <= 1e-13 * abs(want)
    r.to_complex()
    assert r.is_complex and np.array_equal(r.to_dense()[0], A.real.astype(complex))


@pytest.mark.parametrize("mixed", [False, True])
def test_complex_envs_heff_theta(pkg, ctx, mixed):
    """Lupdate / Rupdate / theta / two-site and one-site H_eff on complex
    dense tensors; with mixed=True the environments come from real tensors
    (real operands have no imaginary plane)."""
    say("case complex contractions mixed=%s" % mixed)
    rng = np.random.default_rng(5)
    dims = [1, 4, 9, 6, 1]
    L = 4
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    A = [crand(rng, (dims[i], 2, dims[i + 1])) for i in range(L)]
    Aenv = [a.real.copy() for a in A] if mixed else A
    dev = [pkg.site_from_dense(ctx, a) for a in A]
    denv = [pkg.site_from_dense(ctx, a) for a in Aenv]
    gL = pkg.Lupdate(pkg.network.makeBoundary(ctx, True), denv[0], mpo, 1)
    gR = pkg.Rupdate(pkg.network.makeBoundary(ctx, False), denv[3], mpo, 4)
    oL = od.Lupdate(np.ones((1, 1, 1)), Aenv[0], mpo.H[0])
    oR = od.Rupdate(np.ones((1, 1, 1)), Aenv[3], mpo.H[3])
    assert rel(gL.to_dense()[0], oL) <= 1e-12
    assert rel(gR.to_dense()[0], oR) <= 1e-12
    gL2 = pkg.Lupdate(gL, dev[1], mpo, 2)
    oL2 = od.Lupdate(oL, A[1], mpo.H[1])
    gR2 = pkg.Rupdate(gR, dev[2], mpo, 3)
    oR2 = od.Rupdate(oR, A[2], mpo.H[2])
    say(" Lupdate err %.2e Rupdate err %.2e" % (rel(gL2.to_dense()[0], oL2), rel(gR2.to_dense()[0], oR2)))
    assert rel(gL2.to_dense()[0], oL2) <= 1e-12
    assert rel(gR2.to_dense()[0], oR2) <= 1e-12
    h = C.c_void_p()
    ctx.call("dmrj_theta2", dev[1].h, dev[2].h, C.byref(h))
    th = pkg.DeviceTensor(ctx, h)
    oth = np.tensordot(A[1], A[2], axes=(2, 0))
    assert rel(th.to_dense()[0], oth) <= 1e-12
    h = C.c_void_p()
    ctx.call("dmrj_heff2_apply", gL.h, mpo.device(ctx, 2), mpo.device(ctx, 3), gR.h, th.h, 0, C.byref(h))
    y = pkg.DeviceTensor(ctx, h).to_dense()[0]
    want = od.heff2(oL, oR, oth, mpo.H[1], mpo.H[2])
    say(" heff2 err %.2e" % rel(y, want))
    assert rel(y, want) <= 1e-12
    h = C.c_void_p()
    ctx.call("dmrj_heff1_apply", gL.h, mpo.device(ctx, 2), gR2.h, dev[1].h, 0, C.byref(h))
    y1 = pkg.DeviceTensor(ctx, h).to_dense()[0]
    want1 = od.singlesite_update(oL, oR2, A[1], mpo.H[1])
    assert rel(y1, want1) <= 1e-12
    # H_eff is Hermitian: <x|H y> = conj <y|H x>
    lhs = np.vdot(oth, od.heff2(oL, oR, want, mpo.H[1], mpo.H[2]))
    assert abs(lhs.imag) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("batch", [1, 2, 0])
def test_complex_svd_truncation(pkg, ctx, batch):
    """Complex two-site SVD with charges: kept singular values against numpy
    per sector, left factor orthonormal, product equal to the truncated
    tensor; the sectors' Hermitian eigensolves batched by the cost model,
    one per sector, or all in one."""
    say("case complex SVD, eigensolve batching", batch)
    rng = np.random.default_rng(9)
    q = lambda n, c: np.array([[c]] * n)
    qL = np.vstack([q(3, -1), q(5, 1)])
    qP = np.array([[1], [-1]])
    qR = np.vstack([q(4, -1), q(6, 1)])
    mL, d, mR = 8, 2, 10
    AA = np.zeros((mL, d, d, mR), dtype=complex)
    for a in range(mL):
        for s1 in range(d):
            for s2 in range(d):
                for b in range(mR):
                    if qL[a, 0] + qP[s1, 0] + qP[s2, 0] == qR[b, 0]:
                        AA[a, s1, s2, b] = rng.standard_normal() + 1j * rng.standard_normal()
    dev = pkg.site2_from_dense(ctx, AA, qL, qP, qP, qR, nq=1)
    for m in (0, 5):
        l, r = C.c_void_p(), C.c_void_p()
        te = C.c_double()
        k = C.c_int64()
        sv = np.zeros(64)
        ctx.set_eig_batching(batch)
        try:
            ctx.call("dmrj_svd_trunc2", dev.h, m, 2, 0.0, 1, C.byref(l), C.byref(r), C.byref(te), C.byref(k),
                     pkg._lib.dptr(sv), 64)
        finally:
            ctx.set_eig_batching(1)
        U, s, V, terr, qn = od.fsvd(AA, m, 2, 0.0, qL, qP, qR)
        kk = k.value
        say(" m", m, "kept", kk, "truncerr %.3e %.3e" % (te.value, terr))
        assert kk == len(s)
        # the C ABI returns the kept values sorted; the oracle keeps sector order
        assert np.allclose(np.sort(sv[:kk])[::-1], np.sort(s)[::-1], rtol=1e-12, atol=1e-14)
        assert abs(te.value - terr) <= 1e-12 * max(terr, 1e-300) + 1e-15
        gl = pkg.DeviceTensor(ctx, l)
        gr = pkg.DeviceTensor(ctx, r)
        Ul = gl.to_dense(1)[0].reshape(mL * d, kk, order="F")
        assert np.allclose(Ul.conj().T @ Ul, np.eye(kk), atol=1e-12)
        prod = np.tensordot(gl.to_dense(1)[0], gr.to_dense(1)[0], axes=(2, 0))
        want = np.tensordot(U, s[:, None, None] * V, axes=(2, 0))
        assert rel(prod, want) <= 1e-11


def test_complex_krylov_propagator(pkg, ctx):
    say("case complex krylov")
    rng = np.random.default_rng(11)
    dims = [1, 5, 10, 5, 1]
    L = 4
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    A = [crand(rng, (dims[i], 2, dims[i + 1])) for i in range(L)]
    dev = [pkg.site_from_dense(ctx, a) for a in A]
    gL = pkg.Lupdate(pkg.network.makeBoundary(ctx, True), dev[0], mpo, 1)
    gR = pkg.Rupdate(pkg.network.makeBoundary(ctx, False), dev[3], mpo, 4)
    oL = od.Lupdate(np.ones((1, 1, 1)), A[0], mpo.H[0])
    oR = od.Rupdate(np.ones((1, 1, 1)), A[3], mpo.H[3])
    h = C.c_void_p()
    ctx.call("dmrj_theta2", dev[1].h, dev[2].h, C.byref(h))
    th = pkg.DeviceTensor(ctx, h)
    oth = np.tensordot(A[1], A[2], axes=(2, 0))
    for r, pf in ((12, -0.05j), (6, -0.3j), (4, -0.1 - 0.2j)):
        out, e = pkg.krylovTimeEvol2(gL, gR, th, mpo, 2, r, pf)
        want, we = ot.krylov_expm(lambda t: od.heff2(oL, oR, t, mpo.H[1], mpo.H[2]), oth, r, pf)
        err = rel(out.to_dense()[0], want)
        say(" r", r, "pf", pf, "err %.2e" % err, "E", e, we)
        assert err <= 1e-12
        assert abs(e - we) <= 1e-12 * abs(we)
    # real start vector, real environments, imaginary prefactor
    R = [a.real.copy() for a in A]
    rdev = [pkg.site_from_dense(ctx, a) for a in R]
    rL = pkg.Lupdate(pkg.network.makeBoundary(ctx, True), rdev[0], mpo, 1)
    rR = pkg.Rupdate(pkg.network.makeBoundary(ctx, False), rdev[3], mpo, 4)
    oL = od.Lupdate(np.ones((1, 1, 1)), R[0], mpo.H[0])
    oR = od.Rupdate(np.ones((1, 1, 1)), R[3], mpo.H[3])
    h = C.c_void_p()
    ctx.call("dmrj_theta2", rdev[1].h, rdev[2].h, C.byref(h))
    rth = pkg.DeviceTensor(ctx, h)
    out, e = pkg.krylovTimeEvol2(rL, rR, rth, mpo, 2, 8, -0.2j)
    assert out.is_complex
    want, _ = ot.krylov_expm(lambda t: od.heff2(oL, oR, t, mpo.H[1], mpo.H[2]),
                             np.tensordot(R[1], R[2], axes=(2, 0)), 8, -0.2j)
    assert rel(out.to_dense()[0], want) <= 1e-12


def test_realtime_tdvp_two_sites_closed_form(pkg, ctx):
    """<Sz_1>(t) = cos(t)/2 for two spins from |up down> (one call is the
    exact propagator on L=2 with the reference's step conventions)."""
    say("case real-time TDVP L=2 closed form")
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, 2).H, qphys=qphys)
    psi = pkg.product_state([0, 1], 2, qphys=qphys, ctx=ctx)
    Sz = np.diag([0.5, -0.5])
    Le = Re = None
    dt = 0.3
    for n in range(1, 5):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=4, cutoff=0.0, r=12, prefactor=-1j * dt, Lenv=Le, Renv=Re)
        sz1 = pkg.local_expect(psi, Sz, 1)
        say(" n", n, "Sz1 %.15f" % sz1, "want %.15f" % (0.5 * np.cos(n * dt)))
        assert abs(sz1 - 0.5 * np.cos(n * dt)) < 1e-12


def test_realtime_tdvp_u1_neel_vs_oracle(pkg, ctx):
    """Heisenberg L=10 from the Neel state, U(1), real time dt = 0.05 per
    half-sweep, m=16, cutoff 1e-12, Krylov 12: local <Sz_i>(t) and the energy
    against the oracle for eight half-sweeps."""
    say("case real-time TDVP U(1) L=10")
    L = 10
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    ompo = om.makeMPO(om.heisenbergMPO(), 2, L)
    Sz = np.diag([0.5, -0.5])
    gLe = gRe = oLe = oRe = None
    dt = 0.05
    for hs in range(8):
        E, gLe, gRe = pkg.tdvp_twosite(psi, mpo, m=16, cutoff=1e-12, r=12, prefactor=-1j * dt,
                                       Lenv=gLe, Renv=gRe)
        oE, oLe, oRe = ot.tdvp_twosite(opsi, ompo, m=16, cutoff=1e-12, r=12, prefactor=-1j * dt,
                                       Lenv=oLe, Renv=oRe)
        sz = np.array([pkg.local_expect(psi, Sz, i) for i in range(1, L + 1)])
        osz = np.array([ot.local_expect(opsi, Sz, i) for i in range(1, L + 1)])
        Eg = pkg.expect(psi, mpo)
        say(" half-sweep", hs, "Sz1 %.12f %.12f" % (sz[0], osz[0]), "max dSz %.2e" % np.max(np.abs(sz - osz)),
            "E %.12f" % Eg)
        assert np.max(np.abs(sz - osz)) <= 1e-10
        assert abs(E - oE) <= 1e-10 * max(abs(oE), 1.0)
    assert psi[5].is_complex


def test_realtime_tdvp_graph_matches_direct(pkg, ctx):
    """With graphs on, every Krylov exponential's iteration is captured and
    launched as one graph: the half-sweeps reproduce the direct launches bit
    for bit (tensors and <Sz_i>), and the executable graphs are updated in
    place from bond to bond."""
    say("case real-time TDVP graph vs direct")
    L = 10
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    Sz = np.diag([0.5, -0.5])
    runs = []
    for graph in (0, 1):
        ctx.set_graph(graph)
        try:
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            Le = Re = None
            for _ in range(4):
                _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=16, cutoff=1e-12, r=12, prefactor=-0.05j, Lenv=Le, Renv=Re)
            sz = np.array([pkg.local_expect(psi, Sz, i) for i in range(1, L + 1)])
            dense = [t.to_dense(psi.nq)[0] for t in psi.A]
            stats = ctx.graph_stats()
        finally:
            ctx.set_graph(0)
        runs.append((sz, dense, stats))
    assert np.array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)
    say(" graph stats", runs[1][2])
    assert runs[1][2]["updates"] > 0
