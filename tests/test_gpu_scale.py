"""The configurations at their stated sizes, checked through properties that
need no oracle run: energies that do not rise from sweep to sweep at cutoff 0,
expect(psi, mpo) equal to the reported energy, orthonormal canonical factors,
a truncation error consistent with the overlap of the truncated state, energy
conservation of real-time TDVP, and the reference's own integration case
(test/dmrg_test.jl:82-85: L=100, m=45, 100 sweeps, dense vs U(1) within 1e-4)."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _ramp(pkg, psi, mpo, params, ramp, m, max_grow=8, **kw):
    for mr in ramp:
        pkg.dmrg(psi, mpo, m=mr, sweeps=1, cutoff=0.0, params=params, **kw)
        h = params.history[-1]
        say(" ramp m=%d E=%.12f maxm=%d %.2fs" % (mr, h["energy"], h["biggestm"], h["time"]))
    for _ in range(max_grow):
        if params.history[-1]["biggestm"] >= int(0.95 * m):
            break
        pkg.dmrg(psi, mpo, m=m, sweeps=1, cutoff=0.0, params=params, **kw)
        h = params.history[-1]
        say(" grow m=%d E=%.12f maxm=%d %.2fs" % (m, h["energy"], h["biggestm"], h["time"]))


def _right_orthonormal(psi, sites, tol=1e-12):
    """After a full sweep the centre is site 1: sites 2..L are right-canonical,
    sum_{s,b} A[a,s,b] A[a',s,b] = delta(a,a')."""
    worst = 0.0
    for i in sites:
        A, _, _ = psi[i].to_dense(psi.nq)
        M = A.reshape(A.shape[0], -1)
        worst = max(worst, float(np.max(np.abs(M @ M.T - np.eye(M.shape[0])))))
    assert worst < tol, worst
    return worst


def _truncation_overlap(pkg, ctx, psi, mpo, params, c, m):
    """At bond (c, c+1): theta = psi[c] psi[c+1] (unit norm) and the truncated
    split at bond dimension m; the normalised truncated state theta' obeys
    <theta|theta'> = sqrt(1 - truncerr) (truncerr = discarded sigma^2)."""
    th = C.c_void_p()
    ctx.call("dmrj_theta2", psi[c].h, psi[c + 1].h, C.byref(th))
    th = pkg.DeviceTensor(ctx, th)
    ctx.call("dmrj_normalize", th.h)
    nl, nr, te, kept = C.c_void_p(), C.c_void_p(), C.c_double(), C.c_int64()
    sv = np.zeros(4 * m)
    ctx.call("dmrj_svd_trunc2", th.h, m, 2, 0.0, 1, C.byref(nl), C.byref(nr), C.byref(te), C.byref(kept),
             sv.ctypes.data_as(C.POINTER(C.c_double)), 4 * m)
    L_, R_ = pkg.DeviceTensor(ctx, nl), pkg.DeviceTensor(ctx, nr)
    t2 = C.c_void_p()
    ctx.call("dmrj_theta2", L_.h, R_.h, C.byref(t2))
    t2 = pkg.DeviceTensor(ctx, t2)
    ov = th.dot(t2)
    nrm = t2.norm()
    for t in (th, t2, L_, R_):
        t.free()
    say(" bond %d|%d truncated to %d: truncerr %.3e, <theta|theta'> %.15f, sqrt(1-truncerr) %.15f, |theta'| %.15f"
        % (c, c + 1, kept.value, te.value, ov, np.sqrt(1 - te.value), nrm))
    assert kept.value == m
    assert te.value > 0.0
    assert abs(nrm - 1.0) < 1e-12
    assert abs(ov - np.sqrt(1.0 - te.value)) < 1e-11
    assert abs(np.sum(sv[:m] ** 2) - 1.0) < 1e-12


def test_cfg2_heisenberg_L100_m1024(pkg, ctx):
    """Heisenberg L=100, U(1), seeded product start, m ramped 128 -> 1024."""
    say("case cfg2 Heisenberg L=100 U(1) m=1024")
    L, m = 100, 1024
    rng = np.random.default_rng(0)
    qphys = np.array([[1], [-1]])
    states = list(rng.permutation([0] * (L // 2) + [1] * (L - L // 2)))
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    _ramp(pkg, psi, mpo, params, [128, 256, 512], m)
    pkg.dmrg(psi, mpo, m=m, sweeps=2, cutoff=0.0, params=params)
    es = [h["energy"] for h in params.history]
    say(" energies", ["%.12f" % e for e in es])
    for a, b in zip(es[3:], es[4:]):        # once the bond dimension is at m
        assert b <= a + 1e-10 * abs(a)
    assert max(psi.bond_dims()) == m
    e = pkg.expect(psi, mpo)
    say(" expect %.15f reported %.15f" % (e, params.energy))
    assert abs(e - params.energy) <= 1e-10 * abs(e)
    _right_orthonormal(psi, [2, 25, 50, 51, 75, 99, 100])
    # truncation check at the centre: move the centre there with FP64 steps
    for i in range(1, L // 2):
        pkg.twostep(0, 1, i, i, i + 1, psi, mpo, params.Lenv, params.Renv, params)
        params.Lenv[i + 1] = pkg.Lupdate(params.Lenv[i], psi[i], mpo, i)
        psi.oc += 1
    _truncation_overlap(pkg, ctx, psi, mpo, params, L // 2, 512)
    psi.free()


def test_cfg3_hubbard_L64_m2048(pkg, ctx):
    """Hubbard L=64, U(1)xU(1), m=2048 (size-class batched eigensolves over
    sectors of a few hundred rows, LQ/QR-free augmented path, completion of
    zero singular values at cutoff 0); one step at the centre bond from the
    same state with one eigensolve batch and with the size classes gives the
    same energy and kept spectrum."""
    say("case cfg3 Hubbard L=64 U(1)xU(1) m=2048")
    L, m = 64, 2048
    rng = np.random.default_rng(0)
    qphys = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    states = list(rng.permutation([1] * (L // 2) + [2] * (L - L // 2)))
    mpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=qphys)
    psi = pkg.product_state(states, 4, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    ctx.set_svd_backend(5)          # the ramp on the Gram backend (fast), the checked sweep in FP64
    try:
        _ramp(pkg, psi, mpo, params, [256, 512, 1024], m)
    finally:
        ctx.set_svd_backend(0)
    t0 = time.perf_counter()
    pkg.dmrg(psi, mpo, m=m, sweeps=1, cutoff=0.0, params=params)
    say(" FP64 sweep E=%.12f maxm=%d %.2fs" % (params.energy, params.history[-1]["biggestm"],
                                               time.perf_counter() - t0))
    es = [h["energy"] for h in params.history]
    assert es[-1] <= es[-2] + 1e-10 * abs(es[-2])
    assert params.history[-1]["biggestm"] == m
    e = pkg.expect(psi, mpo)
    say(" expect %.15f reported %.15f" % (e, params.energy))
    assert abs(e - params.energy) <= 1e-10 * abs(e)
    _right_orthonormal(psi, [2, 16, 32, 33, 48, 63, 64])
    # one bond step from the same state (centre moved to L/2) with the two batchings
    for i in range(1, L // 2):
        pkg.twostep(0, 1, i, i, i + 1, psi, mpo, params.Lenv, params.Renv, params)
        params.Lenv[i + 1] = pkg.Lupdate(params.Lenv[i], psi[i], mpo, i)
        psi.oc += 1
    c = L // 2
    res = {}
    for mode in (0, 1):
        ctx.set_eig_batching(mode)
        st = psi.copy()
        p2 = pkg.algvars(st, maxm=m, cutoff=0.0)
        p2.SvNbond = 0
        sv = np.zeros(4 * m)
        pkg.twostep(0, 1, c, c, c + 1, st, mpo, params.Lenv, params.Renv, p2)
        A, _, _ = st[c + 1].to_dense(2)
        res[mode] = (p2.energy, p2.truncerr, np.sort(np.linalg.norm(A.reshape(A.shape[0], -1), axis=1))[::-1])
        st.free()
    ctx.set_eig_batching(1)
    say(" centre step: E one batch %.15f size classes %.15f; truncerr %.3e %.3e"
        % (res[0][0], res[1][0], res[0][1], res[1][1]))
    assert res[0][0] == res[1][0]                       # the Lanczos step precedes the SVD
    assert abs(res[0][1] - res[1][1]) < 1e-12
    assert np.max(np.abs(res[0][2] - res[1][2])) < 1e-10   # kept sigma (rows of D*V)
    psi.free()


def test_cfg4_j1j2_cylinder_m4096_autompo(pkg, ctx):
    """J1-J2 cylinder 4 x 32 (L=128), J2=0.5, U(1), the D=14 MPO of autoMPO,
    one sweep at m=4096 in the precision form (FP16x3 H_eff*psi, FP64
    environments and correction, Gram SVD)."""
    say("case cfg4 J1-J2 cylinder 32x4 m=4096 (autoMPO D=14)")
    import bench
    L, m, Ly = 128, 4096, 4
    Ws = bench.j1j2_autompo(pkg, L // Ly, Ly)
    D = max(W.shape[3] for W in Ws)
    say(" autoMPO link dimension %d" % D)
    assert D == 14
    qphys = np.array([[1], [-1]])
    rng = np.random.default_rng(0)
    states = list(rng.permutation([0] * (L // 2) + [1] * (L - L // 2)))
    mpo = pkg.MPO(Ws, qphys=qphys)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    kw = dict(precision="f16x3", env_precision="f64", correction="sweep")
    ctx.set_svd_backend(5)
    try:
        _ramp(pkg, psi, mpo, params, [256, 1024], m, **kw)
        maxtrunc = params.history[-1]["maxtrunc"]
    finally:
        ctx.set_svd_backend(0)
    assert params.history[-1]["biggestm"] >= int(0.95 * m)
    e = pkg.expect(psi, mpo)
    say(" expect %.15f reported (FP64-corrected) %.15f, maxtrunc %.2e" % (e, params.energy, maxtrunc))
    # the reported energy is the FP64 Rayleigh quotient before the last truncation
    assert abs(e - params.energy) <= 1e-10 * abs(e) + 2.0 * maxtrunc * abs(e)
    _right_orthonormal(psi, [2, 64, 65, 127, 128], tol=1e-11)
    psi.free()


def test_cfg5_tdvp_L100_neel_200_steps(pkg, ctx):
    """Real-time two-site TDVP of the Heisenberg chain L=100 from the Neel
    state, dt=0.05, 200 calls, Krylov 12, m <= 1024, U(1): the energy of the
    evolved state stays that of the Neel state, the state stays normalised,
    and total Sz stays zero."""
    say("case cfg5 TDVP L=100 Neel 200 calls")
    L = 100
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    e0 = pkg.expect(psi, mpo)
    assert abs(e0 - (-(L - 1) / 4.0)) < 1e-12
    Le = Re = None
    t0 = time.perf_counter()
    for k in range(200):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=1024, cutoff=1e-12, r=12, prefactor=-0.05j, Lenv=Le, Renv=Re)
        if k % 50 == 49:
            say(" call %d: %.1fs, max bond %d" % (k + 1, time.perf_counter() - t0, max(psi.bond_dims())))
    e = pkg.expect(psi, mpo)
    Sz = np.diag([0.5, -0.5])
    szc = pkg.local_expect(psi, Sz, L // 2)
    Id = np.eye(2)
    nrm = pkg.local_expect(psi, Id, 1)
    say(" E %.15f (Neel %.15f), <Sz_%d> %.12f" % (e, e0, L // 2, szc))
    assert abs(e - e0) <= 1e-9 * abs(e0)
    assert abs(nrm - 1.0) < 1e-12
    assert szc != 0.5 and abs(szc) <= 0.5
    psi.free()


@pytest.mark.parametrize("method", ["twosite", "3S"])
def test_reference_integration_L100_m45(pkg, ctx, method):
    """test/dmrg_test.jl:82-85 at its size: Heisenberg L=100 from the Neel
    state, m=45, 100 sweeps, cutoff 1e-12, dense and U(1) energies within
    1e-4 of each other."""
    say("case reference integration L=100 m=45 100 sweeps %s" % method)
    L = 100
    qphys = np.array([[1], [-1]])
    energies = []
    for q in (None, qphys):
        mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=q)
        psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=q, ctx=ctx)
        t0 = time.perf_counter()
        E = pkg.dmrg(psi, mpo, m=45, sweeps=100, cutoff=1e-12, method=method)
        say(" %s E=%.12f %.1fs" % ("U(1)" if q is not None else "dense", E, time.perf_counter() - t0))
        energies.append(E)
        psi.free()
    assert abs(energies[0] - energies[1]) < 1e-4
