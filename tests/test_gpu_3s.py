"""3S (single-site DMRG with subspace expansion) on the GPU path against the
oracle (oracle/dmrg3s.py), through the C ABI.

Tolerances: FP64, 1e-12 relative for the Ritz pair, the expansion (compared
through its Gram matrix over the enlarged bond, which does not depend on how
that bond is ordered) and the two-site state after a step (gauge-invariant
product of the two new tensors); per-sweep energies of whole runs 1e-12
relative; dense against quantum-number runs within the reference's 1e-4
(test/dmrg_test.jl:82-85).
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import dmrg3s as o3
from oracle import models as om
from oracle import mps as omps

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(got, want):
    return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300))


def _oracle_state(L, m, oc, qn):
    """Heisenberg state after one two-site sweep, centre carried to oc by
    two-site steps (bond bases sorted by charge, as the device keeps them)."""
    mpo_h = om.makeMPO(om.heisenbergMPO(), 2, L)
    qphys = np.array([[1], [-1]]) if qn else None
    psi = omps.product_mps([i % 2 for i in range(L)], 2, qphys=qphys)
    od.dmrg(psi, mpo_h, m=m, sweeps=1, cutoff=0.0)
    params = od.Params(L, m=m, cutoff=0.0)
    Lenv, Renv = od.makeEnv(psi, mpo_h)
    while psi.oc < oc:
        i = psi.oc
        od.twostep(1, i, i, i + 1, psi, mpo_h, Lenv, Renv, params)
        Lenv[i + 1] = od.Lupdate(Lenv[i], psi[i], mpo_h[i - 1])
        psi.oc += 1
    Lenv, Renv = od.makeEnv(psi, mpo_h)
    return psi, mpo_h, Lenv, Renv


def _upload(pkg, ctx, opsi, mpo_h):
    qn = opsi.qn
    psi = pkg.MPS([a.copy() for a in opsi.A], oc=opsi.oc, qphys=opsi.qphys, qbond=opsi.qbond if qn else None,
                  ctx=ctx, canonical=True)
    mpo = pkg.MPO(mpo_h, qphys=opsi.qphys) if qn else pkg.MPO(mpo_h)
    Lenv, Renv = pkg.makeEnv(psi, mpo)
    return psi, mpo, Lenv, Renv


@pytest.mark.parametrize("qn", [False, True], ids=["dense", "u1"])
def test_lanczos1_and_expansion_match_oracle(pkg, ctx, qn):
    say("case 3S lanczos1/expand3s", "u1" if qn else "dense")
    i = 4
    opsi, mpo_h, oL, oR = _oracle_state(10, 8, i, qn)
    psi, mpo, L, R = _upload(pkg, ctx, opsi, mpo_h)
    W = mpo_h[i - 1]
    A, e = pkg.lanczos1(L[i], R[i], psi[i], mpo, i, 2)
    oe, oA = o3.lanczos_ritz(lambda t: od.singlesite_update(oL[i], oR[i], t, W), opsi[i], 2)
    gA = A.to_dense(psi.nq)[0]
    say(" energy", e, oe, "vector rel", relerr(gA, oA))
    assert abs(e - oe) <= TOL * abs(oe)
    assert relerr(gA, oA) <= TOL
    for direction in (1, -1):
        E = L[i] if direction > 0 else R[i]
        T = pkg.expand3s(E, A, mpo, i, 0.7, direction)
        gT = T.to_dense(psi.nq)[0]        # two-site layout [a, d, 1, b'] / [a', 1, d, b]
        if direction > 0:
            oT = np.concatenate([oA, o3.Lexpand(oA, W, oL[i], 0.7)], axis=2)
            gM = gT.reshape(-1, gT.shape[-1], order="F")
            oM = oT.reshape(-1, oT.shape[2], order="F")
            gG, oG = gM @ gM.T, oM @ oM.T
        else:
            oT = np.concatenate([oA, o3.Rexpand(oA, W, oR[i], 0.7)], axis=0)
            gM = gT.reshape(gT.shape[0], -1, order="F")
            oM = oT.reshape(oT.shape[0], -1, order="F")
            gG, oG = gM.T @ gM, oM.T @ oM
        say(" direction", direction, "enlarged bond", gT.shape, "oracle", oT.shape, "gram rel", relerr(gG, oG))
        assert relerr(gG, oG) <= TOL
        assert np.linalg.matrix_rank(gM, tol=1e-10 * np.abs(gM).max()) == \
            np.linalg.matrix_rank(oM, tol=1e-10 * np.abs(oM).max())
        T.free()
        # alpha = 0: A itself in the two-site layout
        T0 = pkg.expand3s(E, A, mpo, i, 0.0, direction)
        assert np.array_equal(T0.to_dense(psi.nq)[0].reshape(gA.shape, order="F"), gA)
        T0.free()
    A.free()


@pytest.mark.parametrize("qn", [False, True], ids=["dense", "u1"])
def test_lanczos_k_matches_oracle(pkg, ctx, qn):
    """dmrj_lanczos_k: two-site r = 2 equals simplelanczos (algs/DMRG.jl:57-90),
    two-site r = 10 and one-site r = 6 equal the oracle's Ritz pairs."""
    say("case lanczos_k", "u1" if qn else "dense")
    i = 4
    opsi, mpo_h, oL, oR = _oracle_state(10, 8, i, qn)
    psi, mpo, L, R = _upload(pkg, ctx, opsi, mpo_h)
    th = pkg.tdvp._theta(ctx, psi[i], psi[i + 1])
    W1, W2 = mpo_h[i - 1], mpo_h[i]
    oth = np.tensordot(opsi[i], opsi[i + 1], axes=(2, 0))
    e2, _ = od.simplelanczos(oL[i], oR[i + 1], opsi[i], opsi[i + 1], W1, W2)
    for r, want in ((2, e2), (10, None)):
        out, e, k = pkg.lanczos_k(L[i], R[i + 1], th, mpo, i, r)
        oe, ov = o3.lanczos_ritz(lambda t: od.heff2(oL[i], oR[i + 1], t, W1, W2), oth, r)
        gv = out.to_dense(psi.nq)[0].reshape(ov.shape, order="F")
        say(" two-site r", r, "k", k, "energy", e, oe, "simplelanczos", want, "vector rel", relerr(gv, ov))
        assert abs(e - oe) <= TOL * abs(oe)
        if want is not None:
            assert abs(e - want) <= TOL * abs(want)
        assert relerr(gv, ov) <= 1e-10
        out.free()
    out, e, k = pkg.lanczos_k(L[i], R[i], psi[i], mpo, i, 6)
    oe, ov = o3.lanczos_ritz(lambda t: od.singlesite_update(oL[i], oR[i], t, W1), opsi[i], 6)
    say(" one-site r 6 energy", e, oe)
    assert abs(e - oe) <= TOL * abs(oe)
    assert relerr(out.to_dense(psi.nq)[0], ov) <= 1e-10
    out.free()
    th.free()


@pytest.mark.parametrize("qn", [False, True], ids=["dense", "u1"])
@pytest.mark.parametrize("j", [1, -1], ids=["right", "left"])
def test_step3s_matches_oracle(pkg, ctx, qn, j):
    """One step3S: energy, the two-site state after the move (gauge-invariant
    product), truncation error, kept dimension and the updated noise."""
    say("case 3S step", "u1" if qn else "dense", "j", j)
    i, L, m = 5, 12, 6
    opsi, mpo_h, oL, oR = _oracle_state(L, 10, i, qn)
    psi, mpo, Lg, Rg = _upload(pkg, ctx, opsi, mpo_h)
    iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
    op = o3.Params3S(L, m=m, cutoff=1e-10, noise=0.6)
    op.lastenergy = -3.0
    gp = pkg.algvars(psi, maxm=m, cutoff=1e-10)
    gp.noise = 0.6
    gp.lastenergy = -3.0
    deltas = omps.mpo_deltas(mpo_h, opsi.qphys) if qn else None
    o3.step3S(j, i, iL, iR, opsi, mpo_h, oL, oR, op, deltas)
    pkg.step3S(1, j, i, iL, iR, psi, mpo, Lg, Rg, gp)
    gth = np.tensordot(psi[iL].to_dense(psi.nq)[0], psi[iR].to_dense(psi.nq)[0], axes=(2, 0))
    oth = np.tensordot(opsi[iL], opsi[iR], axes=(2, 0))
    say(" energy", gp.energy, op.energy, "theta rel", relerr(gth, oth), "truncerr", gp.truncerr, op.truncerr,
        "m", gp.biggestm, op.biggestm, "noise", gp.noise, op.noise)
    assert abs(gp.energy - op.energy) <= TOL * abs(op.energy)
    assert relerr(gth, oth) <= 1e-11
    assert abs(gp.truncerr - op.truncerr) <= 1e-10 * max(op.truncerr, 1e-14)
    assert gp.biggestm == op.biggestm
    assert gp.noise == pytest.approx(op.noise, rel=1e-15)
    # the moved-to site's left (right) factor is orthonormal
    if j > 0:
        X = psi[iL].to_dense(psi.nq)[0]
        X = X.reshape(-1, X.shape[2], order="F")
        assert np.allclose(X.T @ X, np.eye(X.shape[1]), atol=1e-12)
    else:
        X = psi[iR].to_dense(psi.nq)[0]
        X = X.reshape(X.shape[0], -1, order="F")
        assert np.allclose(X @ X.T, np.eye(X.shape[0]), atol=1e-12)


@pytest.mark.parametrize("qn", [False, True], ids=["dense", "u1"])
def test_dmrg3s_sweeps_match_oracle(pkg, ctx, qn):
    """Whole 3S runs from the Neel state (truncating: m = 12 on L = 14):
    per-sweep energies, kept dimensions, truncation and noise."""
    say("case 3S sweeps L=14 m=12", "u1" if qn else "dense")
    L = 14
    states = [i % 2 for i in range(L)]
    qphys = np.array([[1], [-1]]) if qn else None
    mpo_h = om.makeMPO(om.heisenbergMPO(), 2, L)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    oE, ohist = o3.dmrg3S(opsi, mpo_h, m=12, sweeps=5, cutoff=1e-12)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    mpo = pkg.MPO(mpo_h, qphys=qphys) if qn else pkg.MPO(mpo_h)
    params = pkg.algvars(psi)
    E = pkg.dmrg(psi, mpo, m=12, sweeps=5, cutoff=1e-12, method="3S", params=params)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel,
            "m", h["biggestm"], oh["biggestm"], "noise %.6f %.6f" % (h["noise"], oh["noise"]),
            "t=%.3fs" % h["time"])
        assert rel <= TOL
        assert h["biggestm"] == oh["biggestm"]
        assert h["noise"] == pytest.approx(oh["noise"], rel=1e-12)
        assert np.allclose(h["storeD"], oh["storeD"], rtol=1e-8, atol=1e-12)
    assert len(params.history) == len(ohist)
    assert abs(E - oE) <= TOL * abs(oE)


@pytest.mark.parametrize("batch", [1, 2], ids=["classes", "per_sector"])
@pytest.mark.parametrize("L,m,n", [(24, 32, 6), (20, 160, 4)], ids=["L24m32", "L20m160"])
def test_dmrg3s_u1_longer_runs_match_oracle(pkg, ctx, L, m, n, batch):
    """Longer 3S runs with many small singular values kept (cutoff 1e-12):
    per-sweep energies against the oracle.  At m = 160 the enlarged sectors
    exceed 64 on both sides, so the large-sector SVD path runs (its LQ/QR
    form), with the eigensolves batched by the cost model (one batch at these
    sizes) or one batch per sector."""
    say("case 3S sweeps U(1) L=%d m=%d eigensolve batching %d" % (L, m, batch))
    states = [i % 2 for i in range(L)]
    qphys = np.array([[1], [-1]])
    mpo_h = om.makeMPO(om.heisenbergMPO(), 2, L)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    oE, ohist = o3.dmrg3S(opsi, mpo_h, m=m, sweeps=n, cutoff=1e-12)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    ctx.set_eig_batching(batch)
    try:
        pkg.dmrg(psi, pkg.MPO(mpo_h, qphys=qphys), m=m, sweeps=n, cutoff=1e-12, method="3S", params=params)
    finally:
        ctx.set_eig_batching(1)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel,
            "m", h["biggestm"], oh["biggestm"], "maxtrunc %.2e %.2e" % (h["maxtrunc"], oh["maxtrunc"]),
            "noise %.6f %.6f" % (h["noise"], oh["noise"]), "t=%.3fs" % h["time"])
        assert rel <= TOL
        assert h["biggestm"] == oh["biggestm"]
    assert len(params.history) == len(ohist)


def test_dmrg3s_hubbard_u1xu1_matches_oracle(pkg, ctx):
    """U(1) x U(1) Hubbard L=8 at half filling (two charges, fermion signs in
    the state), m = 24: per-sweep energies and noise against the oracle."""
    say("case 3S Hubbard U(1)xU(1) L=8 m=24")
    L = 8
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    qh = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    hmpo = om.makeMPO(om.hubbardMPO(), 4, L)
    o = omps.product_mps([0] * L, 4)
    omps.apply_ops(o, list(range(1, 5)), Cup.T, trail=F)
    omps.apply_ops(o, list(range(1, 5)), Cdn.T, trail=F)
    omps.product_charges(o, qh)
    psi = pkg.product_state([0] * L, 4, qphys=qh, ctx=ctx)          # the product's own applyOps!
    pkg.applyOps(psi, list(range(1, 5)), Cup.T, trail=F)
    pkg.applyOps(psi, list(range(1, 5)), Cdn.T, trail=F)
    assert all(np.array_equal(a, b) for a, b in zip(psi.to_dense()[0], o.A))
    oE, ohist = o3.dmrg3S(o, hmpo, m=24, sweeps=4, cutoff=1e-12)
    params = pkg.algvars(psi)
    pkg.dmrg(psi, pkg.MPO(hmpo, qphys=qh), m=24, sweeps=4, cutoff=1e-12, method="3S", params=params)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel,
            "m", h["biggestm"], oh["biggestm"], "noise %.6f %.6f" % (h["noise"], oh["noise"]))
        assert rel <= TOL
        assert h["biggestm"] == oh["biggestm"]
        assert h["noise"] == pytest.approx(oh["noise"], rel=1e-12)
    assert len(params.history) == len(ohist)


def test_dmrg3s_resumed_from_checkpoint(pkg, ctx, tmp_path):
    """3S for 2 sweeps, checkpoint (the noise and the last energy go into the
    archive's extra data), load, 2 more sweeps == 4 sweeps in one run."""
    say("case 3S resume")
    L = 16
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(om.makeMPO(om.heisenbergMPO(), 2, L), qphys=qphys)
    states = [i % 2 for i in range(L)]
    full = pkg.algvars(pkg.product_state(states, 2, qphys=qphys, ctx=ctx))
    pkg.dmrg(pkg.product_state(states, 2, qphys=qphys, ctx=ctx), mpo, m=24, sweeps=4, cutoff=1e-12,
             method="3S", params=full)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    first = pkg.algvars(psi)
    pkg.dmrg(psi, mpo, m=24, sweeps=2, cutoff=1e-12, method="3S", params=first)
    p = str(tmp_path / "r3s.npz")
    pkg.save_mps(psi, p, extra={"noise": first.noise, "energy": first.energy})
    back, extra = pkg.load_mps(p, ctx)
    second = pkg.algvars(back)
    second.noise = extra["noise"]
    second.energy = extra["energy"]
    pkg.dmrg(back, mpo, m=24, sweeps=2, cutoff=1e-12, method="3S", params=second)
    got = [h["energy"] for h in first.history] + [h["energy"] for h in second.history]
    want = [h["energy"] for h in full.history]
    say(" energies", ["%.15f" % e for e in got], "uninterrupted", ["%.15f" % e for e in want])
    assert got == want


def test_dmrg3s_dense_vs_qn_reference_criteria(pkg, ctx):
    """test/dmrg_test.jl:7,82-85 with method 3S: spins (Heisenberg) and
    fermions (Hubbard U(1) x U(1), half filling): dense and quantum-number
    runs agree within 1e-4; the Hubbard L=10 energy is within 1e-4 of the
    reference's m=45 value -25.380567747816 (test/model_test.jl:362-364)."""
    say("case 3S dense vs QN")
    L = 32
    states = [i % 2 for i in range(L)]
    mpo_h = om.makeMPO(om.heisenbergMPO(), 2, L)
    qphys = np.array([[1], [-1]])
    Ed = pkg.dmrg(pkg.product_state(states, 2, ctx=ctx), pkg.MPO(mpo_h), m=45, sweeps=12, cutoff=1e-12,
                  method="3S")
    pq = pkg.algvars(pkg.product_state(states, 2, qphys=qphys, ctx=ctx))
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    Eq = pkg.dmrg(psi, pkg.MPO(mpo_h, qphys=qphys), m=45, sweeps=12, cutoff=1e-12, method="3S", params=pq)
    E2 = pkg.dmrg(pkg.product_state(states, 2, qphys=qphys, ctx=ctx), pkg.MPO(mpo_h, qphys=qphys), m=45,
                  sweeps=6, cutoff=1e-12)
    say(" Heisenberg L=32: dense", Ed, "U(1)", Eq, "two-site", E2, "m", [h["biggestm"] for h in pq.history])
    assert abs(Ed - Eq) < 1e-4
    assert abs(Eq - E2) < 1e-4
    # Hubbard
    L = 10
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    qh = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    hmpo = om.makeMPO(om.hubbardMPO(), 4, L)
    out = []
    for q in (None, qh):
        # the Hubbard filling of test/dmrg_test.jl:67-75 by the product's applyOps!
        hpsi = pkg.product_state([0] * L, 4, qphys=q, ctx=ctx)
        pkg.applyOps(hpsi, list(range(1, 6)), Cup.T, trail=F)
        pkg.applyOps(hpsi, list(range(1, 6)), Cdn.T, trail=F)
        out.append(pkg.dmrg(hpsi, pkg.MPO(hmpo, qphys=q) if q is not None else pkg.MPO(hmpo), m=45, sweeps=10,
                            cutoff=1e-12, method="3S"))
    say(" Hubbard L=10: dense", out[0], "U(1)xU(1)", out[1])
    assert abs(out[0] - out[1]) < 1e-4
    assert abs(out[1] - (-25.380567747816)) < 1e-4


def test_3s_entry_points_reject_bad_arguments(pkg, ctx):
    """Error behaviour of the 3S and k-Lanczos entry points: a clean
    DMRJ_ERR_ARG (raised as DMRJError) for the wrong environment side, a
    neighbour that does not share the bond, a two-site tensor without its
    second MPO tensor, and a direction other than +-1."""
    say("case 3S argument errors")
    i = 4
    opsi, mpo_h, oL, oR = _oracle_state(10, 8, i, True)
    psi, mpo, L, R = _upload(pkg, ctx, opsi, mpo_h)
    with pytest.raises(pkg.DMRJError):
        pkg.expand3s(R[i], psi[i], mpo, i, 0.5, 1)        # moving right needs the left environment
    with pytest.raises(pkg.DMRJError):
        pkg.expand3s(L[i], psi[i], mpo, i, 0.5, -1)       # moving left needs the right environment
    with pytest.raises(pkg.DMRJError):
        pkg.expand3s(L[i], psi[i], mpo, i, 0.5, 2)
    T = pkg.expand3s(L[i], psi[i], mpo, i, 0.5, 1)
    T.ctx.call("dmrj_normalize", T.h)
    import ctypes as C
    nl, nr = C.c_void_p(), C.c_void_p()
    te, kept = C.c_double(), C.c_int64()
    sv = np.zeros(64)
    ctx.call("dmrj_svd_trunc2", T.h, 8, 2, 0.0, 1, C.byref(nl), C.byref(nr), C.byref(te), C.byref(kept),
             pkg._lib.dptr(sv), 64)
    left, right = pkg.DeviceTensor(ctx, nl), pkg.DeviceTensor(ctx, nr)
    with pytest.raises(pkg.DMRJError):
        pkg.absorb3s(psi[i], psi[i + 1], 1)               # not an SVD factor with a trivial index
    good = pkg.absorb3s(right, psi[i + 1], 1)
    assert good.shape[0] == right.shape[0]
    th = pkg.tdvp._theta(ctx, psi[i], psi[i + 1])
    with pytest.raises(pkg.DMRJError):
        ctx.call("dmrj_lanczos_k", L[i].h, R[i + 1].h, th.h, mpo.device(ctx, i), None, 2, 0, C.byref(C.c_void_p()),
                 C.byref(C.c_double()), C.byref(C.c_int()))
    for t in (T, left, right, good, th):
        t.free()
