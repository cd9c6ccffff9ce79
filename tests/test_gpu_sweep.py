"""Whole two-site DMRG runs on the GPU path against the oracle and the
reference's constants.

Per-sweep energies: <= 1e-12 relative to the oracle (the FP64 drop-in bar).  Centre
bond spectra (top-8 singular values) and the largest truncation error per
sweep are compared within 1e-8 relative (gauge- and tie-robust quantities).
"""
import json
import os
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models as om
from oracle import mps as omps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_cfg1_dense_heisenberg_L20_m64(pkg, ctx):
    """open Heisenberg L=20, m=64, 6 sweeps, Krylov 2, cutoff 1e-12, seed-0
    N(0,1) initial MPS."""
    say("case cfg1 L=20 m=64 6 sweeps")
    with open(os.path.join(GOLD, "cfg1.json")) as f:
        g = json.load(f)
    L = 20
    psi = pkg.randMPS_normal(L, 2, 64, seed=0, ctx=ctx)
    host, _ = psi.to_dense()
    opsi = omps.seeded_normal_mps(L, 2, 64, seed=0)
    for a, b in zip(host, opsi.A):
        assert np.allclose(a, b, rtol=1e-13, atol=1e-15)
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    params = pkg.algvars(psi)
    E = pkg.dmrg(psi, mpo, m=64, sweeps=6, cutoff=1e-12, params=params)
    for h, gh in zip(params.history, g["sweeps"]):
        rel = abs(h["energy"] - gh["energy"]) / abs(gh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "oracle %.15f" % gh["energy"], "rel %.2e" % rel,
            "maxtrunc", h["maxtrunc"], gh["maxtrunc"], "t=%.3fs" % h["time"])
        assert rel <= 1e-12
        assert np.allclose(h["storeD"], gh["storeD"], rtol=1e-8, atol=1e-12)
        assert abs(h["maxtrunc"] - gh["maxtrunc"]) <= 1e-8 * max(gh["maxtrunc"], 1e-14) + 1e-15
        assert h["biggestm"] == gh["biggestm"]
    Eexp = pkg.expect(psi, mpo)
    assert abs(Eexp - E) < 1e-10


def test_heisenberg_L10_reaches_ed(pkg, ctx):
    say("case L=10 ED")
    L = 10
    psi = pkg.randMPS_normal(L, 2, 32, seed=1, ctx=ctx)
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    E = pkg.dmrg(psi, mpo, m=32, sweeps=8, cutoff=1e-12)
    say(" E", E)
    assert abs(E - (-4.258035207282883)) < 1e-9


def test_cutoff_zero_from_product_state_is_variational(pkg, ctx):
    """Cutoff 0 from the Neel product state keeps numerically zero singular
    values (their vectors completed orthonormally): energies decrease sweep
    by sweep and reach the L=10 exact ground state (test/MPtask_test.jl:323)."""
    say("case cutoff 0 product start L=10")
    L = 10
    qphys = np.array([[1], [-1]])
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    pkg.dmrg(psi, qmpo, m=32, sweeps=6, cutoff=0.0, params=params)
    es = [h["energy"] for h in params.history]
    say(" energies", ["%.12f" % e for e in es], "m", [h["biggestm"] for h in params.history])
    assert all(b <= a + 1e-12 for a, b in zip(es, es[1:]))
    assert abs(es[-1] - (-4.258035207282883)) < 1e-9


def test_u1_heisenberg_L24_vs_oracle(pkg, ctx):
    """U(1) Sz blocks, Neel start; per-sweep energies vs the QN oracle and
    dense-vs-QN agreement (test/dmrg_test.jl:82-85)."""
    say("case U(1) L=24 m=48")
    L = 24
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo_h = pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H
    qmpo = pkg.MPO(mpo_h, qphys=qphys)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    pkg.dmrg(psi, qmpo, m=48, sweeps=4, cutoff=1e-12, params=params)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    oE, ohist = od.dmrg(opsi, om.makeMPO(om.heisenbergMPO(), 2, L), m=48, sweeps=4, cutoff=1e-12)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel,
            "m", h["biggestm"], oh["biggestm"], "t=%.3fs" % h["time"])
        assert rel <= 1e-12
        assert h["biggestm"] == oh["biggestm"]
        assert np.allclose(h["storeD"], oh["storeD"], rtol=1e-8, atol=1e-12)
    # dense path on the same start agrees within the reference's tolerance
    dpsi = pkg.product_state(states, 2, ctx=ctx)
    Ed = pkg.dmrg(dpsi, pkg.makeMPO(pkg.heisenbergMPO(), 2, L), m=48, sweeps=4, cutoff=1e-12)
    assert abs(Ed - params.energy) < 1e-4


def test_hubbard_L10_u1xu1_reference_constant(pkg, ctx):
    """test/model_test.jl:362-364: m=45 DMRG energy -25.380567747816 +- 1e-4."""
    say("case Hubbard L=10 U(1)xU(1)")
    L = 10
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    opsi = omps.product_mps([0] * L, 4)
    omps.apply_ops(opsi, list(range(1, 6)), Cup.T, trail=F)
    omps.apply_ops(opsi, list(range(1, 6)), Cdn.T, trail=F)
    qphys = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    omps.product_charges(opsi, qphys)
    # the GPU start state is built by the product's own applyOps! on the device
    psi = pkg.product_state([0] * L, 4, qphys=qphys, ctx=ctx)
    pkg.applyOps(psi, list(range(1, 6)), Cup.T, trail=F)
    pkg.applyOps(psi, list(range(1, 6)), Cdn.T, trail=F)
    tensors, qb = psi.to_dense()
    assert all(np.array_equal(a, b) for a, b in zip(tensors, opsi.A))
    assert all(np.array_equal(a, b) for a, b in zip(qb, opsi.qbond))
    mpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=qphys)
    params = pkg.algvars(psi)
    E = pkg.dmrg(psi, mpo, m=45, sweeps=10, cutoff=1e-12, params=params)
    oE, ohist = od.dmrg(opsi, om.makeMPO(om.hubbardMPO(), 4, L), m=45, sweeps=10, cutoff=1e-12)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel)
        assert rel <= 1e-12
    assert abs(E - (-25.380567747816)) < 1e-4


def test_autotune_is_result_neutral(pkg, ctx):
    """The grouped-GEMM autotuner picks tile shapes and the XCD-aware order
    deals tiles to work-groups only: every shape sums K in the same order, so
    per-sweep energies and spectra are bit-identical with the tuner on and
    off and with either tile order."""
    say("case autotune on/off U(1) L=32 m=96")
    L = 32
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    runs = []
    for mode, order in ((0, 1), (1, 1), (1, 0)):
        ctx.set_autotune(2)
        ctx.set_autotune(mode)
        ctx.set_tile_order(order)
        try:
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            params = pkg.algvars(psi)
            pkg.dmrg(psi, qmpo, m=96, sweeps=3, cutoff=1e-10, params=params)
        finally:
            ctx.set_tile_order(1)
        st = ctx.autotune_stats()
        say(" autotune", mode, "xcd order", order, "E %.15f" % params.energy, st)
        runs.append(params.history)
        if mode == 1:
            assert st["keys"] > 0 and st["tuned"] > 0
            assert st["hits"] > 0
    ctx.set_autotune(1)
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert a["energy"] == b["energy"]
            assert a["maxtrunc"] == b["maxtrunc"]
            assert list(a["storeD"]) == list(b["storeD"])


@pytest.mark.parametrize("backend", [0, 5])
def test_eig_batching_modes_agree(pkg, ctx, backend):
    """The SVD's batched eigensolves give the same sweeps whether the sectors
    form one batch padded to the largest, one batch each, batches of two
    sectors (mode 3: with several classes of several sectors every per-class
    and per-member buffer offset is used), or the cost-model size classes
    (which keep one batch at these sizes).  L=16 keeps every
    bond below m, so the cutoff alone truncates and no rounding-level change
    can move a degenerate multiplet across the bond-dimension cut.  Per-sweep
    energies within 1e-12 relative for the augmented form (backend 0), 1e-9
    for the Gram form (backend 5, sigma^2 to eps |M|^2); spectra within
    1e-12 / 1e-8 absolute."""
    say("case eigensolve batching modes, backend", backend)
    L = 16
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    runs = []
    for mode in (0, 2, 3, 1):
        ctx.set_eig_batching(mode)
        ctx.set_svd_backend(backend)
        try:
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            params = pkg.algvars(psi)
            pkg.dmrg(psi, qmpo, m=512, sweeps=3, cutoff=1e-10, params=params)
        finally:
            ctx.set_eig_batching(1)
            ctx.set_svd_backend(0)
        say(" mode", mode, "E %.15f" % params.energy)
        runs.append(params.history)
    tol_e, tol_s = (1e-12, 1e-12) if backend == 0 else (1e-9, 1e-8)
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert abs(a["energy"] - b["energy"]) <= tol_e * abs(a["energy"])
            sa, sb = np.asarray(a["storeD"]), np.asarray(b["storeD"])
            n = min(len(sa), len(sb))
            assert np.max(np.abs(sa[:n] - sb[:n])) <= tol_s


def test_graph_captured_lanczos_matches_direct(pkg, ctx):
    """Lanczos steps captured into hipGraphs (instantiated once, then updated
    in place bond after bond) give the same per-sweep energies, truncation
    errors and spectra as direct launches, bit for bit, with the plain and
    the FP64-corrected mixed-precision step."""
    L = 24
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    for precision, correction in (("f64", "none"), ("bf16x3", "bond")):
        runs = []
        for mode in (0, 1):
            say("case graph mode", mode, precision, correction, "U(1) L=24 m=64")
            ctx.set_graph(mode)
            try:
                psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
                params = pkg.algvars(psi)
                pkg.dmrg(psi, qmpo, m=64, sweeps=3, cutoff=1e-10, params=params, precision=precision,
                         correction=correction)
            finally:
                ctx.set_graph(0)
            runs.append(params.history)
            say(" E %.15f" % params.energy, ctx.graph_stats())
        st = ctx.graph_stats()
        assert st["instantiations"] >= 1 and st["updates"] > 0
        for a, b in zip(*runs):
            assert a["energy"] == b["energy"]
            assert a["maxtrunc"] == b["maxtrunc"]
            assert list(a["storeD"]) == list(b["storeD"])


def test_gram_svd_backend_energy(pkg, ctx):
    """The fast SVD backend (Gram eigensolve, singular values to eps |M|^2
    absolute) reaches the same converged ground-state energy as the FP64
    drop-in backend: U(1) L=32, m=64, cutoff 1e-10, 8 sweeps, |dE/E| <= 1e-10."""
    L = 32
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    energies = []
    for backend in (0, 5):
        say("case svd backend", backend, "U(1) L=32 m=64")
        ctx.set_svd_backend(backend)
        try:
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            params = pkg.algvars(psi)
            pkg.dmrg(psi, qmpo, m=64, sweeps=8, cutoff=1e-10, params=params)
        finally:
            ctx.set_svd_backend(0)
        for h in params.history:
            say(" sweep", h["sweep"], "%.15f" % h["energy"], "m", h["biggestm"], "maxtrunc", h["maxtrunc"])
        energies.append(params.energy)
    assert abs(energies[0] - energies[1]) <= 1e-10 * abs(energies[0])


def test_mps_copy_is_independent(pkg, ctx):
    """MPS.copy gives device copies: optimising the copy leaves the original
    state (its tensors and energy) unchanged."""
    say("case MPS copy")
    L = 10
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.dmrg(psi, mpo, m=16, sweeps=1, cutoff=0.0)
    before, qb = psi.to_dense()
    e0 = pkg.expect(psi, mpo)
    cp = psi.copy()
    got, qc = cp.to_dense()
    for a, b in zip(before, got):
        assert np.array_equal(a, b)
    e1 = pkg.dmrg(cp, mpo, m=32, sweeps=2, cutoff=0.0)
    after, _ = psi.to_dense()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert pkg.expect(psi, mpo) == e0
    assert e1 < e0 + 1e-12
    cp.free()


def test_eig_batching_pairs_against_oracle_3s_and_complex(pkg, ctx):
    """Batches of two sectors (eigensolve batching mode 3) on the paths whose
    batches the cost model rarely splits at test sizes: the two-site sweep
    (augmented form) against the oracle per sweep, 3S (its enlarged sectors go
    through the LQ/QR-reduced eigensolve) against mode 0, and real-time TDVP
    (complex Hermitian eigensolves) against the oracle."""
    from oracle import tdvp as ot
    say("case eigensolve batches of two sectors")
    L = 16
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    qmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    ompo = om.makeMPO(om.heisenbergMPO(), 2, L)
    ctx.set_eig_batching(3)
    try:
        psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
        params = pkg.algvars(psi)
        pkg.dmrg(psi, qmpo, m=512, sweeps=3, cutoff=1e-10, params=params)
        opsi = omps.product_mps(states, 2, qphys=qphys)
        oE, ohist = od.dmrg(opsi, ompo, m=512, sweeps=3, cutoff=1e-10)
        for h, oh in zip(params.history, ohist):
            rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
            say(" two-site sweep %d rel %.2e" % (h["sweep"], rel))
            assert rel <= 1e-12
            assert np.max(np.abs(np.asarray(h["storeD"]) - np.asarray(oh["storeD"]))) <= 1e-10
        runs = []
        for mode in (3, 0):
            ctx.set_eig_batching(mode)
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            p3 = pkg.algvars(psi)
            pkg.dmrg(psi, qmpo, m=160, sweeps=3, cutoff=1e-12, params=p3, method="3S")
            runs.append(p3.history)
            say(" 3S mode %d E %.15f" % (mode, p3.energy))
        for a, b in zip(*runs):
            assert abs(a["energy"] - b["energy"]) <= 1e-12 * abs(a["energy"])
        ctx.set_eig_batching(3)
        Lt = 10
        tmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, Lt).H, qphys=qphys)
        tpsi = pkg.product_state([i % 2 for i in range(Lt)], 2, qphys=qphys, ctx=ctx)
        opsi = omps.product_mps([i % 2 for i in range(Lt)], 2, qphys=qphys)
        otm = om.makeMPO(om.heisenbergMPO(), 2, Lt)
        Sz = np.diag([0.5, -0.5])
        gLe = gRe = oLe = oRe = None
        for hs in range(4):
            _, gLe, gRe = pkg.tdvp_twosite(tpsi, tmpo, m=16, cutoff=1e-12, r=12, prefactor=-0.05j, Lenv=gLe, Renv=gRe)
            _, oLe, oRe = ot.tdvp_twosite(opsi, otm, m=16, cutoff=1e-12, r=12, prefactor=-0.05j, Lenv=oLe, Renv=oRe)
        sz = np.array([pkg.local_expect(tpsi, Sz, i) for i in range(1, Lt + 1)])
        osz = np.array([ot.local_expect(opsi, Sz, i) for i in range(1, Lt + 1)])
        say(" TDVP max dSz %.2e" % np.max(np.abs(sz - osz)))
        assert np.max(np.abs(sz - osz)) <= 1e-10
    finally:
        ctx.set_eig_batching(1)
