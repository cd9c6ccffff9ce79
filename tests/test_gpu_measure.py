"""correlationmatrix on GPU-resident MPS (lib/MPO_MPS/correlationmatrix.jl:21-83)
against the oracle's brute force on the full state vector and the
reference's golden matrix (test/MPmeas_test.jl:310-315)."""
import json
import os
import sys

import numpy as np
import pytest

from oracle import measure as ome
from oracle import models as om

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_correlationmatrix_heisenberg_golden(pkg, ctx):
    """Dense MPS of the exact L=10 ground state (the reference's test builds it
    with eigen(mpo) and makeMPS): the reference's own acceptance test and the
    oracle on the same vector to 1e-12."""
    say("case correlationmatrix L=10 golden")
    with open(os.path.join(GOLD, "correlationmatrix_L10.json")) as f:
        chk = np.array(json.load(f)["checkM"])
    E, V = np.linalg.eigh(om.fullH(om.makeMPO(om.heisenbergMPO(), 2, 10)))
    A = ome.vector_to_mps(V[:, 0], 2, 10)
    psi = pkg.MPS(A, oc=1, ctx=ctx)
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    M = pkg.correlationmatrix(psi, Sp, Sm)
    want = ome.correlationmatrix(V[:, 0], 2, 10, Sp, Sm)
    say(" max |M - oracle| %.2e, norm(checkM - M)/100 %.2e" % (np.max(np.abs(M - want)),
                                                              np.linalg.norm(chk - M) / chk.size))
    assert np.linalg.norm(chk - M) / chk.size < 1e-5
    assert np.max(np.abs(M - want)) < 1e-12


def test_correlationmatrix_u1_dmrg_state(pkg, ctx):
    """U(1) block-sparse state from GPU DMRG (L=12, m=24, charged operators
    S+ / S- and neutral Sz Sz) against the oracle on the same state's vector."""
    say("case correlationmatrix U(1) L=12")
    L = 12
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.dmrg(psi, mpo, m=24, sweeps=3, cutoff=1e-12)
    host, _ = psi.to_dense()
    vec = ome.mps_to_vector(host)
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    for name, (a, b) in (("S+S-", (Sp, Sm)), ("S-S+", (Sm, Sp)), ("SzSz", (Sz, Sz))):
        M = pkg.correlationmatrix(psi, a, b)
        want = ome.correlationmatrix(vec, 2, L, a, b)
        err = np.max(np.abs(M - want))
        say(" %s max |M - oracle| %.2e" % (name, err))
        assert err < 1e-12


def test_correlationmatrix_fermions_with_trail(pkg, ctx):
    """Free fermions (Hubbard U = 0, mu = 0, L = 4), dense MPS of the exact
    ground state: <c+_i c_j> with the Jordan-Wigner trail against the oracle
    and the projector on the occupied orbitals."""
    say("case correlationmatrix fermions L=4")
    L = 4
    E, V = np.linalg.eigh(om.fullH(om.makeMPO(om.hubbardMPO(t=1.0, mu=0.0, HubU=0.0), 4, L)))
    psi = pkg.MPS(ome.vector_to_mps(V[:, 0], 4, L), oc=1, ctx=ctx)
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = pkg.fermionOps()
    M = pkg.correlationmatrix(psi, Cup.T, Cup, trail=F)
    want = ome.correlationmatrix(V[:, 0], 4, L, Cup.T, Cup, trail=F)
    h = np.diag(np.ones(L - 1), 1) + np.diag(np.ones(L - 1), -1)
    e, phi = np.linalg.eigh(h)
    occ = phi[:, e < 0]
    say(" max |M - oracle| %.2e, |M - P_occ| %.2e" % (np.max(np.abs(M - want)), np.max(np.abs(M - occ @ occ.T))))
    assert np.max(np.abs(M - want)) < 1e-12
    assert np.max(np.abs(M - occ @ occ.T)) < 1e-12


@pytest.mark.parametrize("qn", [False, True], ids=["dense", "u1"])
def test_correlationmatrix_complex_tdvp_state(pkg, ctx, qn):
    """ComplexF64 state after real-time TDVP from the Neel state (L=8):
    <S+_i S-_j> and <Sz_i Sz_j> against the oracle's brute force on the
    complex amplitude vector (1e-12), Hermitian, not real."""
    say("case correlationmatrix complex", "u1" if qn else "dense")
    L = 8
    qphys = np.array([[1], [-1]]) if qn else None
    mpo_h = om.makeMPO(om.heisenbergMPO(), 2, L)
    mpo = pkg.MPO(mpo_h, qphys=qphys) if qn else pkg.MPO(mpo_h)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    Le = Re = None
    for _ in range(6):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=16, cutoff=1e-12, r=8, prefactor=-1.0j, Lenv=Le, Renv=Re)
    assert any(t.is_complex for t in psi.A)
    host, _ = psi.to_dense()
    vec = ome.mps_to_vector([np.asarray(a, dtype=np.complex128) for a in host])
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    for Cc, Ca in ((Sp, Sm), (Sz, Sz)):
        M = pkg.correlationmatrix(psi, Cc, Ca)
        want = ome.correlationmatrix(vec, 2, L, Cc, Ca)
        err = np.max(np.abs(M - want))
        say(" max |M - oracle| %.2e, max |Im M| %.2e" % (err, np.max(np.abs(M.imag))))
        assert err < 1e-12 * max(1.0, np.max(np.abs(want)))
        assert np.allclose(M, M.conj().T, atol=1e-13)
    assert np.max(np.abs(pkg.correlationmatrix(psi, Sp, Sm).imag)) > 1e-6


def test_correlationmatrix_two_state_golden(pkg, ctx):
    """correlationmatrix(psi, psi, Sp, Sm) as the reference's test calls it
    (test/MPmeas_test.jl:316-318): the golden checkM and the single-state
    form."""
    say("case correlationmatrix(psi, psi) L=10 golden")
    with open(os.path.join(GOLD, "correlationmatrix_L10.json")) as f:
        chk = np.array(json.load(f)["checkM"])
    E, V = np.linalg.eigh(om.fullH(om.makeMPO(om.heisenbergMPO(), 2, 10)))
    A = ome.vector_to_mps(V[:, 0], 2, 10)
    psi = pkg.MPS(A, oc=1, ctx=ctx)
    psi2 = psi.copy()
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    M2 = pkg.correlationmatrix(psi2, psi, Sp, Sm)       # two handles, same state
    M1 = pkg.correlationmatrix(psi, Sp, Sm)
    want = ome.transition_matrix(V[:, 0], V[:, 0], 2, 10, Sp, Sm)
    say(" two-state vs oracle %.2e, vs single-state %.2e" % (np.max(np.abs(M2 - want)), np.max(np.abs(M2 - M1))))
    assert np.linalg.norm(chk - M2) / chk.size < 1e-5
    assert np.max(np.abs(M2 - want)) < 1e-12
    assert np.max(np.abs(M2 - M1)) < 1e-12


def test_two_state_measurements_u1(pkg, ctx):
    """<dualpsi| ... |psi> between two different U(1) states (GPU DMRG at two
    bond dimensions) against the oracle's brute force on their vectors:
    correlationmatrix(dualpsi, psi, S+, S-) and correlation(dualpsi, psi,
    S+, S-), and correlation(psi, Sz, Sz, Sz) (three operators)."""
    say("case two-state measurements U(1) L=8")
    L = 8
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.dmrg(psi, mpo, m=16, sweeps=3, cutoff=1e-12)
    phi = pkg.product_state([(i // 2) % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.dmrg(phi, mpo, m=4, sweeps=1, cutoff=1e-12)
    vpsi = ome.mps_to_vector(psi.to_dense()[0])
    vphi = ome.mps_to_vector(phi.to_dense()[0])
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    M = pkg.correlationmatrix(phi, psi, Sp, Sm)
    want = ome.transition_matrix(vphi, vpsi, 2, L, Sp, Sm)
    err = np.max(np.abs(M - want))
    say(" correlationmatrix(phi, psi) max err %.2e (scale %.2e)" % (err, np.max(np.abs(want))))
    assert err < 1e-12
    W = pkg.correlation(phi, psi, Sp, Sm)
    want = ome.correlation(vphi, vpsi, 2, L, [Sp, Sm])
    err = np.max(np.abs(W - want))
    say(" correlation(phi, psi, S+, S-) max err %.2e" % err)
    assert err < 1e-12
    W3 = pkg.correlation(psi, Sz, Sz, Sz)
    want = ome.correlation(vpsi, vpsi, 2, L, [Sz, Sz, Sz]) / np.vdot(vpsi, vpsi)
    err = np.max(np.abs(W3 - want))
    say(" correlation(psi, Sz, Sz, Sz) max err %.2e" % err)
    assert err < 1e-12


def test_correlation_fermions_with_trail(pkg, ctx):
    """correlation(psi, c+, c; trail=F) on a U(1)xU(1) Hubbard DMRG state
    against the oracle (both orderings of the two sites, the Jordan-Wigner
    trail on the sites left of each operator)."""
    say("case correlation Hubbard L=6 with trail")
    L = 6
    qphys = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    psi = pkg.product_state([0] * L, 4, qphys=qphys, ctx=ctx)
    pkg.applyOps(psi, list(range(1, 4)), Cup.T, trail=F)
    pkg.applyOps(psi, list(range(1, 4)), Cdn.T, trail=F)
    mpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=qphys)
    pkg.dmrg(psi, mpo, m=32, sweeps=3, cutoff=1e-12)
    v = ome.mps_to_vector(psi.to_dense()[0])
    W = pkg.correlation(psi, Cup.T, Cup, trail=F)
    want = ome.correlation(v, v, 4, L, [Cup.T, Cup], trail=F)
    err = np.max(np.abs(W - want))
    say(" max err %.2e" % err)
    assert err < 1e-12
