"""autoMPO (lib/MPO/autoMPO.jl:38-63,430-556) on the host: the MPO of a sum of
operator terms equals the hand-built MPOs of the models and the explicit
operator sums; compression keeps charges and reaches the minimal link
dimensions of the nearest-neighbour models."""
import numpy as np

from oracle import dmrg as od
from oracle import models as om
from oracle import mps as omps


def _heis_terms(pkg, L):
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    terms = []
    for i in range(1, L):
        terms += [pkg.mpoterm(0.5, Sp, i, Sm, i + 1), pkg.mpoterm(0.5, Sm, i, Sp, i + 1),
                  pkg.mpoterm(Sz, i, Sz, i + 1)]
    return terms


def _j1j2_terms(pkg, bonds):
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    terms = []
    for (a, b), J in bonds.items():
        i, j = min(a, b) + 1, max(a, b) + 1
        terms += [pkg.mpoterm(0.5 * J, Sp, i, Sm, j), pkg.mpoterm(0.5 * J, Sm, i, Sp, j), pkg.mpoterm(J, Sz, i, Sz, j)]
    return terms


def test_autompo_heisenberg_matches_hand_built(pkg):
    L = 8
    ref = om.fullH(om.makeMPO(om.heisenbergMPO(), 2, L))
    for q in (None, np.array([[1], [-1]])):
        Ws = pkg.autoMPO(_heis_terms(pkg, L), qphys=q)
        assert max(W.shape[3] for W in Ws) == 5          # the minimal link dimension
        assert np.max(np.abs(om.fullH(Ws) - ref)) < 1e-13
        if q is not None:
            pkg.mpo_deltas(Ws, q)                          # raises unless charge-conserving


def test_autompo_hubbard_with_trail(pkg):
    """Hopping terms with the fermion trail (mpoterm(val, Cup, i, Cup', i+1, F))
    give hubbardMPO (lib/models.jl:264-273) exactly; without the trail they
    do not."""
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = pkg.fermionOps()
    L, t, mu, U = 4, 1.0, -2.0, 4.0
    qh = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])

    def terms(trail):
        out = []
        for i in range(1, L + 1):
            out += [pkg.mpoterm(mu, Ndens, i), pkg.mpoterm(U, Nup @ Ndn, i)]
        for i in range(1, L):
            for c in (Cup, Cdn):
                extra = (F,) if trail else ()
                out += [pkg.mpoterm(-t, c, i, c.T, i + 1, *extra), pkg.mpoterm(t, c.T, i, c, i + 1, *extra)]
        return out
    ref = om.fullH(om.makeMPO(om.hubbardMPO(t=t, mu=mu, HubU=U), 4, L))
    Ws = pkg.autoMPO(terms(True), qphys=qh)
    assert max(W.shape[3] for W in Ws) == 6
    assert np.max(np.abs(om.fullH(Ws) - ref)) < 1e-12
    assert np.max(np.abs(om.fullH(pkg.autoMPO(terms(False))) - ref)) > 0.1


def test_autompo_j1j2_cylinder(pkg):
    """Width-4 J1-J2 cylinder: exact on 3x4 (full matrices) and on 8x4 / 32x4
    through <psi|H|psi> of random MPS; the compressed links (14) are smaller
    than the direct construction's (23)."""
    q = np.array([[1], [-1]])
    mpo, bonds = om.j1j2_cylinder_mpo(3, 4, J1=1.0, J2=0.5)
    Ws = pkg.autoMPO(_j1j2_terms(pkg, bonds), qphys=q)
    assert np.max(np.abs(om.fullH(Ws) - om.fullH(mpo))) < 1e-12
    for Lx in (8, 32):
        mpo, bonds = om.j1j2_cylinder_mpo(Lx, 4, J1=1.0, J2=0.5)
        Ws = pkg.autoMPO(_j1j2_terms(pkg, bonds), qphys=q)
        assert max(W.shape[3] for W in Ws) <= 14 < max(W.shape[3] for W in mpo)
        L = 4 * Lx
        rng = np.random.default_rng(Lx)
        for _ in range(2):
            psi = omps.MPS([rng.standard_normal((1 if i == 0 else 3, 2, 1 if i == L - 1 else 3)) for i in range(L)])
            e1, e2 = od.expect(psi, Ws), od.expect(psi, mpo)
            assert abs(e1 - e2) <= 1e-11 * abs(e2)


def test_autompo_general_terms_against_explicit_sum(pkg):
    """Single-site, two-site at range 3 and three-site terms with a
    non-symmetric operator: the MPO's matrix (fullH: ket j, bra k of W, so it
    is the transpose of the operator product) equals the explicit sum."""
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    L = 5

    def op_at(ops):
        out = np.ones((1, 1))
        for s in range(L, 0, -1):                           # site 1 fastest
            out = np.kron(out, ops.get(s, Id))
        return out
    spec = [(0.7, {2: Sz}), (1.3, {1: Sp, 4: Sm}), (-0.4, {2: Sz, 3: Sp, 5: Sz}), (0.25, {5: Sm})]
    terms = []
    for val, ops in spec:
        args = [val]
        for s in sorted(ops):
            args += [ops[s], s]
        terms.append(pkg.mpoterm(*args))
    Hex = sum(val * op_at(ops) for val, ops in spec)
    Ws = pkg.autoMPO(terms, Ns=L)
    assert np.max(np.abs(om.fullH(Ws) - Hex.T)) < 1e-13
