"""GPU DMRG with MPOs built by autoMPO: the same per-sweep energies as with
the hand-built MPO (Heisenberg U(1)) and the exact ground state of the J1-J2
3x4 cylinder."""
import sys

import numpy as np
import pytest

from oracle import models as om

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_dmrg_with_autompo(pkg, ctx):
    say("case DMRG with autoMPO")
    L = 16
    q = np.array([[1], [-1]])
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    terms = []
    for i in range(1, L):
        terms += [pkg.mpoterm(0.5, Sp, i, Sm, i + 1), pkg.mpoterm(0.5, Sm, i, Sp, i + 1), pkg.mpoterm(Sz, i, Sz, i + 1)]
    runs = []
    for mpo in (pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=q), pkg.MPO(pkg.autoMPO(terms, qphys=q), qphys=q)):
        psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=q, ctx=ctx)
        params = pkg.algvars(psi)
        pkg.dmrg(psi, mpo, m=32, sweeps=4, cutoff=1e-12, params=params)
        runs.append([h["energy"] for h in params.history])
    say(" hand-built", runs[0], "autoMPO", runs[1])
    for a, b in zip(*runs):
        assert abs(a - b) <= 1e-12 * abs(a)
    # J1-J2 3x4 cylinder from its bond list
    mpo_ref, bonds = om.j1j2_cylinder_mpo(3, 4, J1=1.0, J2=0.5)
    terms = []
    for (a, b), J in bonds.items():
        i, j = min(a, b) + 1, max(a, b) + 1
        terms += [pkg.mpoterm(0.5 * J, Sp, i, Sm, j), pkg.mpoterm(0.5 * J, Sm, i, Sp, j), pkg.mpoterm(J, Sz, i, Sz, j)]
    e_ed = np.linalg.eigvalsh(om.fullH(mpo_ref))[0]
    mpo = pkg.MPO(pkg.autoMPO(terms, qphys=q), qphys=q)
    psi = pkg.product_state([i % 2 for i in range(12)], 2, qphys=q, ctx=ctx)
    E = pkg.dmrg(psi, mpo, m=64, sweeps=10, cutoff=1e-14)
    say(" J1-J2 3x4 autoMPO E %.15f ED %.15f" % (E, e_ed))
    assert abs(E - e_ed) <= 1e-11 * abs(e_ed)
