"""applyOps! (lib/MPO_MPS/applyMPO.jl:21-33) on the GPU: the Hubbard filling of
test/dmrg_test.jl:67-75 built on dense and on U(1)xU(1) device states equals
the oracle's restatement (tensors bit for bit, bond charges), and DMRG from
the device-built state reproduces the oracle's per-sweep energies."""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models as om
from oracle import mps as omps

pytestmark = pytest.mark.gpu

QPHYS = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _oracle_filling(L, nup, ndn, charges):
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    o = omps.product_mps([0] * L, 4)
    omps.apply_ops(o, list(range(1, nup + 1)), Cup.T, trail=F)
    omps.apply_ops(o, list(range(1, ndn + 1)), Cdn.T, trail=F)
    if charges:
        omps.product_charges(o, QPHYS)
    return o


@pytest.mark.parametrize("charged", [False, True])
def test_applyops_hubbard_filling_matches_oracle(pkg, ctx, charged):
    say("case applyOps! Hubbard filling L=10 charged=%s" % charged)
    L = 10
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    psi = pkg.product_state([0] * L, 4, qphys=QPHYS if charged else None, ctx=ctx)
    pkg.applyOps(psi, list(range(1, 6)), Cup.T, trail=F)
    pkg.applyOps(psi, list(range(1, 6)), Cdn.T, trail=F)
    o = _oracle_filling(L, 5, 5, charged)
    tensors, qb = psi.to_dense()
    for i in range(L):
        assert tensors[i].shape == o.A[i].shape
        assert np.array_equal(tensors[i], o.A[i]), i
    if charged:
        for b in range(L + 1):
            assert np.array_equal(qb[b], o.qbond[b]), b
    if charged:
        # an operator that moves two different charges has no block form
        with pytest.raises(ValueError):
            pkg.applyOps(psi, [1], Cup.T + Cdn.T)
    psi.free()


def test_dmrg_from_applyops_state_matches_oracle(pkg, ctx):
    """Hubbard L=10 U(1)xU(1) from the applyOps!-built filling: per-sweep
    energies against the oracle run from its own filling, and the
    reference's m=45 value (test/model_test.jl:362-364)."""
    say("case DMRG from the applyOps! filling, Hubbard L=10")
    L = 10
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = om.fermionOps()
    psi = pkg.product_state([0] * L, 4, qphys=QPHYS, ctx=ctx)
    pkg.applyOps(psi, list(range(1, 6)), Cup.T, trail=F)
    pkg.applyOps(psi, list(range(1, 6)), Cdn.T, trail=F)
    mpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=QPHYS)
    params = pkg.algvars(psi)
    E = pkg.dmrg(psi, mpo, m=45, sweeps=10, cutoff=1e-12, params=params)
    o = _oracle_filling(L, 5, 5, True)
    oE, ohist = od.dmrg(o, om.makeMPO(om.hubbardMPO(), 4, L), m=45, sweeps=10, cutoff=1e-12)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel)
        assert rel <= 1e-12
    assert abs(E - (-25.380567747816)) < 1e-4
