"""J1-J2 Heisenberg cylinder (width 4, J2 = 0.5, U(1) Sz), the model of the
precision-variant comparison, at a size the oracle and exact diagonalisation
finish in seconds: GPU FP64 sweeps against the oracle per sweep, then FP64,
BF16x3 and FP16x3 (split precision in H_eff*psi with FP64 environments and
the FP64 energy correction) converged energies against exact diagonalisation
and their centre-bond spectra against the FP64 run."""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models as om
from oracle import mps as omps

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _setup(pkg, Lx, Ly):
    L = Lx * Ly
    qphys = np.array([[1], [-1]])
    bulk, _ = pkg.j1j2_bulk(Lx, Ly, J1=1.0, J2=0.5)
    mpo = pkg.MPO(pkg.makeMPO(bulk, 2, L).H, qphys=qphys)
    states = [i % 2 for i in range(L)]
    return L, qphys, mpo, states


def test_j1j2_cylinder_fp64_vs_oracle(pkg, ctx):
    say("case J1-J2 cylinder 3x4 U(1) m=32 vs oracle")
    Lx, Ly = 3, 4
    L, qphys, mpo, states = _setup(pkg, Lx, Ly)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    params = pkg.algvars(psi)
    pkg.dmrg(psi, mpo, m=32, sweeps=4, cutoff=1e-12, params=params)
    ompo, _ = om.j1j2_cylinder_mpo(Lx, Ly, J1=1.0, J2=0.5)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    _, ohist = od.dmrg(opsi, ompo, m=32, sweeps=4, cutoff=1e-12)
    for h, oh in zip(params.history, ohist):
        rel = abs(h["energy"] - oh["energy"]) / abs(oh["energy"])
        say(" sweep", h["sweep"], "%.15f" % h["energy"], "%.15f" % oh["energy"], "rel %.2e" % rel,
            "m", h["biggestm"], oh["biggestm"])
        assert rel <= 1e-12
        assert h["biggestm"] == oh["biggestm"]
        assert np.allclose(h["storeD"], oh["storeD"], rtol=1e-8, atol=1e-12)


def test_j1j2_precision_variants_vs_ed(pkg, ctx):
    say("case J1-J2 cylinder 3x4 precision variants vs ED")
    Lx, Ly = 3, 4
    L, qphys, mpo, states = _setup(pkg, Lx, Ly)
    ompo, _ = om.j1j2_cylinder_mpo(Lx, Ly, J1=1.0, J2=0.5)
    e_ed = np.linalg.eigvalsh(om.fullH(ompo))[0]
    res = {}
    for prec, envp in (("f64", None), ("bf16x3", "f64"), ("f16x3", "f64")):
        psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
        params = pkg.algvars(psi)
        pkg.dmrg(psi, mpo, m=64, sweeps=10, cutoff=1e-14, params=params, precision=prec, env_precision=envp)
        res[prec] = (params.energy, np.array(params.history[-1]["storeD"]))
        say(" %-6s E %.15f  rel to ED %.2e" % (prec, params.energy, abs(params.energy - e_ed) / abs(e_ed)))
    e64, d64 = res["f64"]
    # m = 64 = 2^(L/2) represents the ground state exactly
    assert abs(e64 - e_ed) / abs(e_ed) <= 1e-11
    for prec, (etol, dtol) in (("f16x3", (1e-10, 1e-6)), ("bf16x3", (1e-8, 1e-4))):
        e, d = res[prec]
        rel = abs(e - e64) / abs(e64)
        n = min(len(d), len(d64))
        dd = np.max(np.abs(d[:n] - d64[:n]))
        say(" %-6s |dE/E| vs FP64 %.2e, centre spectrum max |dS| %.2e" % (prec, rel, dd))
        assert rel <= etol
        assert dd <= dtol
