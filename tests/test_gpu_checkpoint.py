"""Checkpoint / resume of GPU-resident MPS (save_mps / load_mps through
dmrj_site_export / dmrj_site_import): bit-exact round trips of dense, U(1),
U(1)xU(1) and ComplexF64 states, and a DMRG run resumed from a checkpoint that
reproduces the uninterrupted run."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _same(pkg, a, b):
    assert len(a) == len(b) and a.oc == b.oc and a.nq == b.nq and a.d == b.d
    for x, y in zip(a.A, b.A):
        assert x.info() == y.info()
        assert x.is_complex == y.is_complex
        if x.is_complex:
            assert np.array_equal(x.to_dense(a.nq)[0], y.to_dense(b.nq)[0])
        else:
            xa, xl, xr = x.to_dense(a.nq)
            ya, yl, yr = y.to_dense(b.nq)
            assert np.array_equal(xa, ya)
            if a.nq:
                assert np.array_equal(xl, yl) and np.array_equal(xr, yr)


def test_checkpoint_roundtrips(pkg, ctx, tmp_path):
    say("case checkpoint round trips")
    # dense, seeded N(0,1) state
    psi = pkg.randMPS_normal(12, 2, 16, seed=3, ctx=ctx)
    p = str(tmp_path / "dense.npz")
    pkg.save_mps(psi, p, extra={"energy": 1.5})
    back, extra = pkg.load_mps(p, ctx)
    _same(pkg, psi, back)
    assert extra == {"energy": 1.5}
    # U(1) after DMRG
    L = 12
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.dmrg(psi, mpo, m=24, sweeps=2, cutoff=1e-12)
    p = str(tmp_path / "u1.npz")
    pkg.save_mps(psi, p)
    back, _ = pkg.load_mps(p, ctx)
    _same(pkg, psi, back)
    assert abs(pkg.expect(back, mpo) - pkg.expect(psi, mpo)) == 0.0
    # U(1) x U(1) Hubbard
    L = 6
    qh = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    hmpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=qh)
    hpsi = pkg.product_state([1, 2, 1, 2, 1, 2], 4, qphys=qh, ctx=ctx)
    pkg.dmrg(hpsi, hmpo, m=32, sweeps=2, cutoff=1e-12)
    p = str(tmp_path / "u1u1.npz")
    pkg.save_mps(hpsi, p)
    back, _ = pkg.load_mps(p, ctx)
    _same(pkg, hpsi, back)
    # ComplexF64 after real-time TDVP
    L = 8
    tmpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    tpsi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    pkg.tdvp_twosite(tpsi, tmpo, m=16, cutoff=1e-12, r=8, prefactor=-0.1j)
    assert any(t.is_complex for t in tpsi.A)
    p = str(tmp_path / "z.npz")
    pkg.save_mps(tpsi, p)
    back, _ = pkg.load_mps(p, ctx)
    _same(pkg, tpsi, back)
    say(" sizes", {k: os.path.getsize(str(tmp_path / k)) for k in ("dense.npz", "u1.npz", "u1u1.npz", "z.npz")})


def test_dmrg_resumed_from_checkpoint(pkg, ctx, tmp_path):
    """Two sweeps, checkpoint, load, two more sweeps == four sweeps in one
    run: the same per-sweep energies (environments rebuilt by makeEnv)."""
    say("case DMRG resume")
    L = 16
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    states = [i % 2 for i in range(L)]
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    full = pkg.algvars(psi)
    pkg.dmrg(psi, mpo, m=32, sweeps=4, cutoff=1e-12, params=full)
    psi2 = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    first = pkg.algvars(psi2)
    pkg.dmrg(psi2, mpo, m=32, sweeps=2, cutoff=1e-12, params=first)
    p = str(tmp_path / "resume.npz")
    pkg.save_mps(psi2, p, extra={"sweeps": 2, "energy": first.energy})
    back, extra = pkg.load_mps(p, ctx)
    second = pkg.algvars(back)
    pkg.dmrg(back, mpo, m=32, sweeps=2, cutoff=1e-12, params=second)
    got = [h["energy"] for h in first.history] + [h["energy"] for h in second.history]
    want = [h["energy"] for h in full.history]
    say(" energies", ["%.15f" % e for e in got], "uninterrupted", ["%.15f" % e for e in want])
    assert extra["sweeps"] == 2
    assert got == want
