"""Two-site TDVP (imaginary time, real prefactor) on the GPU path against the
oracle (algs/time.jl:182-202,288-362).

Tolerances: Krylov propagator output and Ritz energy <= 1e-12 relative on
identical inputs; whole half-sweeps: energies <= 1e-10 relative and local
<Sz_i> <= 1e-10 absolute (truncated SVDs in between, gauge-free quantities).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import mps as omps
from oracle import tdvp as ot

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_krylov_expm_vs_oracle(pkg, ctx):
    say("case krylov expm2/expm1")
    rng = np.random.default_rng(21)
    dims = [1, 6, 12, 7, 1]
    L = 4
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    A = [rng.standard_normal((dims[i], 2, dims[i + 1])) for i in range(L)]
    dev = [pkg.site_from_dense(ctx, a) for a in A]
    gL = pkg.Lupdate(pkg.network.makeBoundary(ctx, True), dev[0], mpo, 1)
    gR = pkg.Rupdate(pkg.network.makeBoundary(ctx, False), dev[3], mpo, 4)
    oL = od.Lupdate(np.ones((1, 1, 1)), A[0], mpo.H[0])
    oR = od.Rupdate(np.ones((1, 1, 1)), A[3], mpo.H[3])
    h = C.c_void_p()
    ctx.call("dmrj_theta2", dev[1].h, dev[2].h, C.byref(h))
    th = pkg.DeviceTensor(ctx, h)
    oth = np.tensordot(A[1], A[2], axes=(2, 0))
    for r in (1, 2, 6, 12):
        out, e = pkg.krylovTimeEvol2(gL, gR, th, mpo, 2, r, -0.25)
        want, we = ot.krylov_expm(lambda t: od.heff2(oL, oR, t, mpo.H[1], mpo.H[2]), oth, r, -0.25)
        got = out.to_dense()[0]
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        say(" expm2 r", r, "err", err, "E", e, we)
        assert err <= 1e-12
        assert abs(e - we) <= 1e-12 * abs(we)
    gR2 = pkg.Rupdate(gR, dev[2], mpo, 3)
    oR2 = od.Rupdate(oR, A[2], mpo.H[2])
    out, e = pkg.krylovTimeEvol1(gL, gR2, dev[1], mpo, 2, 6, -0.4)
    want, we = ot.krylov_expm(lambda t: od.singlesite_update(oL, oR2, t, mpo.H[1]), A[1], 6, -0.4)
    err = np.max(np.abs(out.to_dense()[0] - want)) / np.max(np.abs(want))
    say(" expm1 err", err)
    assert err <= 1e-12 and abs(e - we) <= 1e-12 * abs(we)


def test_tdvp_twosite_imaginary_time_u1(pkg, ctx):
    """U(1) Heisenberg L=10 from the Neel state: four imaginary-time
    half-sweeps (m=16, Krylov 6) against the oracle."""
    say("case tdvp U(1) L=10")
    L = 10
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    opsi = omps.product_mps(states, 2, qphys=qphys)
    from oracle import models as om
    ompo = om.makeMPO(om.heisenbergMPO(), 2, L)
    Sz = np.diag([0.5, -0.5])
    gLe = gRe = None
    oLe = oRe = None
    for hs in range(4):
        E, gLe, gRe = pkg.tdvp_twosite(psi, mpo, m=16, cutoff=1e-12, r=6, prefactor=-0.5 * (L - 1),
                                       Lenv=gLe, Renv=gRe)
        oE, oLe, oRe = ot.tdvp_twosite(opsi, ompo, m=16, cutoff=1e-12, r=6, prefactor=-0.5 * (L - 1),
                                       Lenv=oLe, Renv=oRe)
        sz = [pkg.local_expect(psi, Sz, i) for i in range(1, L + 1)]
        osz = [ot.local_expect(opsi, Sz, i) for i in range(1, L + 1)]
        say(" half-sweep", hs, "E %.14f %.14f" % (E, oE), "max dSz %.2e" % np.max(np.abs(np.subtract(sz, osz))))
        assert abs(E - oE) <= 1e-10 * abs(oE)
        assert np.max(np.abs(np.subtract(sz, osz))) <= 1e-10
