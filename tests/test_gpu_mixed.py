"""Precision forms: BF16 split precision and FP16 split precision with
per-block scaling in the H_eff / environment contractions, with FP64 Lanczos
scalars, FP64 SVD and FP64 energy correction, against the FP64 drop-in.

Tolerances: one split-precision H_eff application agrees with FP64 to
<= 1e-4 relative for BF16x3 (about 2^-16 per product) and <= 1e-5 for FP16x3
(about 2^-21, FP32 accumulation); the FP64 corrected energy of a
split-precision run, after FP64 polish sweeps, matches the converged FP64
energy to |dE/E| <= 1e-10.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_heff2_bf16x3_vs_f64(pkg, ctx):
    import ctypes as Ct
    say("case heff2 bf16x3")
    rng = np.random.default_rng(3)
    L = 4
    mpo = pkg.makeMPO(pkg.heisenbergMPO(), 2, L)
    dims = [1, 48, 96, 40, 1]
    A = [rng.standard_normal((dims[i], 2, dims[i + 1])) for i in range(L)]
    dev = [pkg.site_from_dense(ctx, a) for a in A]
    Lenv = pkg.network.makeBoundary(ctx, True)
    L2 = pkg.Lupdate(Lenv, dev[0], mpo, 1)
    Renv = pkg.network.makeBoundary(ctx, False)
    R3 = pkg.Rupdate(Renv, dev[3], mpo, 4)
    th = Ct.c_void_p()
    ctx.call("dmrj_theta2", dev[1].h, dev[2].h, Ct.byref(th))
    th = pkg.DeviceTensor(ctx, th)
    out = {}
    for prec in (0, 1, 2):
        h = Ct.c_void_p()
        ctx.call("dmrj_heff2_apply", L2.h, mpo.device(ctx, 2), mpo.device(ctx, 3), R3.h, th.h, prec, Ct.byref(h))
        out[prec] = pkg.DeviceTensor(ctx, h).to_dense()[0]
    err = np.max(np.abs(out[1] - out[0])) / np.max(np.abs(out[0]))
    err16 = np.max(np.abs(out[2] - out[0])) / np.max(np.abs(out[0]))
    say(" rel err bf16x3 %.2e f16x3 %.2e" % (err, err16))
    assert err < 1e-4
    assert err16 < 1e-5


@pytest.mark.parametrize("precision", ["bf16x3", "f16x3"])
def test_mixed_dmrg_u1_heisenberg_converges_to_f64(pkg, ctx, precision):
    """U(1) Heisenberg L=32, m=64: FP64 reference run vs split-precision
    sweeps (FP64 energy correction each sweep) followed by FP64 polish
    sweeps; the converged energies agree to 1e-10 relative."""
    L = 32
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    say("case mixed U(1) L=32 m=64", precision)
    p64 = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    a64 = pkg.algvars(p64)
    pkg.dmrg(p64, mpo, m=64, sweeps=10, cutoff=1e-12, params=a64)
    pmx = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    amx = pkg.algvars(pmx)
    pkg.dmrg(pmx, mpo, m=64, sweeps=8, cutoff=1e-12, params=amx, precision=precision, correction="sweep")
    pkg.dmrg(pmx, mpo, m=64, sweeps=2, cutoff=1e-12, params=amx, precision="f64")
    for h64, hmx in zip(a64.history, amx.history):
        say(" sweep", h64["sweep"], "%.14f" % h64["energy"], "%.14f" % hmx["energy"], hmx["precision"],
            "res", hmx["residual"], "rel %.2e" % (abs(h64["energy"] - hmx["energy"]) / abs(h64["energy"])))
    e64, emx = a64.history[-1]["energy"], amx.history[-1]["energy"]
    assert abs(e64 - emx) / abs(e64) <= 1e-10
    # the corrected split-precision energy of each sweep is an FP64 Rayleigh quotient
    assert all(h["residual"] is not None for h in amx.history[:8])


@pytest.mark.parametrize("precision,tol", [("f16x3", 1e-10), ("bf16x3", 1e-8)])
def test_mixed_heff_fp64_env_energy_without_polish(pkg, ctx, precision, tol):
    """Split-precision H_eff*psi with FP64 environments and the FP64 energy
    correction at the end of every sweep: the corrected energy is a Rayleigh
    quotient of a state whose error is first order in the product error, so
    it meets the FP64 energy at second order without any FP64 polish sweep
    (FP16x3: ~2^-22 per product, |dE/E| <= 1e-10; BF16x3: ~2^-16, <= 1e-8)."""
    L = 32
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    say("case mixed H_eff, FP64 env", precision)
    p64 = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    a64 = pkg.algvars(p64)
    pkg.dmrg(p64, mpo, m=64, sweeps=8, cutoff=1e-12, params=a64)
    pmx = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
    amx = pkg.algvars(pmx)
    pkg.dmrg(pmx, mpo, m=64, sweeps=8, cutoff=1e-12, params=amx, precision=precision, correction="sweep",
             env_precision="f64")
    for h64, hmx in zip(a64.history, amx.history):
        say(" sweep", h64["sweep"], "%.14f" % h64["energy"], "%.14f" % hmx["energy"],
            "rel %.2e" % (abs(h64["energy"] - hmx["energy"]) / abs(h64["energy"])))
    e64, emx = a64.history[-1]["energy"], amx.history[-1]["energy"]
    assert abs(e64 - emx) / abs(e64) <= tol


def test_f16x3_planes_sweeps_same_bits(pkg, ctx):
    """Mixed sweeps (FP16x3 H_eff*psi, FP64 environments, Gram SVD) with the
    FP16x3 GEMMs on pre-split operand planes and with the on-load split:
    identical energies, truncation errors and spectra, sweep by sweep (the
    planes are shared by the segments that use one operand block, whose
    scaling is the same for all of them)."""
    L = 24
    qphys = np.array([[1], [-1]])
    states = [i % 2 for i in range(L)]
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    say("case FP16x3 planes vs on-load split, U(1) L=24 m=128")
    runs = []
    for planes in (1, 0):
        ctx.set_f16_presplit(planes)
        ctx.set_svd_backend(5)
        try:
            psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
            params = pkg.algvars(psi)
            pkg.dmrg(psi, mpo, m=128, sweeps=3, cutoff=1e-12, params=params, precision="f16x3",
                     correction="bond", env_precision="f64")
        finally:
            ctx.set_f16_presplit(1)
            ctx.set_svd_backend(0)
        say(" planes", planes, "E %.15f" % params.energy)
        runs.append(params.history)
    for a, b in zip(runs[0], runs[1]):
        assert a["energy"] == b["energy"]
        assert a["maxtrunc"] == b["maxtrunc"]
        assert list(a["storeD"]) == list(b["storeD"])
