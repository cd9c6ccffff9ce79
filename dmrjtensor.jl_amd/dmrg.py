"""Two-site and 3S DMRG drivers on the GPU path.

Mirrors algs/DMRG.jl and algs/optimizeMPS.jl:
  dmrg           algs/DMRG.jl:549-589 (method dispatch: "twosite", "3S")
  dmrg_twosite   algs/DMRG.jl:675-691
  dmrg3S         algs/DMRG.jl:655-671 (nsites = 1: the same bond pairs)
  step3S         algs/DMRG.jl:322-360 (dmrj_lanczos1 -> dmrj_expand3s ->
                 dmrj_normalize -> dmrj_svd_trunc2 -> dmrj_absorb3s)
  alpha_update   algs/DMRG.jl:288-303 (noise control, on the host)
  optmps         algs/optimizeMPS.jl:759-849 (sweep driver, plugin kwargs)
  algvars        algs/optimizeMPS.jl:36-106, defaults :514-577, SvNbond :614
  twostep        algs/DMRG.jl:362-405 (one C-ABI call: normalise, lanczos2,
                 truncated SVD; the reference's unused twosite_update is not
                 computed)
  SvNcheck       algs/DMRG.jl:172-210
  dmrgcvg        algs/DMRG.jl:610-652
  startval       algs/optimizeMPS.jl:255-258 (single-site H_eff at the centre)

The host loop only moves handles; all arithmetic runs in libdmrjhip.
"""
import ctypes as C
import math
import time

import numpy as np

from . import _lib
from .network import DeviceTensor, Lupdate, Rupdate, makeEnv


class algvars:
    """Parameter/state record of the two-site path (field names of the
    reference's algvars)."""

    def __init__(self, psi, maxm=0, minm=2, sweeps=1, cutoff=0.0, goal=0.0, silent=True,
                 SvNbond=None, precision="f64", correction="none", env_precision=None):
        self.maxm = maxm
        self.minm = minm
        self.sweeps = sweeps
        self.cutoff = cutoff
        self.goal = goal
        self.silent = silent
        self.SvNbond = (len(psi) + 1) // 2 if SvNbond is None else SvNbond
        self.precision = precision
        # arithmetic of the environment updates (None: same as precision)
        self.env_precision = env_precision
        # FP64 energy correction of split-precision runs: "none", "bond" (every
        # bond step) or "sweep" (the last bond step of each sweep, whose energy
        # is the one a sweep reports)
        self.correction = correction
        self.correct_now = False
        self.residual = None
        self.energy = 0.0
        self.lastenergy = 0.0
        self.truncerr = 0.0
        self.maxtrunc = 0.0
        self.biggestm = 0
        self.entropy = 0.0
        self.lastentropy = 0.0
        self.storeD = [0.0] * 8
        self.maxshowD = 8
        self.cvgE = True
        self.efficient = False
        self.Lenv = None
        self.Renv = None
        # 3S: Krylov vectors per site and the noise (subspace expansion)
        # control of algvars (algs/optimizeMPS.jl:528-537)
        self.method = "twosite"
        self.r = 2
        self.noise = 1.0
        self.noise_goal = 0.3
        self.noise_incr = 0.01
        self.history = []
        self.sweep_times = []


def singlesite_update(Lenv, Renv, A, mpo, i, precision="f64"):
    """H_1site A (algs/optimizeMPS.jl:503-511)."""
    ctx = A.ctx
    h = C.c_void_p()
    ctx.call("dmrj_heff1_apply", Lenv.h, mpo.device(ctx, i), Renv.h, A.h, _lib.PREC[precision],
             C.byref(h))
    return DeviceTensor(ctx, h)


def startval(Lenv, Renv, psi, mpo, i):
    """<psi_oc|H_1site|psi_oc> (algs/optimizeMPS.jl:255-258)."""
    H = singlesite_update(Lenv, Renv, psi[i], mpo, i)
    e = psi[i].dot(H)
    H.free()
    return e


def SvNcheck(i, j, svals, Ns, params):
    """algs/DMRG.jl:172-210 (allSvNbond = false)."""
    if params.maxtrunc < params.truncerr:
        params.maxtrunc = params.truncerr
    if params.biggestm < len(svals):
        params.biggestm = len(svals)
    if i == min(max(2, params.SvNbond + 1), Ns) and j < 0:
        D = np.sort(np.asarray(svals))[::-1]
        p = D ** 2
        params.entropy = float(-np.sum(np.where(p > 0, p * np.log(np.where(p > 0, p, 1.0)), 0.0)))
        params.storeD = [float(x) for x in D[:min(len(D), params.maxshowD)]]


def twostep(n, j, i, iL, iR, psi, mpo, Lenv, Renv, params):
    """One bond step (algs/DMRG.jl:362-405) through dmrj_twostep."""
    ctx = psi.ctx
    cap = max(8, 4 * (params.maxm if params.maxm > 0 else 4096))
    prec = _lib.PREC[params.precision]
    if params.precision != "f64" and (params.correction == "bond" or
                                      (params.correction == "sweep" and params.correct_now)):
        prec |= _lib.CORRECT
    sv = np.zeros(cap)
    nl, nr = C.c_void_p(), C.c_void_p()
    e, te = C.c_double(), C.c_double()
    kept = C.c_int64()
    ctx.call("dmrj_twostep", Lenv[iL].h, Renv[iR].h, psi[iL].h, psi[iR].h, mpo.device(ctx, iL),
             mpo.device(ctx, iR), 1 if j > 0 else -1, int(params.maxm), int(params.minm),
             float(params.cutoff), prec, C.byref(nl), C.byref(nr), C.byref(e),
             C.byref(te), C.byref(kept), _lib.dptr(sv), cap)
    psi[iL].free()
    psi[iR].free()
    psi[iL] = DeviceTensor(ctx, nl)
    psi[iR] = DeviceTensor(ctx, nr)
    params.energy = e.value
    if prec & _lib.CORRECT:
        r = C.c_double()
        ctx.call("dmrj_ctx_last_residual", C.byref(r))
        params.residual = r.value
    params.truncerr = te.value
    params.biggestm = max(params.biggestm, int(kept.value))
    SvNcheck(i, j, sv[:min(cap, kept.value)], len(psi), params)


def lanczos1(Lenv, Renv, A, mpo, i, r, precision="f64"):
    """TensorPACK lanczos(A, mpo[i]; r, m=1, updatefct=singlesite_update) as
    step3S calls it (algs/DMRG.jl:328): (Ritz vector, lowest Ritz value)."""
    ctx = A.ctx
    h = C.c_void_p()
    e = C.c_double()
    k = C.c_int()
    ctx.call("dmrj_lanczos1", Lenv.h, Renv.h, A.h, mpo.device(ctx, i), int(r), _lib.PREC[precision] & 15,
             C.byref(h), C.byref(e), C.byref(k))
    return DeviceTensor(ctx, h), e.value


def lanczos_k(Lenv, Renv, x, mpo, i, r, precision="f64"):
    """k-vector Lanczos ground state (lowest Ritz pair of the r-vector Krylov
    space) of H_eff on a one-site tensor at site i or a two-site tensor at
    the bond (i, i+1): simplelanczos (algs/DMRG.jl:57-90) with r vectors.
    Returns (Ritz vector, energy, Krylov dimension used)."""
    ctx = x.ctx
    h = C.c_void_p()
    e = C.c_double()
    k = C.c_int()
    two = x.kind == 1
    ctx.call("dmrj_lanczos_k", Lenv.h, Renv.h, x.h, mpo.device(ctx, i), mpo.device(ctx, i + 1) if two else None,
             int(r), _lib.PREC[precision] & 15, C.byref(h), C.byref(e), C.byref(k))
    return DeviceTensor(ctx, h), e.value, k.value


def expand3s(E, A, mpo, i, alpha, direction):
    """Lexpand (direction 1, E = Lenv[i]) / Rexpand (direction -1, E =
    Renv[i]) joined to A (algs/DMRG.jl:273-286); alpha = 0: A alone."""
    ctx = A.ctx
    h = C.c_void_p()
    ctx.call("dmrj_expand3s", E.h, A.h, mpo.device(ctx, i), float(alpha), int(direction), C.byref(h))
    return DeviceTensor(ctx, h)


def absorb3s(F, N, direction):
    """New centre of moveR!/moveL! with condition = true
    (lib/MPS/move.jl:124-139,175-190)."""
    ctx = F.ctx
    h = C.c_void_p()
    ctx.call("dmrj_absorb3s", F.h, N.h, int(direction), C.byref(h))
    return DeviceTensor(ctx, h)


def alpha_update(alpha, truncerr, cutoff, lastEnergy, currenergy, noise_goal, dalpha, avgE=None):
    """algs/DMRG.jl:288-303, with IEEE division as in Julia (x/0 = +-Inf,
    0/0 = NaN, which compares false)."""
    avgE = currenergy if avgE is None else avgE
    trunc_check = truncerr <= max(cutoff, 1e-12)
    errEst = abs(currenergy - lastEnergy) <= max(abs(currenergy * cutoff), 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.float64(lastEnergy) / np.float64(currenergy) - 1.0
        sitecond = bool(np.abs(np.float64(truncerr * avgE) / ratio) > noise_goal)
    if currenergy <= lastEnergy and (sitecond or errEst or trunc_check):
        alpha *= 1.0 - dalpha
    else:
        alpha *= 1.0 + dalpha
    return max(min(alpha, 1.0), 1e-42)


def step3S(n, j, i, iL, iR, psi, mpo, Lenv, Renv, params):
    """One site of 3S (algs/DMRG.jl:322-360): single-site Lanczos, subspace
    expansion while the noise exceeds noise_incr, truncated SVD of the
    expanded tensor (scaled to unit norm) and the move of the centre."""
    A, energy = lanczos1(Lenv[i], Renv[i], psi[i], mpo, i, params.r, params.precision)
    params.energy = energy
    noise = params.noise
    alpha = noise if noise > params.noise_incr else 0.0
    direction = 1 if j > 0 else -1
    T = expand3s(Lenv[iL] if j > 0 else Renv[iR], A, mpo, i, alpha, direction)
    A.free()
    T.ctx.call("dmrj_normalize", T.h)
    ctx = psi.ctx
    cap = max(8, 4 * (params.maxm if params.maxm > 0 else 4096))
    sv = np.zeros(cap)
    nl, nr = C.c_void_p(), C.c_void_p()
    te = C.c_double()
    kept = C.c_int64()
    ctx.call("dmrj_svd_trunc2", T.h, int(params.maxm), int(params.minm), float(params.cutoff), direction,
             C.byref(nl), C.byref(nr), C.byref(te), C.byref(kept), _lib.dptr(sv), cap)
    T.free()
    left, right = DeviceTensor(ctx, nl), DeviceTensor(ctx, nr)
    if j > 0:
        newR = absorb3s(right, psi[iR], 1)
        right.free()
        psi[iL].free()
        psi[iR].free()
        psi[iL], psi[iR] = left, newR
    else:
        newL = absorb3s(left, psi[iL], -1)
        left.free()
        psi[iL].free()
        psi[iR].free()
        psi[iL], psi[iR] = newL, right
    params.noise = alpha_update(noise, te.value, params.cutoff, params.lastenergy, params.energy,
                                params.noise_goal, params.noise_incr)
    params.truncerr = te.value
    params.biggestm = max(params.biggestm, int(kept.value))
    SvNcheck(i, j, sv[:min(cap, kept.value)], len(psi), params)


def dmrgcvg(n, timer, params):
    """algs/DMRG.jl:610-652 (cvgE = true)."""
    if not params.silent:
        print("Sweep %d (back and forth): %.5f sec" % (n, timer))
        print("Largest truncation = %s, m = %d" % (params.maxtrunc, params.biggestm))
        print("Energy at sweep %d is %.12f" % (n, params.energy))
        print(flush=True)
    params.history.append(dict(sweep=n, energy=float(params.energy), maxtrunc=float(params.maxtrunc),
                               biggestm=int(params.biggestm), entropy=float(params.entropy),
                               storeD=list(params.storeD), time=timer, precision=params.precision,
                               residual=params.residual, noise=float(params.noise)))
    params.biggestm = 0
    params.truncerr = params.maxtrunc
    params.maxtrunc = 0.0
    checkE = abs(params.energy - params.lastenergy)
    if checkE < params.goal:
        return True
    params.lastenergy = params.energy
    return False


def optmps(psi, mpo, params, stepfct=twostep, cvgfct=dmrgcvg, measfct=startval):
    """Sweep driver (algs/optimizeMPS.jl:759-849) with the same pluggable
    step / convergence / measurement functions."""
    Ns = len(psi)
    if params.Lenv is None or params.Renv is None:
        params.Lenv, params.Renv = makeEnv(psi, mpo)
    Lenv, Renv = params.Lenv, params.Renv
    if params.energy == 0.0 and measfct is not None:
        params.energy = measfct(Lenv[psi.oc], Renv[psi.oc], psi, mpo, psi.oc)
    params.lastenergy = params.energy
    Nsteps = 2 * (Ns - 1)
    if psi.oc == 1:
        j = 1
    elif psi.oc == Ns:
        j = -1
    else:
        j = 1
    envp = params.env_precision or params.precision
    psi.ctx.call("dmrj_ctx_set_env_precision", _lib.PREC[envp] & 15)
    try:
        _sweeps(psi, mpo, params, stepfct, cvgfct, Lenv, Renv, Nsteps, j)
    finally:
        psi.ctx.call("dmrj_ctx_set_env_precision", 0)
    return params.energy


def _sweeps(psi, mpo, params, stepfct, cvgfct, Lenv, Renv, Nsteps, j):
    Ns = len(psi)
    for n in range(1, params.sweeps + 1):
        t0 = time.perf_counter()
        for ns in range(Nsteps):
            params.correct_now = ns == Nsteps - 1
            i = psi.oc
            iL, iR = (i, i + 1) if j > 0 else (i - 1, i)
            stepfct(n, j, i, iL, iR, psi, mpo, Lenv, Renv, params)
            if j > 0:
                Lenv[i + 1] = Lupdate(Lenv[i], psi[i], mpo, i)
            else:
                Renv[i - 1] = Rupdate(Renv[i], psi[i], mpo, i)
            psi.oc += j
            if (psi.oc == Ns and j > 0) or (psi.oc == 1 and j < 0):
                j = -j
        psi.ctx.sync()
        timer = time.perf_counter() - t0
        params.sweep_times.append(timer)
        if cvgfct(n, timer, params):
            break


def dmrg(psi, mpo, m=0, minm=2, sweeps=1, cutoff=0.0, silent=True, goal=0.0, method="twosite",
         SvNbond=None, precision="f64", correction=None, params=None, env_precision=None, r=None,
         noise=None, noise_goal=None, noise_incr=None):
    """dmrg(psi, mpo; m, minm, sweeps, cutoff, silent, goal, method)
    (algs/DMRG.jl:549-589).  Optimises ``psi`` in place and returns the
    energy; the per-sweep report is in ``params.history`` when a params
    object is passed.  method "3S" runs dmrg3S (r, noise, noise_goal,
    noise_incr as in the reference's keywords)."""
    if method not in ("twosite", "3S"):
        raise ValueError("DMRG method (%s) not defined on the GPU path" % method)
    if params is None:
        params = algvars(psi)
    params.method = method
    for name, val in (("r", r), ("noise", noise), ("noise_goal", noise_goal), ("noise_incr", noise_incr)):
        if val is not None:
            setattr(params, name, val)
    if method == "3S" and precision != "f64":
        raise ValueError("3S runs in FP64 on the GPU path")
    params.maxm, params.minm, params.sweeps, params.cutoff = m, minm, sweeps, cutoff
    params.silent, params.goal, params.precision = silent, goal, precision
    if precision not in ("f64", "bf16x3", "f16x3"):
        raise ValueError("precision must be 'f64', 'bf16x3' or 'f16x3'")
    if env_precision not in (None, "f64", "bf16x3", "f16x3"):
        raise ValueError("env_precision must be None, 'f64', 'bf16x3' or 'f16x3'")
    params.env_precision = env_precision
    params.correction = correction if correction is not None else ("none" if precision == "f64" else "sweep")
    if SvNbond is not None:
        params.SvNbond = SvNbond
    return optmps(psi, mpo, params, stepfct=step3S if method == "3S" else twostep)


def dmrg3S(psi, mpo, **kw):
    """dmrg3S (algs/DMRG.jl:655-671)."""
    return dmrg(psi, mpo, method="3S", **kw)
