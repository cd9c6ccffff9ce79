"""dmrjtensor.jl_amd: MI355X-native DMRG hot path (two-site and 3S).

Host-side mirror (Python) of the reference's API for the two-site DMRG path
(bakerte/DMRJtensor.jl: algs/DMRG.jl, algs/optimizeMPS.jl, lib/MPO_MPS,
lib/Env, lib/models.jl).  All arithmetic runs in libdmrjhip (HIP/CDNA4
kernels behind include/dmrj.h); this package only plans and moves handles.
The directory name contains a dot, so it is loaded with ``load_package()``
from ``build.py`` or imported as ``dmrjtensor_jl_amd`` after registration.
"""
from . import _lib
from ._lib import Context, DMRJError, default_context
from .models import fermionOps, heisenbergMPO, hubbardMPO, j1j2_bulk, spinOps, XXZ
from .network import (MPO, MPS, DeviceTensor, Lupdate, Rupdate, applyOps, environment, expect, makeEnv,
                      makeMPO, makeq, mpo_deltas, product_state, randMPS_normal, site2_from_dense,
                      site_from_dense)
from .dmrg import (absorb3s, algvars, alpha_update, dmrg, dmrg3S, expand3s, lanczos1, lanczos_k, optmps,
                   singlesite_update,
                   startval, step3S, twostep)
from .tdvp import krylovTimeEvol1, krylovTimeEvol2, local_expect, tdvp_twosite
from .measure import correlation, correlationmatrix
from .checkpoint import load_mps, save_mps
from .autompo import autoMPO, mpoterm

__all__ = [
    "Context", "DMRJError", "default_context", "applyOps", "fermionOps", "heisenbergMPO", "hubbardMPO",
    "j1j2_bulk", "spinOps", "XXZ", "MPO", "MPS", "DeviceTensor", "Lupdate", "Rupdate",
    "environment", "expect", "makeEnv", "makeMPO", "makeq", "mpo_deltas", "product_state",
    "randMPS_normal", "site2_from_dense", "site_from_dense", "algvars", "dmrg", "optmps",
    "singlesite_update", "startval", "twostep", "dmrg3S", "step3S", "lanczos1", "lanczos_k", "expand3s", "absorb3s",
    "alpha_update", "krylovTimeEvol1", "krylovTimeEvol2", "local_expect",
    "tdvp_twosite", "correlation", "correlationmatrix", "save_mps", "load_mps", "autoMPO", "mpoterm",
]
