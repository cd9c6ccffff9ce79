"""The C-ABI library loads and exports every symbol include/dmrj.h declares;
host-side helpers behave (CPU only, no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dmrj.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dmrj_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for must in ["dmrj_ctx_create", "dmrj_heff2_apply", "dmrj_lanczos2", "dmrj_env_left",
                 "dmrj_env_right", "dmrj_svd_trunc2", "dmrj_twostep", "dmrj_last_error"]:
        assert must in syms


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg._lib.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert lib.dmrj_version() >= 100
    # ctypes signatures cover the whole header
    for s in declared_symbols():
        if s != "dmrj_last_error":
            assert s in pkg._lib.SIGNATURES, s


def test_ctx_create_without_gpu_fails_cleanly(pkg):
    lib = pkg._lib.load()
    h = ctypes.c_void_p()
    import subprocess
    import sys
    # no GPU in this container: creation must return an error code, not crash
    code = ("import ctypes,sys; lib=ctypes.CDLL(%r); h=ctypes.c_void_p(); "
            "rc=lib.dmrj_ctx_create(0, ctypes.byref(h)); print(rc)" % pkg._lib.LIB_PATH)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0
    rc = int(out.stdout.strip().splitlines()[-1])
    if rc == 0:
        pytest.skip("a GPU is present")
    assert rc < 0


def test_null_arguments_rejected(pkg):
    lib = pkg._lib.load()
    assert lib.dmrj_ctx_create(0, None) == -1
    assert lib.dmrj_tensor_info(None, None) == -1
    assert lib.dmrj_ctx_sync(None) == -1


def test_host_mpo_matches_oracle(pkg):
    from oracle import models as om
    a = pkg.makeMPO(pkg.heisenbergMPO(), 2, 6)
    b = om.makeMPO(om.heisenbergMPO(), 2, 6)
    for x, y in zip(a.H, b):
        assert np.array_equal(x, y)
    a = pkg.makeMPO(pkg.hubbardMPO(), 4, 5)
    b = om.makeMPO(om.hubbardMPO(), 4, 5)
    for x, y in zip(a.H, b):
        assert np.array_equal(x, y)
    mats, bonds = pkg.j1j2_bulk(3, 4)
    ob, obonds = om.j1j2_cylinder_mpo(3, 4)
    assert bonds == obonds
    for x, y in zip(pkg.makeMPO(mats, 2).H, ob):
        assert np.array_equal(x, y)


def test_host_mpo_deltas(pkg):
    from oracle import mps as omps
    from oracle import models as om
    q = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
    a = pkg.mpo_deltas(pkg.makeMPO(pkg.hubbardMPO(), 4, 6).H, q)
    b = omps.mpo_deltas(om.makeMPO(om.hubbardMPO(), 4, 6), q)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
