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


def test_makeq_host_assigns_charges_drops_and_randomises(pkg):
    """makeqMPS's charge assignment (assignflux!: the charge of the largest
    entry of every right-link column, first maximum with the physical index
    innermost), the dropping of entries off the charge blocks, and the
    refilling of a tensor left empty, on hand-labelled MPS (CPU only)."""
    from dmrjtensor_jl_amd.network import makeq_host
    q = [[1], [-1]]
    # Neel product state: bond charges are the running sums
    A = [np.zeros((1, 2, 1)) for _ in range(4)]
    for i in range(4):
        A[i][0, i % 2, 0] = 1.0
    t, qb = makeq_host(A, q)
    assert [int(x[0, 0]) for x in qb] == [0, 1, 0, 1, 0]
    assert all(np.array_equal(a, b) for a, b in zip(t, A))
    # two right-link states; column 0 is dominated by (a=0, s=1), column 1 by
    # (a=0, s=0); a small entry of column 0 at (a=0, s=0) breaks the charge
    # rule and is dropped
    B = np.zeros((1, 2, 2))
    B[0, 1, 0] = 0.9
    B[0, 0, 0] = 0.1
    B[0, 0, 1] = 0.5
    t, qb = makeq_host([B], q)
    assert qb[1].ravel().tolist() == [-1, 1]
    assert t[0][0, 0, 0] == 0.0 and t[0][0, 1, 0] == 0.9 and t[0][0, 0, 1] == 0.5
    # ties go to the first entry with the left index outer and the physical
    # index inner: (a=0, s=1) before (a=1, s=0), whose charges differ here
    T = np.zeros((2, 2, 1))
    T[0, 1, 0] = 1.0
    T[1, 0, 0] = 1.0
    t, qb = makeq_host([T], q, qleft=[[0], [0]])
    assert int(qb[1][0, 0]) == -1
    # an all-zero tensor is refilled on its allowed blocks only
    Z = np.zeros((2, 2, 1))
    t, qb = makeq_host([np.ones((1, 2, 2)) * [[[1.0, 0.0], [0.0, 1.0]]], Z], q, seed=3)
    ok = (qb[1][:, None, None, 0] + np.array([1, -1])[None, :, None]) == qb[2][None, None, :, 0]
    assert np.any(t[1][ok] != 0.0) and np.all(t[1][~ok] == 0.0)


def test_load_mps_rejects_corrupted_archives(pkg, tmp_path):
    """A truncated or inconsistent checkpoint raises ValueError before any
    array reaches the library (CPU only: no device call happens)."""
    import json as _json
    FORMAT = pkg.checkpoint.FORMAT if hasattr(pkg, "checkpoint") else "dmrj-mps-1"

    def write(path, qL_rows=2, qP_rows=2, data_n=8, info_n=8, nq_site=1, nq_head=1):
        head = {"format": FORMAT, "L": 1, "oc": 1, "nq": nq_head, "d": [2], "qphys": [[1], [-1]], "extra": {}}
        info = np.array([nq_site, 2, 2, 2, 8, 0, 0, 0][:info_n], dtype=np.int64)
        arrays = {"header": np.frombuffer(_json.dumps(head).encode(), dtype=np.uint8), "A1_info": info,
                  "A1_qL": np.zeros((qL_rows, 1), np.int32), "A1_qP": np.zeros((qP_rows, 1), np.int32),
                  "A1_qR": np.zeros((2, 1), np.int32), "A1_data": np.zeros(data_n)}
        with open(path, "wb") as f:
            np.savez(f, **arrays)
        return path

    load = pkg.load_mps
    for kw in ({"qL_rows": 1}, {"qP_rows": 3}, {"data_n": 7}, {"info_n": 4}, {"nq_site": 2}):
        p = write(str(tmp_path / "bad.npz"), **kw)
        with pytest.raises(ValueError):
            load(p)


def test_block_jacobi_panel_width(pkg):
    """The block-Jacobi panel (rows x 2bw doubles of dynamic LDS) at the edge
    of the 160 KiB work-group LDS with 256 bytes of static LDS: 639 rows fit
    width 16, 640 rows need width 8, 1280 rows width 4 (host arithmetic)."""
    lib = pkg._lib.load()
    lds, st = 160 * 1024, 256
    assert lib.dmrj_bj_panel_width(639, lds, st) == 16
    assert lib.dmrj_bj_panel_width(640, lds, st) == 8
    assert lib.dmrj_bj_panel_width(1280, lds, st) == 4
    assert lib.dmrj_bj_panel_width(5200, lds, st) == 0
    assert lib.dmrj_bj_panel_width(640, lds, 0) == 16
