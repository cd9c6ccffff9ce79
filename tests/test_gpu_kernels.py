"""Kernel self-tests through the C ABI: MFMA FP64 lane layout and the grouped
GEMM (all tile shapes, transposes, ragged sizes) against numpy FP64."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mfma_f64_lane_layout(pkg, ctx):
    out = np.zeros(768)
    ctx.call("dmrj_selftest_mfma_f64", pkg._lib.dptr(out))
    rows = out[:256].reshape(64, 4)      # a = lane+1, b = 1: row sums of A
    cols = out[256:512].reshape(64, 4)   # a = 1, b = lane+1: column sums of B
    enc = out[512:].reshape(64, 4)
    # operands: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k;
    # result: item r of lane l holds D[4r + l/16][l%16]
    want_rows = np.zeros((64, 4))
    want_cols = np.zeros((64, 4))
    want = np.zeros((64, 4))
    for l in range(64):
        for r in range(4):
            i, j = 4 * r + l // 16, l % 16
            want_rows[l, r] = 4 * i + 100
            want_cols[l, r] = 4 * j + 100
            want[l, r] = (i + 1) + 1000 * (j + 1)
    assert np.array_equal(rows, want_rows)
    assert np.array_equal(cols, want_cols)
    assert np.array_equal(enc, want)


@pytest.mark.parametrize("cfg", [0, 1, 2])
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (33, 70, 17), (130, 129, 200), (5, 300, 0)])
def test_grouped_gemm_f64(pkg, ctx, cfg, tA, tB, M, N, K):
    say("gemm cfg", cfg, "tA", tA, "tB", tB, "MNK", M, N, K)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    As = np.asfortranarray(A.T if tA else A)
    Bs = np.asfortranarray(B.T if tB else B)
    C = np.zeros((M, N), order="F")
    ctx.call("dmrj_selftest_gemm", cfg, M, N, K, tA, tB, pkg._lib.dptr(As.ravel(order="F")),
             pkg._lib.dptr(Bs.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
    want = A @ B
    err = np.max(np.abs(C - want)) / max(1.0, np.max(np.abs(want)))
    if err > 1e-13 and M == N:
        say(" err", err, "vs transpose", np.max(np.abs(C - want.T)), "zero?", np.max(np.abs(C)))
    assert err <= 1e-13, err
