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


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
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


@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(33, 70, 17), (130, 129, 200), (64, 31, 333)])
def test_gemm_shapes_same_bits(pkg, ctx, tA, tB, M, N, K):
    """Every tile shape accumulates each output element in the same K order:
    identical bits."""
    say("gemm bits tA", tA, "tB", tB, "MNK", M, N, K)
    rng = np.random.default_rng(M + 5 * N + 11 * K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    As = np.asfortranarray(A.T if tA else A)
    Bs = np.asfortranarray(B.T if tB else B)
    out = []
    for cfg in range(5):
        C = np.zeros((M, N), order="F")
        ctx.call("dmrj_selftest_gemm", cfg, M, N, K, tA, tB, pkg._lib.dptr(As.ravel(order="F")),
                 pkg._lib.dptr(Bs.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
        out.append(C)
    for cfg in range(1, 5):
        assert np.array_equal(out[0], out[cfg]), cfg


@pytest.mark.parametrize("cfg,M,N,K,prec", [(0, 1600, 1600, 40, 0), (1, 2100, 2100, 24, 0), (0, 1600, 1600, 40, 2),
                                            (1, 2100, 2100, 24, 1)])
def test_gemm_tile_orders_more_tiles_than_workgroups(pkg, ctx, cfg, M, N, K, prec):
    """More tiles than resident work-groups (several tiles per work-group of
    the persistent grid): both tile orders (0 descending work, 1 XCD streams)
    run every tile exactly once and give the same bits, and the result
    matches numpy."""
    say("tile orders cfg", cfg, "MNK", M, N, K, "precision", prec)
    rng = np.random.default_rng(M + K)
    A = np.asfortranarray(rng.standard_normal((M, K)))
    B = np.asfortranarray(rng.standard_normal((K, N)))
    out = []
    try:
        for order in (0, 1):
            ctx.set_tile_order(order)
            C = np.zeros((M, N), order="F")
            ctx.call("dmrj_selftest_gemm", cfg | (prec << 4), M, N, K, 0, 0, pkg._lib.dptr(A.ravel(order="F")),
                     pkg._lib.dptr(B.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
            out.append(C)
    finally:
        ctx.set_tile_order(1)
    want = A @ B
    # FP64: rounding level; split forms: their error bounds relative to |A||B|
    tol = {0: 1e-13, 1: 6e-5, 2: 5e-6}[prec]
    assert np.max(np.abs(out[0] - want) / (np.abs(A) @ np.abs(B))) <= tol
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("backend", [0, pytest.param(1, marks=pytest.mark.xfail(
    reason="rocSOLVER gesdd loses the small singular values of graded spectra; kept only as a "
           "selectable backend, never the default", strict=False)), 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(24, 40), (64, 50), (128, 128), (90, 200), (400, 360)])
def test_svd_backends_graded(pkg, ctx, backend, shape):
    """Two-site split of a graded tensor (singular values 1 .. 1e-9): every SVD
    backend's spectrum vs LAPACK (numpy), absolute error relative to sigma_max
    and relative error of the kept values above 1e-6."""
    from oracle import dmrg as od
    R, C = shape
    rng = np.random.default_rng(R + C)
    k = min(R, C)
    U, _ = np.linalg.qr(rng.standard_normal((R, k)))
    V, _ = np.linalg.qr(rng.standard_normal((C, k)))
    s = np.logspace(0, -9, k)
    M = (U * s) @ V.T
    M /= np.linalg.norm(M)
    AA = M.reshape(R // 2, 2, 2, C // 2, order="F")
    t = pkg.site2_from_dense(ctx, AA)
    ctx.set_svd_backend(backend)
    try:
        import ctypes as Ct
        l, r = Ct.c_void_p(), Ct.c_void_p()
        te = Ct.c_double()
        kk = Ct.c_int64()
        sv = np.zeros(4096)
        ctx.call("dmrj_svd_trunc2", t.h, 0, 2, 0.0, 1, Ct.byref(l), Ct.byref(r), Ct.byref(te), Ct.byref(kk),
                 pkg._lib.dptr(sv), 4096)
    finally:
        ctx.set_svd_backend(0)
    sv = sv[:kk.value]
    want = np.linalg.svd(AA.reshape(R, C, order="F"), compute_uv=False)
    want = want / np.linalg.norm(want)
    n = min(len(sv), len(want))
    abs_err = np.max(np.abs(sv[:n] - want[:n]))
    big = want[:n] > 1e-6
    rel_err = np.max(np.abs(sv[:n][big] - want[:n][big]) / want[:n][big])
    say(" backend", backend, "shape", shape, "abs/sigma_max %.2e" % abs_err, "rel(sigma>1e-6) %.2e" % rel_err)
    if backend == 5:
        # Gram eigensolve: sigma^2 to eps |M|^2 absolute, sigma to eps |M|^2 / sigma
        big = want[:n] > 1e-3
        assert abs_err < 5e-7
        assert np.max(np.abs(sv[:n][big] - want[:n][big]) / want[:n][big]) < 1e-9
        return
    # one-sided Jacobi accumulates O(n eps) rounding over its rotations
    assert abs_err < (1e-13 if backend == 4 else 1e-14)


def _split(pkg, ctx, t, m, direction, cap=4096):
    import ctypes as Ct
    l, r = Ct.c_void_p(), Ct.c_void_p()
    te = Ct.c_double()
    kk = Ct.c_int64()
    sv = np.zeros(cap)
    ctx.call("dmrj_svd_trunc2", t.h, m, 2, 0.0, direction, Ct.byref(l), Ct.byref(r), Ct.byref(te), Ct.byref(kk),
             pkg._lib.dptr(sv), cap)
    return pkg.DeviceTensor(ctx, l), pkg.DeviceTensor(ctx, r), te.value, sv[:kk.value]


@pytest.mark.parametrize("direction", [1, -1])
@pytest.mark.parametrize("nq", [0, 1])
def test_svd_block_jacobi_split(pkg, ctx, direction, nq):
    """Backend 4 (block one-sided Jacobi + GEMM for the other factor) against
    numpy per sector: kept spectrum, truncation error, orthonormal factor on
    the side the direction asks for, product equal to the best rank-k
    approximation (renormalised)."""
    from oracle import dmrg as od
    rng = np.random.default_rng(17 + nq + direction)
    if nq == 0:
        mL, mR = 150, 170
        qL = qP = qR = None
    else:
        qL = np.array([[-2]] * 30 + [[0]] * 130 + [[2]] * 40)
        qP = np.array([[1], [-1]])
        qR = np.array([[-2]] * 50 + [[0]] * 120 + [[2]] * 35)
        mL, mR = len(qL), len(qR)
    AA = rng.standard_normal((mL, 2, 2, mR))
    if nq:
        for a in range(mL):
            for s1 in range(2):
                for s2 in range(2):
                    bad = qL[a, 0] + qP[s1, 0] + qP[s2, 0] != qR[:, 0]
                    AA[a, s1, s2, bad] = 0.0
    # graded spectrum: scale by a decaying profile along the columns
    AA *= np.logspace(0, -6, mR)[None, None, None, :]
    AA /= np.linalg.norm(AA)
    t = pkg.site2_from_dense(ctx, AA, qL, qP, qP, qR, nq=nq)
    for m in (0, 120):
        ctx.set_svd_backend(4)
        try:
            gl, gr, te, sv = _split(pkg, ctx, t, m, direction)
        finally:
            ctx.set_svd_backend(0)
        U, s, V, terr, _ = od.fsvd(AA, m, 2, 0.0, qL, qP, qR)
        say(" bj nq", nq, "dir", direction, "m", m, "kept", len(sv), "te %.3e %.3e" % (te, terr))
        assert len(sv) == len(s)
        # the C ABI returns the kept values sorted; the oracle keeps sector order
        assert np.allclose(sv, np.sort(s)[::-1], rtol=1e-11, atol=1e-14)
        assert abs(te - terr) <= 1e-10 * max(terr, 1e-300) + 1e-15
        A = gl.to_dense(nq)[0]
        B = gr.to_dense(nq)[0]
        k = A.shape[2]
        if direction > 0:
            X = A.reshape(-1, k, order="F")
            assert np.allclose(X.T @ X, np.eye(k), atol=1e-12)
        else:
            X = B.reshape(k, -1, order="F")
            assert np.allclose(X @ X.T, np.eye(k), atol=1e-12)
        prod = np.tensordot(A, B, axes=(2, 0))
        want = np.tensordot(U, s[:, None, None] * V, axes=(2, 0))
        err = np.max(np.abs(prod - want)) / np.max(np.abs(want))
        say("  product err %.2e" % err)
        assert err <= 1e-11


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (33, 70, 17), (130, 129, 200), (1, 300, 65)])
def test_grouped_gemm_f16x3_planes_same_bits(pkg, ctx, cfg, tA, tB, M, N, K):
    """FP16x3 on pre-split operand planes (default) and with the split done by
    every tile on load: the same scaled halves, K steps and MFMA sequence, so
    the same bits, for every tile shape, transpose and ragged edge."""
    rng = np.random.default_rng(M + 3 * N + 7 * K + 11 * tA + 13 * tB)
    A = rng.standard_normal((M, K)) * 1e3
    B = rng.standard_normal((K, N)) * 1e-4
    As = np.asfortranarray(A.T if tA else A)
    Bs = np.asfortranarray(B.T if tB else B)
    out = []
    for planes in (1, 0):
        ctx.set_f16_presplit(planes)
        try:
            C = np.zeros((M, N), order="F")
            ctx.call("dmrj_selftest_gemm", cfg | (2 << 4), M, N, K, tA, tB, pkg._lib.dptr(As.ravel(order="F")),
                     pkg._lib.dptr(Bs.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
        finally:
            ctx.set_f16_presplit(1)
        out.append(C)
    assert np.array_equal(out[0], out[1])
    err = np.max(np.abs(out[0] - A @ B) / (np.abs(A) @ np.abs(B)))
    assert err < 5e-6


@pytest.mark.parametrize("cfg", [0, 2, 3])
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K,scaleA,scaleB", [(33, 70, 17, 1.0, 1.0), (130, 129, 200, 3e-9, 2e6),
                                                 (64, 64, 300, 1e5, 1e-7)])
def test_grouped_gemm_f16x3_scaled(pkg, ctx, cfg, tA, tB, M, N, K, scaleA, scaleB):
    """FP16 split precision with per-block power-of-two scaling: operands far
    outside FP16's range still give a few 2^-22 of the |A||B| scale (FP32 accumulation)."""
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    A = rng.standard_normal((M, K)) * scaleA
    B = rng.standard_normal((K, N)) * scaleB
    As = np.asfortranarray(A.T if tA else A)
    Bs = np.asfortranarray(B.T if tB else B)
    C = np.zeros((M, N), order="F")
    ctx.call("dmrj_selftest_gemm", cfg | (2 << 4), M, N, K, tA, tB, pkg._lib.dptr(As.ravel(order="F")),
             pkg._lib.dptr(Bs.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
    want = A @ B
    scale = np.abs(A) @ np.abs(B)
    err = np.max(np.abs(C - want) / scale)
    say(" f16x3 cfg", cfg, tA, tB, M, N, K, scaleA, scaleB, "err/|A||B| %.2e" % err)
    assert err < 5e-6


@pytest.mark.parametrize("direction", [1, -1])
def test_svd_gram_split(pkg, ctx, direction):
    """Backend 5 (Gram eigensolve + GEMM for the other factor), U(1) tensor:
    the canonical factor is orthonormal to working precision, the product is
    the best rank-k approximation (renormalised) to 1e-10, the large
    singular values match LAPACK to 1e-10."""
    from oracle import dmrg as od
    rng = np.random.default_rng(29 + direction)
    qL = np.array([[-2]] * 30 + [[0]] * 130 + [[2]] * 40)
    qP = np.array([[1], [-1]])
    qR = np.array([[-2]] * 50 + [[0]] * 120 + [[2]] * 35)
    mL, mR = len(qL), len(qR)
    AA = rng.standard_normal((mL, 2, 2, mR))
    for a in range(mL):
        for s1 in range(2):
            for s2 in range(2):
                AA[a, s1, s2, qL[a, 0] + qP[s1, 0] + qP[s2, 0] != qR[:, 0]] = 0.0
    AA *= np.logspace(0, -4, mR)[None, None, None, :]
    AA /= np.linalg.norm(AA)
    t = pkg.site2_from_dense(ctx, AA, qL, qP, qP, qR, nq=1)
    ctx.set_svd_backend(5)
    try:
        gl, gr, te, sv = _split(pkg, ctx, t, 120, direction)
    finally:
        ctx.set_svd_backend(0)
    U, s, V, terr, _ = od.fsvd(AA, 120, 2, 0.0, qL, qP, qR)
    ss = np.sort(s)[::-1]
    big = ss > 1e-3 * ss[0]
    say(" gram dir", direction, "kept", len(sv), "te %.3e %.3e" % (te, terr))
    assert len(sv) == len(s)
    assert np.allclose(sv[big], ss[big], rtol=1e-10)
    A = gl.to_dense(1)[0]
    B = gr.to_dense(1)[0]
    k = A.shape[2]
    if direction > 0:
        X = A.reshape(-1, k, order="F")
        assert np.allclose(X.T @ X, np.eye(k), atol=1e-13)
    else:
        X = B.reshape(k, -1, order="F")
        assert np.allclose(X @ X.T, np.eye(k), atol=1e-13)
    prod = np.tensordot(A, B, axes=(2, 0))
    want = np.tensordot(U, s[:, None, None] * V, axes=(2, 0))
    assert np.max(np.abs(prod - want)) / np.max(np.abs(want)) <= 1e-10


@pytest.mark.parametrize("backend", [0, 5])
@pytest.mark.parametrize("direction", [1, -1])
@pytest.mark.parametrize("mL,mR,rank", [(60, 70, 3), (50, 50, 45), (300, 280, 5)])
def test_svd_rank_deficient_orthonormal_completion(pkg, ctx, backend, direction, mL, mR, rank):
    """A low-rank two-site tensor split with m = 0 and cutoff 0 keeps all
    min(R, C) values, most of them zero.  Their singular vectors are not
    determined by the data; the canonical factor must still be orthonormal
    (the reference's LAPACK SVD returns an orthonormal completion) and the
    product must equal the tensor.  The shapes need one to nineteen blocks of
    32 completed vectors, with the kept vectors fewer or more than a block."""
    rng = np.random.default_rng(41 + backend + direction + rank)
    M = rng.standard_normal((2 * mL, rank)) @ rng.standard_normal((rank, 2 * mR))
    AA = (M / np.linalg.norm(M)).reshape(mL, 2, 2, mR, order="F")
    t = pkg.site2_from_dense(ctx, AA)
    ctx.set_svd_backend(backend)
    try:
        gl, gr, te, sv = _split(pkg, ctx, t, 0, direction)
    finally:
        ctx.set_svd_backend(0)
    A = gl.to_dense()[0]
    B = gr.to_dense()[0]
    k = A.shape[2]
    say(" rank", rank, "shape", (mL, mR), "backend", backend, "dir", direction, "kept", k, "sv[:4]", sv[:4])
    assert k == 2 * min(mL, mR)
    if direction > 0:
        X = A.reshape(-1, k, order="F")
        err = np.max(np.abs(X.T @ X - np.eye(k)))
    else:
        X = B.reshape(k, -1, order="F")
        err = np.max(np.abs(X @ X.T - np.eye(k)))
    prod = np.tensordot(A, B, axes=(2, 0))
    perr = np.max(np.abs(prod - AA)) / np.max(np.abs(AA))
    say("  orthonormality err %.2e product err %.2e" % (err, perr))
    assert err < 1e-12
    assert perr < 1e-12


def test_mfma_bf16_lane_layout(pkg, ctx):
    import ctypes as Ct
    out = np.zeros(768, dtype=np.float32)
    ctx.call("dmrj_selftest_mfma_bf16", out.ctypes.data_as(Ct.POINTER(Ct.c_float)))
    rows = out[:256].reshape(64, 4)
    cols = out[256:512].reshape(64, 4)
    enc = out[512:].reshape(64, 4)
    # operands: A[i][k] from lane i + 16(k/8) element k%8, B[k][j] from lane
    # j + 16(k/8); result: item r of lane l holds D[4(l/16) + r][l%16]
    wr = np.zeros((64, 4))
    wc = np.zeros((64, 4))
    we = np.zeros((64, 4))
    for l in range(64):
        for r in range(4):
            i, j = 4 * (l // 16) + r, l % 16
            wr[l, r] = 8 * (4 * i + 100)
            wc[l, r] = 8 * (4 * j + 100)
            we[l, r] = (i + 1) + 16 * (j + 1)
    if not (np.array_equal(rows, wr) and np.array_equal(cols, wc) and np.array_equal(enc, we)):
        say("bf16 probe rows:", rows.astype(int).tolist())
        say("bf16 probe cols:", cols.astype(int).tolist())
        say("bf16 probe enc:", enc.astype(int).tolist())
    assert np.array_equal(rows, wr)
    assert np.array_equal(cols, wc)
    assert np.array_equal(enc, we)


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (33, 70, 17), (130, 129, 200)])
def test_grouped_gemm_bf16x3(pkg, ctx, cfg, tA, tB, M, N, K):
    """Split-precision GEMM: error bounded by ~2^-16 of the |A||B| scale."""
    rng = np.random.default_rng(M + 5 * N + 11 * K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    As = np.asfortranarray(A.T if tA else A)
    Bs = np.asfortranarray(B.T if tB else B)
    C = np.zeros((M, N), order="F")
    ctx.call("dmrj_selftest_gemm", cfg | (1 << 4), M, N, K, tA, tB, pkg._lib.dptr(As.ravel(order="F")),
             pkg._lib.dptr(Bs.ravel(order="F")), C.ctypes.data_as(pkg._lib._dp))
    want = A @ B
    scale = np.abs(A) @ np.abs(B)
    err = np.max(np.abs(C - want) / scale)
    say(" bf16x3 cfg", cfg, tA, tB, M, N, K, "err/|A||B| %.2e" % err)
    assert err < 6e-5


def test_block_jacobi_panel_boundary_640_rows(pkg, ctx):
    """Backend 4 on a sector of exactly 640 rows: the 16-wide panel would fill
    the whole 160 KiB of LDS, so the launcher narrows it (width 8) with the
    kernel's static LDS read from its attributes; the spectrum of a graded
    640 x 600 matrix matches LAPACK."""
    say("case block Jacobi 640 rows")
    R, C = 640, 600
    rng = np.random.default_rng(640)
    k = min(R, C)
    U, _ = np.linalg.qr(rng.standard_normal((R, k)))
    V, _ = np.linalg.qr(rng.standard_normal((C, k)))
    s = np.logspace(0, -6, k)
    M = (U * s) @ V.T
    M /= np.linalg.norm(M)
    AA = M.reshape(R // 2, 2, 2, C // 2, order="F")
    t = pkg.site2_from_dense(ctx, AA)
    ctx.set_svd_backend(4)
    try:
        l, r, te, sv = _split(pkg, ctx, t, 0, 1)
    finally:
        ctx.set_svd_backend(0)
    want = np.linalg.svd(AA.reshape(R, C, order="F"), compute_uv=False)
    want = want / np.linalg.norm(want)
    n = min(len(sv), len(want))
    err = np.max(np.abs(sv[:n] - want[:n]))
    say(" kept", len(sv), "abs err %.2e" % err, "jacobi stats", ctx.svd_stats())
    assert err < 1e-13
    for x in (l, r, t):
        x.free()
