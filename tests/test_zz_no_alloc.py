"""dmrj_dots_noalloc (the allocation-free inner products) against dmrj_dot,
through the C ABI, on two calls in a row, the second with fewer and smaller
inputs.  The output buffer is device memory the caller owns, allocated here
with hipMalloc from the HIP runtime the library itself runs on."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DevBuf:
    """A caller-owned device buffer of n doubles (plain HIP runtime calls)."""

    def __init__(self, n):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0
        nan = np.full(n, np.nan)
        assert self.hip.hipMemcpy(self.p, nan.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0  # H2D

    def read(self, k):
        out = np.zeros(k)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * k), 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _dot(ctx, x, y):
    out = C.c_double()
    ctx.call("dmrj_dot", x.h, y.h, C.byref(out))
    return out.value


def test_dots_noalloc_match_dmrj_dot(pkg, ctx):
    rng = np.random.default_rng(7)
    buf = _DevBuf(16)
    try:
        # second call: fewer pairs of smaller tensors, into the same scratch
        for (n_pairs, mL, mR) in ((12, 300, 257), (5, 40, 33)):
            print("case dots_noalloc pairs=%d mL=%d mR=%d" % (n_pairs, mL, mR), flush=True)
            xs, ys = [], []
            for p in range(n_pairs):
                xs.append(pkg.site_from_dense(ctx, rng.standard_normal((mL + p, 2, mR))))
                ys.append(pkg.site_from_dense(ctx, rng.standard_normal((mL + p, 2, mR))))
            xa = (C.c_void_p * n_pairs)(*[x.h for x in xs])
            ya = (C.c_void_p * n_pairs)(*[y.h for y in ys])
            ctx.call("dmrj_dots_noalloc", n_pairs, xa, ya, buf.p)
            ctx.sync()
            got = buf.read(n_pairs)
            want = np.array([_dot(ctx, x, y) for x, y in zip(xs, ys)])
            # same kernels, same grid and summation order: bit-identical
            assert np.array_equal(got, want), (got, want)
            dense = np.array([np.sum(x.to_dense()[0] * y.to_dense()[0]) for x, y in zip(xs, ys)])
            assert np.allclose(got, dense, rtol=1e-12, atol=0)
            for t in xs + ys:
                t.free()
    finally:
        buf.free()


def test_dots_noalloc_rejects_too_many_pairs(pkg, ctx):
    buf = _DevBuf(32)
    a = pkg.site_from_dense(ctx, np.ones((2, 2, 2)))
    try:
        xa = (C.c_void_p * 17)(*([a.h] * 17))
        lib = pkg._lib.load()
        rc = lib.dmrj_dots_noalloc(ctx.h, 17, xa, xa, buf.p)
        assert rc == -1
        assert b"npairs" in lib.dmrj_last_error(ctx.h)
        assert np.all(np.isnan(buf.read(32)))   # nothing was written
    finally:
        a.free()
        buf.free()
