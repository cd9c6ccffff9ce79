"""ctypes binding of libdmrjhip (include/dmrj.h).

The shared library is built in-tree (``lib/libdmrjhip.so``) by
``build.py``; loading fails loudly when it is missing.  There is no CPU
fallback: every numerical operation of the package runs through this library
on the GPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdmrjhip.so")

PREC = {"f64": 0, "bf16x3": 1, "f16x3": 2, "f64c": 16, "bf16x3c": 17, "f16x3c": 18}
CORRECT = 16

_p = C.c_void_p
_i64 = C.c_int64
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)

# name -> argtypes (all functions return int, except dmrj_last_error)
SIGNATURES = {
    "dmrj_version": [],
    "dmrj_ctx_create": [C.c_int, C.POINTER(_p)],
    "dmrj_ctx_destroy": [_p],
    "dmrj_ctx_sync": [_p],
    "dmrj_ctx_memstats": [_p, _i64p],
    "dmrj_ctx_set_gemm_cfg": [_p, C.c_int],
    "dmrj_ctx_set_eig_batching": [_p, C.c_int],
    "dmrj_ctx_set_autotune": [_p, C.c_int],
    "dmrj_ctx_set_tile_order": [_p, C.c_int],
    "dmrj_ctx_set_f16_presplit": [_p, C.c_int],
    "dmrj_ctx_svd_stats": [_p, _dp, C.c_int],
    "dmrj_ctx_autotune_stats": [_p, _dp],
    "dmrj_ctx_set_graph": [_p, C.c_int],
    "dmrj_ctx_marker": [_p, C.c_int],
    "dmrj_ctx_graph_stats": [_p, _dp],
    "dmrj_ctx_flops": [_p, _dp, C.c_int],
    "dmrj_ctx_profile": [_p, C.c_int],
    "dmrj_ctx_set_svd_backend": [_p, C.c_int],
    "dmrj_ctx_set_env_precision": [_p, C.c_int],
    "dmrj_ctx_last_residual": [_p, _dp],
    "dmrj_selftest_mfma_bf16": [_p, C.POINTER(C.c_float)],
    "dmrj_ctx_kernel_stats": [_p, C.c_int, _dp, C.c_int],
    "dmrj_site_from_dense": [_p, C.c_int, _i64, _i64, _i64, _dp, _i32p, _i32p, _i32p, C.POINTER(_p)],
    "dmrj_site2_from_dense": [_p, C.c_int, _i64, _i64, _i64, _i64, _dp, _i32p, _i32p, _i32p, _i32p,
                              C.POINTER(_p)],
    "dmrj_tensor_info": [_p, _i64p],
    "dmrj_tensor_to_dense": [_p, _p, _dp, _i32p, _i32p],
    "dmrj_tensor_to_dense_z": [_p, _p, _dp, _i32p, _i32p],
    "dmrj_site_from_dense_z": [_p, C.c_int, _i64, _i64, _i64, _dp, _i32p, _i32p, _i32p, C.POINTER(_p)],
    "dmrj_site2_from_dense_z": [_p, C.c_int, _i64, _i64, _i64, _i64, _dp, _i32p, _i32p, _i32p, _i32p,
                                C.POINTER(_p)],
    "dmrj_tensor_is_complex": [_p],
    "dmrj_tensor_to_complex": [_p, _p],
    "dmrj_dotc": [_p, _p, _p, _dp],
    "dmrj_dotu": [_p, _p, _p, _dp],
    "dmrj_tensor_free": [_p, _p],
    "dmrj_tensor_copy": [_p, _p, C.POINTER(_p)],
    "dmrj_site_export": [_p, _p, _i64p, _i32p, _i32p, _i32p, _dp],
    "dmrj_site_import": [_p, C.c_int, _i64, _i64, _i64, _i32p, _i32p, _i32p, _dp, _i64, C.c_int, C.POINTER(_p)],
    "dmrj_mpo_create": [_p, C.c_int, _i64, _i64, _i64, _dp, _i32p, _i32p, _i32p, C.POINTER(_p)],
    "dmrj_mpo_free": [_p, _p],
    "dmrj_env_boundary": [_p, C.c_int, C.c_int, _i32p, C.POINTER(_p)],
    "dmrj_env_left": [_p, _p, _p, _p, C.POINTER(_p)],
    "dmrj_env_right": [_p, _p, _p, _p, C.POINTER(_p)],
    "dmrj_theta2": [_p, _p, _p, C.POINTER(_p)],
    "dmrj_heff2_apply": [_p, _p, _p, _p, _p, _p, C.c_int, C.POINTER(_p)],
    "dmrj_heff1_apply": [_p, _p, _p, _p, _p, C.c_int, C.POINTER(_p)],
    "dmrj_lanczos2": [_p, _p, _p, _p, _p, _p, _p, C.c_int, C.POINTER(_p), _dp],
    "dmrj_krylov_expm2": [_p, _p, _p, _p, _p, _p, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(_p), _dp,
                          C.POINTER(C.c_int)],
    "dmrj_krylov_expm1": [_p, _p, _p, _p, _p, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(_p), _dp,
                          C.POINTER(C.c_int)],
    "dmrj_lanczos1": [_p, _p, _p, _p, _p, C.c_int, C.c_int, C.POINTER(_p), _dp, C.POINTER(C.c_int)],
    "dmrj_lanczos_k": [_p, _p, _p, _p, _p, _p, C.c_int, C.c_int, C.POINTER(_p), _dp, C.POINTER(C.c_int)],
    "dmrj_expand3s": [_p, _p, _p, _p, C.c_double, C.c_int, C.POINTER(_p)],
    "dmrj_absorb3s": [_p, _p, _p, C.c_int, C.POINTER(_p)],
    "dmrj_svd_trunc2": [_p, _p, _i64, _i64, C.c_double, C.c_int, C.POINTER(_p), C.POINTER(_p), _dp,
                        _i64p, _dp, _i64],
    "dmrj_twostep": [_p, _p, _p, _p, _p, _p, _p, C.c_int, _i64, _i64, C.c_double, C.c_int,
                     C.POINTER(_p), C.POINTER(_p), _dp, _dp, _i64p, _dp, _i64],
    "dmrj_dot": [_p, _p, _p, _dp],
    "dmrj_bj_panel_width": [C.c_int, _i64, _i64],
    "dmrj_env_left2": [_p, _p, _p, _p, _p, C.POINTER(_p)],
    "dmrj_env_right2": [_p, _p, _p, _p, _p, C.POINTER(_p)],
    "dmrj_site_apply_op": [_p, _p, _dp, _i32p, _i32p, C.POINTER(_p)],
    "dmrj_dots_noalloc": [_p, C.c_int, C.POINTER(_p), C.POINTER(_p), _p],
    "dmrj_norm": [_p, _p, _dp],
    "dmrj_scale": [_p, _p, C.c_double],
    "dmrj_axpy": [_p, C.c_double, _p, _p],
    "dmrj_normalize": [_p, _p],
    "dmrj_selftest_gemm": [_p, C.c_int, _i64, _i64, _i64, C.c_int, C.c_int, _dp, _dp, _dp],
    "dmrj_selftest_mfma_f64": [_p, _dp],
    "dmrj_selftest_copy": [_p, _i64],
}

_lib = None


class DMRJError(RuntimeError):
    pass


def load():
    """Load libdmrjhip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DMRJError("libdmrjhip.so is not built (%s); run build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.dmrj_last_error.argtypes = [_p]
    lib.dmrj_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_i32p)


def charges_arg(q, n, nq):
    """Charges as a contiguous int32 [n][nq] array (or None for nq == 0)."""
    if nq == 0 or q is None:
        return None
    a = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(n, nq))
    return a


class Context:
    """One device context = one HIP stream (dmrj_ctx)."""

    def __init__(self, device=0):
        self.lib = load()
        h = _p()
        rc = self.lib.dmrj_ctx_create(device, C.byref(h))
        if rc != 0:
            raise DMRJError("dmrj_ctx_create failed (%d): no usable GPU" % rc)
        self.h = h

    def check(self, rc, what=""):
        if rc != 0:
            msg = self.lib.dmrj_last_error(self.h)
            raise DMRJError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def call(self, name, *args):
        self.check(getattr(self.lib, name)(self.h, *args), name)

    def sync(self):
        self.call("dmrj_ctx_sync")

    def memstats(self):
        out = (C.c_int64 * 3)()
        self.call("dmrj_ctx_memstats", out)
        return dict(live=out[0], cached=out[1], peak=out[2])

    def flops(self, reset=False):
        out = (C.c_double * 2)()
        self.call("dmrj_ctx_flops", out, 1 if reset else 0)
        return out[0], out[1]

    def profile(self, enable=True):
        self.call("dmrj_ctx_profile", 1 if enable else 0)

    KINDS = {"gemm_f64": 0, "mix": 1, "copy": 2, "svd": 3, "gemm_lp": 4, "graph": 5, "gemm_f16": 6, "f16_prep": 7}

    def kernel_stats(self, kind, reset=False):
        out = (C.c_double * 4)()
        self.call("dmrj_ctx_kernel_stats", self.KINDS.get(kind, kind), out, 1 if reset else 0)
        return dict(launches=int(out[0]), ms=out[1], flops=out[2], bytes=out[3])

    def set_svd_backend(self, mode):
        self.call("dmrj_ctx_set_svd_backend", int(mode))

    def set_gemm_cfg(self, cfg):
        self.call("dmrj_ctx_set_gemm_cfg", int(cfg))

    def set_eig_batching(self, mode):
        self.call("dmrj_ctx_set_eig_batching", int(mode))

    def svd_stats(self, reset=False):
        """Block-Jacobi SVD: calls and total sweeps since the last reset."""
        out = np.zeros(2)
        self.call("dmrj_ctx_svd_stats", dptr(out), 1 if reset else 0)
        return dict(calls=int(out[0]), sweeps=int(out[1]))

    def set_f16_presplit(self, on):
        self.call("dmrj_ctx_set_f16_presplit", 1 if on else 0)

    def set_tile_order(self, mode):
        """1 = XCD-aware grouped-GEMM tile order (default), 0 = descending work only."""
        self.call("dmrj_ctx_set_tile_order", int(mode))

    def set_autotune(self, mode):
        """1 = autotune GEMM tile shapes per block-size histogram, 0 = heuristic, 2 = clear cache."""
        self.call("dmrj_ctx_set_autotune", int(mode))

    def autotune_stats(self):
        out = np.zeros(5)
        self.call("dmrj_ctx_autotune_stats", dptr(out))
        return dict(keys=int(out[0]), hits=int(out[1]), tuned=int(out[2]), untunable=int(out[3]),
                    tune_ms=float(out[4]))

    def marker(self, ident):
        """Enqueue the empty k_region_marker kernel (cuts kernel traces)."""
        self.call("dmrj_ctx_marker", int(ident))

    def set_graph(self, mode):
        """1 = capture each Lanczos step into a hipGraph, 0 = direct launches."""
        self.call("dmrj_ctx_set_graph", int(mode))

    def graph_stats(self):
        out = np.zeros(2)
        self.call("dmrj_ctx_graph_stats", dptr(out))
        return dict(instantiations=int(out[0]), updates=int(out[1]))

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmrj_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default
