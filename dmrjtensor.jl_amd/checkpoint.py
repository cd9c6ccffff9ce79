"""Checkpoint and resume of a GPU-resident MPS.

The reference keeps long runs restartable by writing tensors to disk
(largeMPS / tensor2disc, lib/MPS/largeMPS.jl:19-81) and by handing the
environments back through params.Lenv / params.Renv
(algs/optimizeMPS.jl:273-334).  This package stores an MPS in its own format:
one NumPy ``.npz`` archive (no pickled objects) holding, per site, the block
storage of the device tensor exactly as the library keeps it (real plane then
imaginary plane) and the per-index charges of both bonds and of the physical
index, plus a JSON header (length, centre, charges, physical dimensions and
any scalars of the run the caller adds).  Loading rebuilds every tensor bit
for bit (dmrj_site_import); the environments are rebuilt by makeEnv, so a
resumed dmrg continues exactly where the saved one stopped.
"""
import ctypes as C
import json

import numpy as np

from . import _lib
from .network import MPS, DeviceTensor

FORMAT = "dmrj-mps-1"


def _export(t):
    ctx = t.ctx
    info = np.zeros(8, dtype=np.int64)
    ctx.call("dmrj_site_export", t.h, info.ctypes.data_as(_lib._i64p), None, None, None, None)
    nq, mL, P, mR, n = (int(x) for x in info[:5])
    qL = np.zeros((mL, max(nq, 1)), dtype=np.int32)
    qP = np.zeros((P, max(nq, 1)), dtype=np.int32)
    qR = np.zeros((mR, max(nq, 1)), dtype=np.int32)
    data = np.zeros(max(n, 1))
    ctx.call("dmrj_site_export", t.h, info.ctypes.data_as(_lib._i64p), _lib.iptr(qL), _lib.iptr(qP),
             _lib.iptr(qR), _lib.dptr(data))
    return info, qL, qP, qR, data[:n]


def save_mps(psi, path, extra=None):
    """Write psi (and the JSON-serialisable dict ``extra``) to ``path`` (.npz)."""
    arrays = {}
    for i, t in enumerate(psi.A, start=1):
        info, qL, qP, qR, data = _export(t)
        arrays["A%d_info" % i] = info
        arrays["A%d_qL" % i] = qL
        arrays["A%d_qP" % i] = qP
        arrays["A%d_qR" % i] = qR
        arrays["A%d_data" % i] = data
    head = {"format": FORMAT, "L": len(psi), "oc": int(psi.oc), "nq": int(psi.nq), "d": [int(x) for x in psi.d],
            "qphys": None if psi.qphys is None else psi.qphys.tolist(), "extra": extra or {}}
    arrays["header"] = np.frombuffer(json.dumps(head).encode(), dtype=np.uint8)
    with open(path, "wb") as f:
        np.savez(f, **arrays)


def _site_arrays(z, path, i, nq_head):
    """One site's arrays from the archive, checked against its info record
    before anything reaches the library (which reads mL*nq, P*nq and mR*nq
    charges and n values through raw pointers)."""
    info = np.asarray(z["A%d_info" % i])
    if info.ndim != 1 or info.size < 6:
        raise ValueError("%s: site %d: info record has %d entries, expected 8" % (path, i, info.size))
    nq, mL, P, mR, n, cplx = (int(x) for x in info[:6])
    if nq != nq_head or not 0 <= nq <= 2:
        raise ValueError("%s: site %d carries nq=%d, the header says %d" % (path, i, nq, nq_head))
    if min(mL, P, mR) < 1 or n < 0 or cplx not in (0, 1):
        raise ValueError("%s: site %d: bad dimensions %s" % (path, i, info[:6].tolist()))
    w = max(nq, 1)
    out = []
    for name, rows in (("qL", mL), ("qP", P), ("qR", mR)):
        q = np.ascontiguousarray(z["A%d_%s" % (i, name)], dtype=np.int32)
        if q.shape != (rows, w):
            raise ValueError("%s: site %d: %s has shape %s, expected %s" % (path, i, name, q.shape, (rows, w)))
        out.append(q)
    data = np.ascontiguousarray(z["A%d_data" % i], dtype=np.float64)
    if data.ndim != 1 or data.size != n:
        raise ValueError("%s: site %d holds %d values, expected %d" % (path, i, data.size, n))
    return (nq, mL, P, mR, n, cplx), out[0], out[1], out[2], data


def load_mps(path, ctx=None):
    """Read an MPS written by save_mps onto the device; returns (psi, extra).
    Every array is checked against the header and the site records first; a
    truncated or inconsistent archive raises ValueError before any device
    call."""
    with np.load(path, allow_pickle=False) as z:
        head = json.loads(bytes(z["header"]).decode())
        if head.get("format") != FORMAT:
            raise ValueError("%s: not a %s archive" % (path, FORMAT))
        L = int(head["L"])
        if L < 1 or len(head.get("d", [])) != L:
            raise ValueError("%s: header length and physical dimensions disagree" % path)
        sites = [_site_arrays(z, path, i, int(head["nq"])) for i in range(1, L + 1)]
    ctx = ctx or _lib.default_context()
    tensors = []
    for (nq, mL, P, mR, n, cplx), qL, qP, qR, data in sites:
        buf = data if n else np.zeros(1)
        h = C.c_void_p()
        ctx.call("dmrj_site_import", nq, mL, P, mR, _lib.iptr(qL), _lib.iptr(qP), _lib.iptr(qR),
                 _lib.dptr(buf), n, cplx, C.byref(h))
        tensors.append(DeviceTensor(ctx, h))
    psi = MPS.__new__(MPS)
    psi.ctx = ctx
    psi.qphys = None if head["qphys"] is None else np.asarray(head["qphys"], dtype=np.int64)
    psi.nq = head["nq"]
    psi.oc = head["oc"]
    psi.d = list(head["d"])
    psi.A = tensors
    return psi, head["extra"]
