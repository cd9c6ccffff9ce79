"""Device-resident MPS / MPO / environment containers.

Mirrors the reference's container types and constructors:
  matrixproductstate / MPS         lib/types.jl:84-87, lib/MPS/MPS.jl:17-37
  matrixproductoperator / makeMPO  lib/types.jl:99-101, lib/MPO/makeMPO.jl:59-92
  environment / makeEnv            lib/types.jl:113-115, lib/Env/makeEnv.jl:196-222
  MPS(Qlabels, psi, mpo)           lib/MPS/MPS.jl:259-263,301-305 (makeqMPS /
                                   makeqMPO, lib/MPS/makeqMPS.jl:26-112,
                                   lib/MPO/makeqMPO.jl:23-93)
  Lupdate / Rupdate                lib/MPO_MPS/Lupdate.jl:17-26, Rupdate.jl:17-26
  expect                           lib/MPO_MPS/expect.jl:19-45

Tensors live in HBM behind opaque handles of libdmrjhip; the host keeps only
shapes and charge labels.  Host-side numpy is used for one-off input
preparation (initial states, MPO matrices, canonical form of a user-supplied
state), never inside a sweep.
"""
import ctypes as C

import numpy as np

from . import _lib


class DeviceTensor:
    """Opaque device tensor (MPS tensor, two-site tensor or environment)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def info(self):
        d = (C.c_int64 * 9)()
        self.ctx.check(self.ctx.lib.dmrj_tensor_info(self.h, d), "dmrj_tensor_info")
        return list(d)

    def copy(self):
        """Device copy (dmrj_tensor_copy)."""
        h = C.c_void_p()
        self.ctx.call("dmrj_tensor_copy", self.h, C.byref(h))
        return DeviceTensor(self.ctx, h)

    @property
    def kind(self):
        return self.info()[0]

    @property
    def shape(self):
        i = self.info()
        if i[0] == 1:
            return (i[1], i[2], i[3], i[4])
        return (i[1], i[2], i[3])

    @property
    def nelem(self):
        return self.info()[5]

    @property
    def is_complex(self):
        r = self.ctx.lib.dmrj_tensor_is_complex(self.h)
        if r < 0:
            raise _lib.DMRJError("dmrj_tensor_is_complex: null handle")
        return bool(r)

    def to_complex(self):
        """Promote in place to ComplexF64 (zero imaginary part)."""
        self.ctx.call("dmrj_tensor_to_complex", self.h)
        return self

    def to_dense(self, nq=0):
        """Dense column-major copy with the reference's index order, plus the
        charges of the first and last index (sector-sorted basis).  Complex
        tensors come back as complex128 arrays."""
        shp = self.shape
        z = self.is_complex
        out = np.zeros(int(np.prod(shp)) * (2 if z else 1), dtype=np.float64)
        q0 = np.zeros((shp[0], max(nq, 1)), dtype=np.int32)
        q1 = np.zeros((shp[-1], max(nq, 1)), dtype=np.int32)
        self.ctx.call("dmrj_tensor_to_dense_z" if z else "dmrj_tensor_to_dense", self.h, _lib.dptr(out),
                      _lib.iptr(q0) if nq else None, _lib.iptr(q1) if nq else None)
        if z:
            out = out.view(np.complex128)
        arr = out.reshape(shp, order="F")
        return arr, (q0[:, :nq] if nq else None), (q1[:, :nq] if nq else None)

    def norm(self):
        r = C.c_double()
        self.ctx.call("dmrj_norm", self.h, C.byref(r))
        return r.value

    def dot(self, other):
        """Re <self|other>."""
        r = C.c_double()
        self.ctx.call("dmrj_dot", self.h, other.h, C.byref(r))
        return r.value

    def dotc(self, other):
        """<self|other> = sum conj(self) other (complex)."""
        out = np.zeros(2)
        self.ctx.call("dmrj_dotc", self.h, other.h, _lib.dptr(out))
        return complex(out[0], out[1])

    def dotu(self, other):
        """sum self other without conjugation (complex)."""
        out = np.zeros(2)
        self.ctx.call("dmrj_dotu", self.h, other.h, _lib.dptr(out))
        return complex(out[0], out[1])

    def free(self):
        if self.h:
            self.ctx.lib.dmrj_tensor_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


def _host_buffer(A):
    """Column-major float64 buffer; complex arrays as interleaved (re, im)."""
    if np.iscomplexobj(A):
        z = np.asfortranarray(np.asarray(A, dtype=np.complex128)).ravel(order="F")
        return np.ascontiguousarray(z).view(np.float64), True
    return np.asfortranarray(np.asarray(A, dtype=np.float64)).ravel(order="F"), False


def site_from_dense(ctx, A, qL=None, qP=None, qR=None, nq=0):
    mL, d, mR = np.shape(A)
    buf, z = _host_buffer(A)
    h = C.c_void_p()
    ql = _lib.charges_arg(qL, mL, nq)
    qp = _lib.charges_arg(qP, d, nq)
    qr = _lib.charges_arg(qR, mR, nq)
    ctx.call("dmrj_site_from_dense_z" if z else "dmrj_site_from_dense", nq, mL, d, mR, _lib.dptr(buf),
             _lib.iptr(ql), _lib.iptr(qp), _lib.iptr(qr), C.byref(h))
    return DeviceTensor(ctx, h)


def site2_from_dense(ctx, A, qL=None, qP1=None, qP2=None, qR=None, nq=0):
    mL, d1, d2, mR = np.shape(A)
    buf, z = _host_buffer(A)
    h = C.c_void_p()
    ctx.call("dmrj_site2_from_dense_z" if z else "dmrj_site2_from_dense", nq, mL, d1, d2, mR, _lib.dptr(buf),
             _lib.iptr(_lib.charges_arg(qL, mL, nq)), _lib.iptr(_lib.charges_arg(qP1, d1, nq)),
             _lib.iptr(_lib.charges_arg(qP2, d2, nq)), _lib.iptr(_lib.charges_arg(qR, mR, nq)),
             C.byref(h))
    return DeviceTensor(ctx, h)


# ----------------------------------------------------------------------- MPO
def makeMPO(H, d, Ns=None):
    """lib/MPO/makeMPO.jl:59-92.  ``H`` is a bulk matrix, a list of them, or a
    function i -> bulk matrix (1-based i, as the reference).  Returns an MPO
    of W[a1, j, k, a2] arrays with W[l,j,k,m] = H[j + l*d, k + m*d]."""
    if callable(H):
        mats = [H(i) for i in range(1, Ns + 1)]
    elif isinstance(H, (list, tuple)):
        mats = list(H)
    else:
        mats = [H] * Ns
    Ws = []
    for Hi in mats:
        Hi = np.asarray(Hi, dtype=np.float64)
        a1, a2 = Hi.shape[0] // d, Hi.shape[1] // d
        Ws.append(Hi.reshape(a1, d, a2, d).transpose(0, 1, 3, 2).copy())
    Ws[0] = Ws[0][-1:].copy()
    Ws[-1] = Ws[-1][..., :1].copy()
    return MPO(Ws)


class MPO:
    """matrixproductoperator: host W tensors, optional charges, and lazily
    created device handles per context."""

    def __init__(self, Ws, qphys=None):
        self.H = [np.asarray(W, dtype=np.float64) for W in Ws]
        self.qphys = None if qphys is None else np.asarray(qphys, dtype=np.int64)
        self.deltas = None
        if self.qphys is not None:
            self.deltas = mpo_deltas(self.H, self.qphys)
        self._dev = {}

    def __len__(self):
        return len(self.H)

    def __getitem__(self, i):  # 1-based like the reference
        return self.H[i - 1]

    @property
    def nq(self):
        return 0 if self.qphys is None else self.qphys.shape[1]

    def device(self, ctx, i):
        """Device handle of site i (1-based)."""
        key = (id(ctx), i)
        if key not in self._dev:
            W = self.H[i - 1]
            DL, d, _, DR = W.shape
            nq = self.nq
            buf = np.asfortranarray(W).ravel(order="F")
            h = C.c_void_p()
            qp = _lib.charges_arg(self.qphys, d, nq)
            dl = _lib.charges_arg(self.deltas[i - 1], DL, nq) if nq else None
            dr = _lib.charges_arg(self.deltas[i], DR, nq) if nq else None
            ctx.call("dmrj_mpo_create", nq, DL, d, DR, _lib.dptr(buf), _lib.iptr(qp), _lib.iptr(dl),
                     _lib.iptr(dr), C.byref(h))
            self._dev[key] = (ctx, h)
        return self._dev[key][1]

    def __del__(self):
        for (ctx, h) in getattr(self, "_dev", {}).values():
            try:
                if ctx.h:
                    ctx.lib.dmrj_mpo_free(ctx.h, h)
            except Exception:
                pass


def mpo_deltas(Ws, qphys):
    """Charges of every MPO link (lib/MPO/makeqMPO.jl:53-74 restated):
    W[w,j,k,u] != 0 requires delta[u] = delta[w] + q[k] - q[j]; the left edge
    link carries charge 0."""
    qphys = np.asarray(qphys, dtype=np.int64)
    nq = qphys.shape[1]
    out = [np.zeros((Ws[0].shape[0], nq), dtype=np.int64)]
    for W in Ws:
        dl = out[-1]
        dr = np.zeros((W.shape[3], nq), dtype=np.int64)
        seen = np.zeros(W.shape[3], dtype=bool)
        for (w, j, k, u) in np.argwhere(W != 0.0):
            v = dl[w] + qphys[k] - qphys[j]
            if not seen[u]:
                dr[u], seen[u] = v, True
            elif not np.array_equal(dr[u], v):
                raise ValueError("MPO does not conserve the given charges")
        out.append(dr)
    return out


# ----------------------------------------------------------------------- MPS
def _qr(A):
    a, d, b = A.shape
    Q, R = np.linalg.qr(A.reshape(a * d, b))
    return Q.reshape(a, d, -1), R


def _lq(A):
    a, d, b = A.shape
    Q, R = np.linalg.qr(A.reshape(a, d * b).T)
    return R.T, Q.T.reshape(-1, d, b)


def canonicalize(tensors, oc):
    """Host-side QR/LQ gauge fixing of a user-supplied dense MPS (move! with
    m=0, cutoff=0, lib/MPS/move.jl:124-197), used once before upload."""
    A = [np.array(t, dtype=np.float64) for t in tensors]
    L = len(A)
    for i in range(oc - 1):
        Q, R = _qr(A[i])
        A[i] = Q
        A[i + 1] = np.tensordot(R, A[i + 1], axes=(1, 0))
    for i in range(L - 1, oc - 1, -1):
        Lm, Q = _lq(A[i])
        A[i] = Q
        A[i - 1] = np.tensordot(A[i - 1], Lm, axes=(2, 0))
    return A


class MPS:
    """matrixproductstate: device tensors A[1..L], orthogonality centre oc,
    optional charges (qphys [d, nq]; bond charges are carried by the device
    tensors)."""

    def __init__(self, tensors, oc=1, qphys=None, qbond=None, ctx=None, canonical=False):
        self.ctx = ctx or _lib.default_context()
        self.qphys = None if qphys is None else np.asarray(qphys, dtype=np.int64)
        nq = 0 if self.qphys is None else self.qphys.shape[1]
        self.nq = nq
        self.oc = oc
        hosts = [np.asarray(t, dtype=np.float64) for t in tensors]
        if not canonical and qbond is None:
            hosts = canonicalize(hosts, oc)
        self.d = [h.shape[1] for h in hosts]
        if nq:
            if qbond is None:
                raise ValueError("quantum-number MPS needs bond charges")
            hosts, qbond = _sort_bonds(hosts, [np.asarray(q, dtype=np.int64) for q in qbond])
        self.A = []
        for i, h in enumerate(hosts):
            if nq:
                self.A.append(site_from_dense(self.ctx, h, qbond[i], self.qphys, qbond[i + 1], nq))
            else:
                self.A.append(site_from_dense(self.ctx, h))

    def __len__(self):
        return len(self.A)

    def __getitem__(self, i):
        return self.A[i - 1]

    def __setitem__(self, i, v):
        self.A[i - 1] = v

    def to_dense(self):
        """Host copies of all tensors (and the bond charges when nq > 0)."""
        out, qb = [], []
        for i, t in enumerate(self.A):
            a, q0, q1 = t.to_dense(self.nq)
            out.append(a)
            if self.nq:
                if i == 0:
                    qb.append(q0)
                qb.append(q1)
        return out, (qb if self.nq else None)

    def bond_dims(self):
        return [t.shape[2] for t in self.A]

    def copy(self):
        """Independent device copy of the state (same centre and charges)."""
        c = MPS.__new__(MPS)
        c.ctx, c.qphys, c.nq, c.oc, c.d = self.ctx, self.qphys, self.nq, self.oc, list(self.d)
        c.A = [t.copy() for t in self.A]
        return c

    def free(self):
        for t in self.A:
            t.free()
        self.A = []


def _sort_bonds(hosts, qbond):
    """Permute every bond basis so that its charges are sorted (stable)."""
    hosts = [h.copy() for h in hosts]
    qbond = [q.copy() for q in qbond]
    for b in range(len(qbond)):
        q = qbond[b]
        perm = np.lexsort(q.T[::-1]) if q.shape[1] > 0 else np.arange(q.shape[0])
        if np.all(perm == np.arange(len(perm))):
            continue
        qbond[b] = q[perm]
        if b > 0:
            hosts[b - 1] = hosts[b - 1][:, :, perm]
        if b < len(hosts):
            hosts[b] = hosts[b][perm]
    return hosts, qbond


def product_state(states, d, qphys=None, ctx=None):
    """Product state |s_1 ... s_L> (test/dmrg_test.jl:55-60); with charges the
    bond charges are the running sums of the physical charges."""
    A = []
    for s in states:
        t = np.zeros((1, d, 1))
        t[0, s, 0] = 1.0
        A.append(t)
    qbond = None
    if qphys is not None:
        qphys = np.asarray(qphys, dtype=np.int64)
        q = np.zeros((1, qphys.shape[1]), dtype=np.int64)
        qbond = [q.copy()]
        for s in states:
            q = q + qphys[s][None, :]
            qbond.append(q.copy())
    return MPS(A, oc=1, qphys=qphys, qbond=qbond, ctx=ctx, canonical=True)


def _op_charge(op, qphys):
    """Charge moved by a one-site operator: q(k) - q(j) for its non-zero
    entries O[k, j] (one value, or the operator mixes charges)."""
    nz = np.argwhere(np.asarray(op) != 0.0)
    if len(nz) == 0:
        return np.zeros(qphys.shape[1], dtype=np.int64)
    dq = {tuple(int(v) for v in qphys[k] - qphys[j]) for k, j in nz}
    if len(dq) != 1:
        raise ValueError("applyOps!: the operator does not carry one definite charge")
    return np.array(next(iter(dq)), dtype=np.int64)


def _apply_site(psi, i, op, dqL, dqR):
    op = np.asarray(op, dtype=np.float64)
    d = psi.d[i - 1]
    if op.shape != (d, d):
        raise ValueError("applyOps!: operator must be %d x %d" % (d, d))
    nq = psi.nq
    h = C.c_void_p()
    buf = np.ascontiguousarray(op.ravel(order="F"))
    l = np.ascontiguousarray(dqL, dtype=np.int32) if nq else None
    r = np.ascontiguousarray(dqR, dtype=np.int32) if nq else None
    psi.ctx.call("dmrj_site_apply_op", psi[i].h, _lib.dptr(buf), _lib.iptr(l), _lib.iptr(r), C.byref(h))
    psi[i].free()
    psi[i] = DeviceTensor(psi.ctx, h)


def applyOps(psi, sites, Op, trail=None):
    """applyOps!(psi, sites, Op; trail) (lib/MPO_MPS/applyMPO.jl:21-33) in place
    on the GPU: for every listed site p in order, psi[p] = Op psi[p] (Op[bra,
    ket] acting on the physical index), then, when a trail is given, the trail
    on every site left of p.  On a charged MPS a charged operator moves the
    bond charges right of p by its charge (the tensors right of p are
    relabelled, their data unchanged), so the state keeps the block rule
    q(a) + q(s) = q(b).  Returns psi."""
    Op = np.asarray(Op, dtype=np.float64)
    tr = None if trail is None else np.asarray(trail, dtype=np.float64)
    L = len(psi)
    nq = psi.nq
    zero = np.zeros(nq, dtype=np.int64)
    for p in sites:
        if not 1 <= p <= L:
            raise ValueError("applyOps!: site %d outside 1..%d" % (p, L))
        dq = _op_charge(Op, psi.qphys) if nq else zero
        _apply_site(psi, p, Op, zero, dq)
        if nq and np.any(dq != 0):
            for i in range(p + 1, L + 1):
                _apply_site(psi, i, np.eye(psi.d[i - 1]), dq, dq)
        if tr is not None and not np.array_equal(tr, np.ones((1, 1))):
            if nq and np.any(_op_charge(tr, psi.qphys) != 0):
                raise ValueError("applyOps!: the trail must be neutral")
            for i in range(1, p):
                _apply_site(psi, i, tr, zero, zero)
    return psi


def randMPS_normal(L, d, m, seed=0, ctx=None):
    """Random dense MPS: i.i.d. N(0,1) entries from numpy default_rng(seed),
    chi_b = min(d^b, d^(L-b), m), filled site 1 -> L in column-major order,
    right-canonicalised, oc = 1, site 1 normalised (compare
    lib/MPS/randMPS.jl:27-56, which draws uniform entries)."""
    rng = np.random.default_rng(seed)
    chi = [min(d ** b, d ** (L - b), m) for b in range(L + 1)]
    A = []
    for i in range(L):
        shape = (chi[i], d, chi[i + 1])
        A.append(rng.standard_normal(int(np.prod(shape))).reshape(shape, order="F"))
    A = canonicalize(A, 1)
    A[0] = A[0] / np.linalg.norm(A[0])
    return MPS(A, oc=1, ctx=ctx, canonical=True)


def makeq_host(psi_tensors, Qlabels, seed=0, qleft=None):
    """Host part of makeqMPS (lib/MPS/makeqMPS.jl:26-112) for a dense MPS:
    the charge of every right-link state is that of the largest-magnitude
    entry of its column (assignflux!, lib/MPS/MPS.jl:206-219: the first
    maximum with the left index outer and the physical index inner; an
    all-zero column takes the charge of its (1, 1) entry), entries off the charge-allowed blocks are dropped as the Qtens
    construction drops them, and a tensor left without any non-zero allowed
    entry is refilled with N(0,1) values (numpy default_rng(seed)) on its
    allowed blocks, as randomize=true does (makeqMPS.jl:66-67).  qleft: the
    charges of the first bond (default zero).  Returns (tensors, bond
    charges)."""
    qphys = np.asarray(Qlabels, dtype=np.int64)
    nq = qphys.shape[1]
    rng = np.random.default_rng(seed)
    q = np.zeros((psi_tensors[0].shape[0], nq), dtype=np.int64) if qleft is None else \
        np.asarray(qleft, dtype=np.int64).reshape(psi_tensors[0].shape[0], nq)
    qbond = [q]
    out = []
    for A in psi_tensors:
        A = np.array(A, dtype=np.float64)
        a, d, b = A.shape
        mag = np.abs(A).reshape(a * d, b)                  # rows (a, s), s fastest: the loop order of assignflux!
        rows = np.argmax(mag, axis=0)                      # first maximum in that order
        qin = (qbond[-1][:, None, :] + qphys[None, :, :])  # [a, s, nq]
        qr = qin.reshape(a * d, nq)[rows]
        ok = np.all(qin[:, :, None, :] == qr[None, None, :, :], axis=3)   # [a, s, b]
        A[~ok] = 0.0
        if not np.any(A != 0.0):
            A[ok] = rng.standard_normal(int(ok.sum()))
        out.append(A)
        qbond.append(qr)
    return out, qbond


def _dense_norm2(tensors):
    E = np.ones((1, 1))
    for A in tensors:
        E = np.einsum("xy,xsb,ysc->bc", E, A, A)
    return float(np.trace(E))


def makeq(psi_tensors, Qlabels, oc=1, ctx=None, newnorm=True, seed=0):
    """MPS(Qlabels, psi) (lib/MPS/MPS.jl:259-263, makeqMPS lib/MPS/makeqMPS.jl:
    26-112) for a dense host MPS: charges and dropped entries by makeq_host,
    then (newnorm) the tensor at the centre rescaled so that <psi|psi> is
    that of the dense input (makeqMPS.jl:81-84)."""
    tensors, qbond = makeq_host(psi_tensors, Qlabels, seed=seed)
    if newnorm:
        n0, n1 = _dense_norm2(psi_tensors), _dense_norm2(tensors)
        if n1 > 0.0:
            tensors[oc - 1] = tensors[oc - 1] * np.sqrt(n0 / n1)
    return MPS(tensors, oc=oc, qphys=np.asarray(Qlabels, dtype=np.int64), qbond=qbond, ctx=ctx, canonical=True)


# --------------------------------------------------------------- environment
class environment:
    """Vector of environment tensors, one per site (lib/types.jl:113-115).
    ``V[i]`` (1-based) is the left (or right) environment of site i."""

    def __init__(self, L):
        self.V = [None] * (L + 1)

    def __getitem__(self, i):
        return self.V[i]

    def __setitem__(self, i, v):
        old = self.V[i]
        self.V[i] = v
        if old is not None and old is not v and all(x is not old for x in self.V):
            old.free()

    def __len__(self):
        return len(self.V) - 1


def makeBoundary(ctx, left, nq=0, q=None):
    h = C.c_void_p()
    qa = None
    if nq:
        qa = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(1, nq))
    ctx.call("dmrj_env_boundary", 1 if left else 0, nq, _lib.iptr(qa), C.byref(h))
    return DeviceTensor(ctx, h)


def Lupdate(Lenv, A, mpo, i):
    """Lenv of site i+1 from Lenv of site i (lib/MPO_MPS/Lupdate.jl:17-26)."""
    ctx = A.ctx
    h = C.c_void_p()
    ctx.call("dmrj_env_left", Lenv.h, A.h, mpo.device(ctx, i), C.byref(h))
    return DeviceTensor(ctx, h)


def Rupdate(Renv, A, mpo, i):
    """Renv of site i-1 from Renv of site i (lib/MPO_MPS/Rupdate.jl:17-26)."""
    ctx = A.ctx
    h = C.c_void_p()
    ctx.call("dmrj_env_right", Renv.h, A.h, mpo.device(ctx, i), C.byref(h))
    return DeviceTensor(ctx, h)


def total_charge(psi):
    if not psi.nq:
        return None
    _, _, qr = psi[len(psi)].to_dense(psi.nq)
    return qr[0]


def makeEnv(psi, mpo):
    """lib/Env/makeEnv.jl:196-222: Lenv[1], Renv[L] boundaries, Lupdate! up to
    oc-1, Rupdate! from L down to oc+1."""
    ctx = psi.ctx
    L = len(psi)
    Lenv, Renv = environment(L), environment(L)
    nq = psi.nq
    Lenv[1] = makeBoundary(ctx, True, nq, np.zeros(nq, dtype=np.int64) if nq else None)
    Renv[L] = makeBoundary(ctx, False, nq, total_charge(psi) if nq else None)
    for i in range(1, psi.oc):
        Lenv[i + 1] = Lupdate(Lenv[i], psi[i], mpo, i)
    for i in range(L, psi.oc, -1):
        Renv[i - 1] = Rupdate(Renv[i], psi[i], mpo, i)
    return Lenv, Renv


def expect(psi, mpo):
    """<psi|H|psi> by a right-to-left sweep of Rupdate (expect.jl:19-45)."""
    ctx = psi.ctx
    L = len(psi)
    R = makeBoundary(ctx, False, psi.nq, total_charge(psi) if psi.nq else None)
    for i in range(L, 0, -1):
        R2 = Rupdate(R, psi[i], mpo, i)
        R.free()
        R = R2
    val, _, _ = R.to_dense(psi.nq)
    R.free()
    return float(np.real(val.ravel()[0]))
