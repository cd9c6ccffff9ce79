"""MPS containers, initial states, canonical form and charges (oracle).

TEST INFRASTRUCTURE ONLY.  Follows lib/types.jl:84-87 (matrixproductstate),
lib/MPS/randMPS.jl:27-56, lib/MPS/move.jl:124-197 (moveR!/moveL! by QR/LQ),
lib/MPS/MPS.jl:206-219 (assignflux!) and lib/MPS/makeqMPS.jl:26-112.

Index convention: A[alpha, s, beta], shape (chi_left, d, chi_right).
Charges (optional, for quantum-number runs): every bond b = 0..L carries an
integer array qbond[b] of shape (chi_b, nq); the physical index carries
qphys of shape (d, nq).  A block A[a, s, b] may be non-zero only when
qbond_left[a] + qphys[s] == qbond_right[b] (flux rule used throughout).
"""
import numpy as np


class MPS:
    """Oracle MPS: list of tensors, orthogonality centre ``oc`` (1-based as in
    lib/types.jl:84), optional bond/physical charges."""

    def __init__(self, tensors, oc=1, qphys=None, qbond=None):
        self.A = [np.array(t, dtype=np.float64 if not np.iscomplexobj(t) else np.complex128)
                  for t in tensors]
        self.oc = oc
        self.qphys = None if qphys is None else np.asarray(qphys, dtype=np.int64)
        self.qbond = None if qbond is None else [np.asarray(q, dtype=np.int64) for q in qbond]

    def __len__(self):
        return len(self.A)

    def __getitem__(self, i):          # 1-based, as the reference's getindex
        return self.A[i - 1]

    def __setitem__(self, i, v):
        self.A[i - 1] = v

    def copy(self):
        return MPS([a.copy() for a in self.A], self.oc, self.qphys,
                   None if self.qbond is None else [q.copy() for q in self.qbond])

    @property
    def qn(self):
        return self.qbond is not None


def seeded_normal_mps(L, d, m, seed=0):
    """Random MPS with i.i.d. N(0,1) entries from numpy ``default_rng(seed)``,
    bond dimensions chi_b = min(d^b, d^(L-b), m), tensors filled site 1 -> L,
    each in column-major (Fortran) order.  Right-canonicalised by LQ from site
    L down to 2 (lib/MPS/move.jl:124-146), oc = 1, then site 1 normalised."""
    rng = np.random.default_rng(seed)
    chi = [min(d ** b, d ** (L - b), m) for b in range(L + 1)]
    A = []
    for i in range(L):
        shape = (chi[i], d, chi[i + 1])
        flat = rng.standard_normal(int(np.prod(shape)))
        A.append(flat.reshape(shape, order='F'))
    psi = MPS(A, oc=L)
    move(psi, 1)
    psi.A[0] = psi.A[0] / np.linalg.norm(psi.A[0])
    return psi


def product_mps(states, d, qphys=None, dtype=np.float64):
    """Product state |s_1 s_2 ...> (test/dmrg_test.jl:55-60).  With qphys the
    bond charges are the running sums of the physical charges."""
    L = len(states)
    A = []
    for s in states:
        t = np.zeros((1, d, 1), dtype=dtype)
        t[0, s, 0] = 1.0
        A.append(t)
    qbond = None
    if qphys is not None:
        qphys = np.asarray(qphys, dtype=np.int64)
        nq = qphys.shape[1]
        q = np.zeros((1, nq), dtype=np.int64)
        qbond = [q.copy()]
        for s in states:
            q = q + qphys[s][None, :]
            qbond.append(q.copy())
    return MPS(A, oc=1, qphys=qphys, qbond=qbond)


def _qr_site(A):
    """moveR! by QR (lib/MPS/move.jl:175-182): A[(a,s), b] = Q R."""
    a, d, b = A.shape
    M = A.reshape(a * d, b)
    Q, R = np.linalg.qr(M)
    return Q.reshape(a, d, Q.shape[1]), R


def _lq_site(A):
    """moveL! by LQ (lib/MPS/move.jl:126-131): A[a, (s,b)] = L Q."""
    a, d, b = A.shape
    M = A.reshape(a, d * b)
    Q, R = np.linalg.qr(M.T)
    return R.T, Q.T.reshape(Q.shape[1], d, b)


def move(psi, newoc):
    """lib/MPS/move.jl:19-49 (movecenter!) using QR/LQ (m=0, cutoff=0).
    Dense MPS only (the charge-labelled oracle states are built canonical)."""
    assert not psi.qn
    while psi.oc < newoc:
        i = psi.oc
        Q, R = _qr_site(psi[i])
        psi[i] = Q
        psi[i + 1] = np.tensordot(R, psi[i + 1], axes=(1, 0))
        psi.oc += 1
    while psi.oc > newoc:
        i = psi.oc
        Lm, Q = _lq_site(psi[i])
        psi[i] = Q
        psi[i - 1] = np.tensordot(psi[i - 1], Lm, axes=(2, 0))
        psi.oc -= 1
    return psi


def apply_ops(psi, sites, op, trail=None):
    """lib/MPO_MPS/applyMPO.jl:21-33 (applyOps!): apply ``op`` to each listed
    site of a product state, with a Jordan-Wigner ``trail`` on sites to the
    left.  The operator acts as O|s> with O[bra, ket] in the usual matrix
    convention."""
    for i in sites:
        A = psi[i]
        psi[i] = np.einsum('kj,ajb->akb', op, A)
        if trail is not None:
            for w in range(1, i):
                psi[w] = np.einsum('kj,ajb->akb', trail, psi[w])
    return psi


def product_charges(psi, qphys):
    """Attach bond charges to a product state (bond dim 1 everywhere) from the
    occupied physical state of every site (lib/MPS/MPS.jl:206-219 picks the
    charge of the largest-magnitude entry; for a product state that is the
    occupied state)."""
    qphys = np.asarray(qphys, dtype=np.int64)
    nq = qphys.shape[1]
    q = np.zeros((1, nq), dtype=np.int64)
    qbond = [q.copy()]
    for i in range(1, len(psi) + 1):
        A = psi[i]
        assert A.shape[0] == 1 and A.shape[2] == 1
        s = int(np.argmax(np.abs(A[0, :, 0])))
        q = q + qphys[s][None, :]
        qbond.append(q.copy())
    psi.qphys = qphys
    psi.qbond = qbond
    return psi


def mpo_deltas(mpo, qphys):
    """Charge carried by every MPO link (lib/MPO/makeqMPO.jl:53-74 restated):
    W[w, j, k, u] may be non-zero only when delta[u] = delta[w] + q[k] - q[j];
    the left edge link has delta 0.  Returns a list of L+1 arrays (D_b, nq)."""
    qphys = np.asarray(qphys, dtype=np.int64)
    nq = qphys.shape[1]
    deltas = [np.zeros((mpo[0].shape[0], nq), dtype=np.int64)]
    for W in mpo:
        Dl, d, _, Dr = W.shape
        dl = deltas[-1]
        dr = np.zeros((Dr, nq), dtype=np.int64)
        seen = np.zeros(Dr, dtype=bool)
        nz = np.argwhere(np.abs(W) > 0)
        for (w, j, k, u) in nz:
            val = dl[w] + qphys[k] - qphys[j]
            if not seen[u]:
                dr[u] = val
                seen[u] = True
            elif not np.array_equal(dr[u], val):
                raise ValueError("MPO does not conserve the charge")
        deltas.append(dr)
    return deltas
