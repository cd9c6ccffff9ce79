"""Correlation matrix and general correlations on full state vectors (oracle).

TEST INFRASTRUCTURE ONLY.  Restates lib/MPO_MPS/correlationmatrix.jl:21-83 and
lib/MPO_MPS/correlation.jl:74-237 by brute force on the 2^L (d^L) amplitude
vector.  correlationmatrix:
    rho[i, j] = <psi| (Cc F)_i F_{i+1} ... F_{j-1} (Ca)_j |psi>   (i < j)
    rho[i, i] = <psi| (Cc Ca)_i |psi>,   rho[j, i] = conj(rho[i, j])
with operators acting on the ket (the reference contracts Cc's second index
with the ket's physical index, :35-36,44) and F only when a trail is given.

Pinned (tests/test_oracle_pins.py) by the reference's golden matrix of the
L=10 Heisenberg ground state (test/MPmeas_test.jl:310-315, with the
reference's tolerance norm/length < 1e-5: the golden itself is within 7e-6 of
the exact state) and by exact identities: the nearest-neighbour correlations
of that state sum to its energy, its diagonal is 1/2; on the Neel product
state rho(S+, S-) is diag(1, 0, 1, ...); for free fermions the matrix with
the Jordan-Wigner trail is the projector on the occupied orbitals (without
the trail it is not).  Vector index convention: site 1
is the fastest index (Julia's column-major reshape of the MPS product).

transition_matrix: the same sums between two vectors, <bra| ... |ket>, not
divided by any norm (correlationmatrix(dualpsi, psi, Cc, Ca; trail),
correlationmatrix.jl:21-62).

correlation(bra, ket, ops, trail) (correlation.jl:92-237, default sites):
    omega[p1, ..., pn] = <bra| X |ket>,  X = O1_{p1} T1_{<p1} On_{pn} Tn_{<pn}
    ... O2_{p2} T2_{<p2} applied to the ket in the reference's order (operator
    n and its trail first, operator 1 last), T_g the trail of operator g on
    every site left of p_g (none when not given).  Pinned by: the i < j and
    i = j entries equal correlationmatrix (one trail shared by both
    operators); for free fermions with the Jordan-Wigner trail the whole
    matrix, i > j included, is the occupied-orbital projector (a dropped
    trail fails); on the Neel state a three-point Sz function is the product
    of the local values.
"""
import numpy as np


def mps_to_vector(tensors):
    """Amplitude vector of an open MPS [a, s, b] with site 1 fastest."""
    v = np.ones((1, 1))
    for A in tensors:
        # v[rest, a] x A[a, s, b] -> v[rest + s * |rest|, b]
        t = np.tensordot(v, A, axes=(1, 0))          # [rest, s, b]
        v = t.reshape(-1, A.shape[2], order="F")
    return v[:, 0]


def vector_to_mps(vec, d, L):
    """Left-canonical open MPS of a full vector (site 1 fastest), by
    successive SVDs; the norm sits in the last tensor."""
    out = []
    rest = np.asarray(vec).reshape(-1, 1)
    chi = 1
    for site in range(L - 1):
        # rest has shape [chi * d^(L-site), 1] with (chi, s_site, remaining) order
        M = rest.reshape(chi * d, -1, order="F")
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
        k = int(np.sum(S > 1e-14 * S[0])) or 1
        out.append(U[:, :k].reshape(chi, d, k, order="F"))
        rest = (S[:k, None] * Vt[:k]).reshape(-1, 1, order="F")
        chi = k
    out.append(rest.reshape(chi, d, 1, order="F"))
    return out


def _apply(vec, op, site, d, L):
    """op (d x d) on the ket at `site` (1-based)."""
    t = vec.reshape((d,) * L, order="F")
    t = np.moveaxis(np.tensordot(op, t, axes=(1, site - 1)), 0, site - 1)
    return t.reshape(-1, order="F")


def correlationmatrix(vec, d, L, Cc, Ca, trail=None):
    Cc = np.asarray(Cc)
    Ca = np.asarray(Ca)
    rho = np.zeros((L, L), dtype=np.result_type(vec, Cc, Ca))
    nrm = np.vdot(vec, vec)
    for i in range(1, L + 1):
        rho[i - 1, i - 1] = np.vdot(vec, _apply(vec, Cc @ Ca, i, d, L)) / nrm
    FCc = Cc @ trail if trail is not None else Cc
    for i in range(1, L):
        for j in range(i + 1, L + 1):
            w = _apply(vec, Ca, j, d, L)
            if trail is not None:
                for k in range(i + 1, j):
                    w = _apply(w, trail, k, d, L)
            w = _apply(w, FCc, i, d, L)
            v = np.vdot(vec, w) / nrm
            rho[i - 1, j - 1] = v
            rho[j - 1, i - 1] = np.conj(v)
    return rho


def transition_matrix(bra, ket, d, L, Cc, Ca, trail=None):
    """<bra| (Cc F)_i F ... F (Ca)_j |ket> for i < j, <bra| (Cc Ca)_i |ket> on
    the diagonal, and the lower triangle as the reference fills it:
    rho[j, i] = conj(rho[i, j])."""
    Cc = np.asarray(Cc)
    Ca = np.asarray(Ca)
    rho = np.zeros((L, L), dtype=np.result_type(bra, ket, Cc, Ca))
    for i in range(1, L + 1):
        rho[i - 1, i - 1] = np.vdot(bra, _apply(ket, Cc @ Ca, i, d, L))
    FCc = Cc @ trail if trail is not None else Cc
    for i in range(1, L):
        for j in range(i + 1, L + 1):
            w = _apply(ket, Ca, j, d, L)
            if trail is not None:
                for k in range(i + 1, j):
                    w = _apply(w, trail, k, d, L)
            w = _apply(w, FCc, i, d, L)
            v = np.vdot(bra, w)
            rho[i - 1, j - 1] = v
            rho[j - 1, i - 1] = np.conj(v)
    return rho


def correlation(bra, ket, d, L, ops, trail=None):
    """omega[p1-1, ..., pn-1] over all position tuples (see the module
    docstring); trail: None, one matrix for every operator, or a tuple with
    one entry per operator (None = no trail for that operator)."""
    n = len(ops)
    if trail is None or isinstance(trail, np.ndarray):
        trails = [trail] * n
    else:
        trails = [trail[g % len(trail)] for g in range(n)]
    omega = np.zeros((L,) * n, dtype=np.result_type(bra, ket, *ops))
    for pos in np.ndindex(*(L,) * n):
        w = ket
        for g in list(range(1, n)) [::-1] + [0]:
            p = pos[g] + 1
            if g > 0:
                w = _apply(w, ops[g], p, d, L)
                if trails[g] is not None:
                    for s in range(1, p):
                        w = _apply(w, trails[g], s, d, L)
            else:
                if trails[0] is not None:
                    for s in range(1, p):
                        w = _apply(w, trails[0], s, d, L)
                w = _apply(w, ops[0], p, d, L)
        omega[pos] = np.vdot(bra, w)
    return omega
