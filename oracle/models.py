"""Local operators and MPOs (oracle).  TEST INFRASTRUCTURE ONLY.

Follows lib/models.jl:30-55 (qubitOps), :95-105 (spinOps), :142-164
(fermionOps), :237-244 (heisenbergMPO), :264-273 (hubbardMPO) and
lib/MPO/makeMPO.jl:59-92 (makeMPO).

Index convention of an MPO tensor: W[a1, j, k, a2] with j the ket physical
index and k the bra physical index, built as
W[l, j, k, m] = H[j + l*d, k + m*d] (0-based form of makeMPO.jl:76).
The first site keeps the last row and the last site the first column
(makeMPO.jl:85-86).
"""
import numpy as np


def qubitOps(d=2):
    """lib/models.jl:30-55: Pauli-like operators; state 0 has the largest m."""
    s = (d - 1) / 2
    O = np.zeros((d, d))
    Id = np.eye(d)
    oz = O.copy()
    op = O.copy()
    for q, m in enumerate(np.arange(s, -s - 1, -1)):
        oz[q, q] = 2 * m
        if m + 1 <= s:
            op[q - 1, q] = np.sqrt(s * (s + 1) - m * (m + 1))
    om = op.T.copy()
    ox = op + om
    return ox, oz, Id, op, om, O


def spinOps(s=0.5):
    """lib/models.jl:95-105: returns Sp, Sm, Sz, O, Id (real operators only)."""
    d = int(round(2 * s + 1))
    ox, oz, Id, op, om, O = qubitOps(d)
    return op, om, oz * s, O, Id


def fermionOps():
    """lib/models.jl:142-164: states |0>, |up>, |dn>, |updn>."""
    O = np.zeros((4, 4))
    Id = np.eye(4)
    Cup = O.copy()
    Cup[0, 1] = 1.0
    Cup[2, 3] = 1.0
    Cdn = O.copy()
    Cdn[0, 2] = 1.0
    Cdn[1, 3] = -1.0
    Nup = Cup.T @ Cup
    Ndn = Cdn.T @ Cdn
    Ndens = Nup + Ndn
    F = Id.copy()
    F[1, 1] = -1.0
    F[2, 2] = -1.0
    return Cup, Cdn, F, Nup, Ndn, Ndens, O, Id


def _blockmatrix(rows):
    return np.block([[np.asarray(x, dtype=np.float64) for x in r] for r in rows])


def heisenbergMPO(J=1.0, spinmag=0.5):
    """lib/models.jl:237-244: lower-triangular bulk matrix, D=5."""
    Sp, Sm, Sz, O, Id = spinOps(spinmag)
    return _blockmatrix([[Id, O, O, O, O],
                         [J * 0.5 * Sm, O, O, O, O],
                         [J * 0.5 * Sp, O, O, O, O],
                         [J * Sz, O, O, O, O],
                         [O, Sp, Sm, Sz, Id]])


def hubbardMPO(t=1.0, mu=-2.0, HubU=4.0):
    """lib/models.jl:264-273: lower-triangular bulk matrix, D=6."""
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = fermionOps()
    onsite = mu * Ndens + HubU * Nup @ Ndn
    return _blockmatrix([[Id, O, O, O, O, O],
                         [-t * Cup.T, O, O, O, O, O],
                         [t * Cup, O, O, O, O, O],
                         [-t * Cdn.T, O, O, O, O, O],
                         [t * Cdn, O, O, O, O, O],
                         [onsite, Cup @ F, Cup.T @ F, Cdn @ F, Cdn.T @ F, Id]])


def makeMPO(H, d, Ns):
    """lib/MPO/makeMPO.jl:59-92 for a uniform bulk matrix (or a list of them).

    Returns a list of W[a1, j, k, a2] arrays.
    """
    Hs = H if isinstance(H, (list, tuple)) else [H] * Ns
    out = []
    for i in range(Ns):
        Hi = np.asarray(Hs[i], dtype=np.float64)
        a1 = Hi.shape[0] // d
        a2 = Hi.shape[1] // d
        # G[l, j, k, m] = H[j + l*d, k + m*d]
        G = Hi.reshape(a1, d, a2, d).transpose(0, 1, 3, 2).copy()
        out.append(G)
    out[0] = out[0][-1:, :, :, :].copy()
    out[-1] = out[-1][:, :, :, :1].copy()
    return out


def j1j2_cylinder_mpo(Lx, Ly, J1=1.0, J2=0.5):
    """J1-J2 Heisenberg model on an Lx x Ly cylinder (periodic in y), snaked
    column-major (site index = y + Ly*x), as a finite-state-machine MPO in the
    same lower-triangular convention as heisenbergMPO.

    Bonds: nearest neighbours (J1) along x and along y (periodic), next-nearest
    (diagonal) neighbours (J2).  Each two-site term S_a.S_b with a < b is
    carried through channels that are keyed by the distance b - a; the bulk
    dimension is D = 2 + 3*R with R the largest distance.  The reference builds
    such MPOs with autoMPO (lib/MPO/autoMPO.jl); this restatement writes the
    finite-state machine directly.
    """
    Sp, Sm, Sz, O, Id = spinOps(0.5)
    N = Lx * Ly
    bonds = {}

    def add(a, b, J):
        if a == b:
            return
        a, b = min(a, b), max(a, b)
        bonds[(a, b)] = bonds.get((a, b), 0.0) + J

    for x in range(Lx):
        for y in range(Ly):
            i = y + Ly * x
            # vertical (periodic)
            if Ly > 1:
                yy = (y + 1) % Ly
                if not (Ly == 2 and y == 1):
                    add(i, yy + Ly * x, J1)
            if x + 1 < Lx:
                add(i, y + Ly * (x + 1), J1)
                if J2 != 0.0 and Ly > 1:
                    add(i, (y + 1) % Ly + Ly * (x + 1), J2)
                    if Ly > 2:
                        add(i, (y - 1) % Ly + Ly * (x + 1), J2)
    R = max(b - a for (a, b) in bonds)
    D = 2 + 3 * R
    # channel layout: 0 = done (identity to the right), 1..R = S+ travelling,
    # R+1..2R = S- travelling, 2R+1..3R = Sz travelling, D-1 = not started.
    ops = []
    for i in range(N):
        H = [[O for _ in range(D)] for _ in range(D)]
        H[0][0] = Id
        H[D - 1][D - 1] = Id
        for r in range(1, R + 1):
            # a string started r sites ago arrives here: close with coupling
            J = bonds.get((i - r, i), 0.0)
            if J != 0.0:
                H[r][0] = J * 0.5 * Sm
                H[R + r][0] = J * 0.5 * Sp
                H[2 * R + r][0] = J * Sz
            if r < R:
                H[r][r + 1] = Id
                H[R + r][R + r + 1] = Id
                H[2 * R + r][2 * R + r + 1] = Id
        H[D - 1][1] = Sp
        H[D - 1][R + 1] = Sm
        H[D - 1][2 * R + 1] = Sz
        ops.append(_blockmatrix(H))
    return makeMPO(ops, 2, N), bonds


def fullH(mpo):
    """lib/MPO/fullmpo.jl:17-57: dense Hamiltonian of an MPO (small chains only).

    Basis ordering: site 1 is the fastest index (column-major, as Julia).
    Returned as H[bra, ket] with ket index j and bra index k of W.
    """
    T = mpo[0][0]  # (j, k, a2)
    T = T.reshape(T.shape[0], T.shape[1], T.shape[2])
    # T[ket_multi, bra_multi, a]
    for W in mpo[1:]:
        # new[(jprev, j), (kprev, k), a2] = sum_a T[jprev, kprev, a] W[a, j, k, a2]
        X = np.einsum('pqa,ajkb->jpkqb', T, W)
        n1 = X.shape[0] * X.shape[1]
        n2 = X.shape[2] * X.shape[3]
        T = X.reshape(n1, n2, X.shape[4])
    Hkj = T[:, :, 0]          # [ket, bra]
    # multi-index built as (j_new slowest ... j_1 fastest)
    return Hkj.T.copy()        # [bra, ket]
