"""Local operators and bulk MPO matrices (host side, built once per run).

Mirrors lib/models.jl: qubitOps (:30-55), spinOps (:95-105), fermionOps
(:142-164), heisenbergMPO (:237-244), hubbardMPO (:264-273), and the
lower-triangular bulk-matrix convention consumed by makeMPO
(lib/MPO/makeMPO.jl:59-92).  The J1-J2 cylinder MPO (a model the reference
builds with autoMPO, lib/MPO/autoMPO.jl) is written here as an explicit
finite-state machine in the same convention.
"""
import numpy as np


def qubitOps(d=2):
    """Returns (ox, oz, Id, op, om, O); state 0 carries the largest m."""
    s = (d - 1) / 2
    O = np.zeros((d, d))
    Id = np.eye(d)
    oz, op = O.copy(), O.copy()
    for q in range(d):
        mval = s - q
        oz[q, q] = 2 * mval
        if q > 0:
            op[q - 1, q] = np.sqrt(s * (s + 1) - mval * (mval + 1))
    om = op.T.copy()
    return op + om, oz, Id, op, om, O


def spinOps(s=0.5):
    """Sp, Sm, Sz, O, Id for spin s (lib/models.jl:95-105)."""
    d = int(round(2 * s + 1))
    ox, oz, Id, op, om, O = qubitOps(d)
    return op, om, oz * s, O, Id


def fermionOps():
    """Cup, Cdn, F, Nup, Ndn, Ndens, O, Id on |0>,|up>,|dn>,|updn>
    (lib/models.jl:142-164)."""
    O = np.zeros((4, 4))
    Id = np.eye(4)
    Cup = O.copy()
    Cup[0, 1] = Cup[2, 3] = 1.0
    Cdn = O.copy()
    Cdn[0, 2] = 1.0
    Cdn[1, 3] = -1.0
    Nup = Cup.T @ Cup
    Ndn = Cdn.T @ Cdn
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    return Cup, Cdn, F, Nup, Ndn, Nup + Ndn, O, Id


def _bm(rows):
    return np.block([[np.asarray(x, dtype=np.float64) for x in r] for r in rows])


def heisenbergMPO(i=1, spinmag=0.5, J=1.0):
    Sp, Sm, Sz, O, Id = spinOps(spinmag)
    return _bm([[Id, O, O, O, O],
                [J * 0.5 * Sm, O, O, O, O],
                [J * 0.5 * Sp, O, O, O, O],
                [J * Sz, O, O, O, O],
                [O, Sp, Sm, Sz, Id]])


XXZ = heisenbergMPO


def hubbardMPO(i=1, t=1.0, mu=-2.0, HubU=4.0):
    Cup, Cdn, F, Nup, Ndn, Ndens, O, Id = fermionOps()
    onsite = mu * Ndens + HubU * Nup @ Ndn
    return _bm([[Id, O, O, O, O, O],
                [-t * Cup.T, O, O, O, O, O],
                [t * Cup, O, O, O, O, O],
                [-t * Cdn.T, O, O, O, O, O],
                [t * Cdn, O, O, O, O, O],
                [onsite, Cup @ F, Cup.T @ F, Cdn @ F, Cdn.T @ F, Id]])


def j1j2_bulk(Lx, Ly, J1=1.0, J2=0.5):
    """Per-site bulk matrices of the J1-J2 Heisenberg model on an Lx x Ly
    cylinder (periodic along y), sites snaked column-major (i = y + Ly*x).

    Channels: 0 = done, 1..R = S+ string of length r, R+1..2R = S- string,
    2R+1..3R = Sz string, D-1 = not started; D = 3R + 2 with R the largest
    site distance of a coupled pair.
    """
    Sp, Sm, Sz, O, Id = spinOps(0.5)
    N = Lx * Ly
    bonds = {}

    def add(a, b, J):
        if a != b:
            key = (min(a, b), max(a, b))
            bonds[key] = bonds.get(key, 0.0) + J

    for x in range(Lx):
        for y in range(Ly):
            i = y + Ly * x
            if Ly > 2 or (Ly == 2 and y == 0):
                add(i, (y + 1) % Ly + Ly * x, J1)
            if x + 1 < Lx:
                add(i, y + Ly * (x + 1), J1)
                if J2 != 0.0 and Ly > 1:
                    add(i, (y + 1) % Ly + Ly * (x + 1), J2)
                    if Ly > 2:
                        add(i, (y - 1) % Ly + Ly * (x + 1), J2)
    R = max(b - a for (a, b) in bonds)
    D = 3 * R + 2
    mats = []
    for i in range(N):
        H = [[O] * D for _ in range(D)]
        H[0][0] = Id
        H[D - 1][D - 1] = Id
        for r in range(1, R + 1):
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
        mats.append(_bm(H))
    return mats, bonds
