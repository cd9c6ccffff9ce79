"""Writes tests/golden/cfg1.json: the oracle's per-sweep report of the
L=20 open Heisenberg chain (J=1, d=2, D=5), m=64, 6 sweeps, cutoff 1e-12,
minm=2, from the seed-0 N(0,1) initial MPS, plus the exact ground-state energy
(scipy eigsh in the Sz=0 sector).  The reference itself cannot run here (no
Julia, TensorPACK not vendored), so this fixture pins the oracle against
regressions; the ED value pins where the sweeps converge."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import dmrg as od  # noqa: E402
from oracle import models as om  # noqa: E402
from oracle import mps as omps  # noqa: E402


def ed_sz0(L):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    states = [s for s in range(1 << L) if bin(s).count("1") == L // 2]
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for s in states:
        i = index[s]
        diag = 0.0
        for b in range(L - 1):
            u, v = (s >> b) & 1, (s >> (b + 1)) & 1
            if u == v:
                diag += 0.25
            else:
                diag -= 0.25
                t = s ^ (1 << b) ^ (1 << (b + 1))
                rows.append(i)
                cols.append(index[t])
                vals.append(0.5)
        rows.append(i)
        cols.append(i)
        vals.append(diag)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    return float(sla.eigsh(H, k=1, which="SA")[0][0])


def main():
    psi = omps.seeded_normal_mps(20, 2, 64, seed=0)
    first4 = psi[1].ravel()[:4].tolist()
    mpo = om.makeMPO(om.heisenbergMPO(), 2, 20)
    E, hist = od.dmrg(psi, mpo, m=64, sweeps=6, cutoff=1e-12)
    out = dict(config="open Heisenberg L=20 J=1 d=2 D=5 m=64 minm=2 sweeps=6 cutoff=1e-12 seed=0",
               init_site1_first4=first4, sweeps=hist, final_energy=E, ed_energy=ed_sz0(20))
    with open(os.path.join(os.path.dirname(__file__), "cfg1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out["sweeps"][-1]), out["ed_energy"])


if __name__ == "__main__":
    main()
