"""Writes fs_bcc_lattice.json: closed-form energies of perfect and
homogeneously strained BCC lattices under the Finnis-Sinclair potential for Fe
(M. W. Finnis and J. E. Sinclair, Phil. Mag. A 50 (1984) 45, Table 1).

Every value comes from the analytic functions by a direct sum over the
lattice vectors within the cutoff -- no spline, no table, no code of the
library or the oracle -- so the fixture pins the whole chain (table
generation, spline, stencil, passes, energy reduction) from outside.

    rho(r)  = (r - d)^2 + beta (r - d)^3 / d   for r < d
    V(r)    = (r - c)^2 (c0 + c1 r + c2 r^2)   for r < c
    E_atom  = -A sqrt(sum_j rho(r_j)) + 1/2 sum_j V(r_j)

Run from the repository root:  python3 tests/golden/make_fs_lattice.py
"""
import json
import os

import numpy as np

D, A, BETA, C = 3.569745, 1.828905, 1.8, 3.40
C0, C1, C2 = 1.2371147, -0.3592185, -0.0385607
RC = 5.6

# (name, lattice constant, diagonal strain)
CASES = [
    ("a2.810", 2.810, (0.0, 0.0, 0.0)),
    ("a2.855", 2.855, (0.0, 0.0, 0.0)),
    ("a2.8665", 2.8665, (0.0, 0.0, 0.0)),
    ("a2.900", 2.900, (0.0, 0.0, 0.0)),
    ("a2.855_tetragonal", 2.855, (0.01, -0.005, -0.005)),
    ("a2.855_orthorhombic", 2.855, (0.008, -0.003, 0.002)),
]


def lattice_vectors(a, strain, rmax):
    n = int(np.ceil(rmax / a)) + 2
    g = np.arange(-n, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    corner = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(float)
    pts = np.concatenate([corner, corner + 0.5]) * a * (1.0 + np.asarray(strain))
    r = np.linalg.norm(pts, axis=1)
    return r[(r > 0.0) & (r < rmax)]


def energy(a, strain):
    r = lattice_vectors(a, strain, RC)
    q = r - D
    rho = np.where(r < D, q * q * (1.0 + BETA * q / D), 0.0).sum()
    p = r - C
    pair = np.where(r < C, p * p * (C0 + C1 * r + C2 * r * r), 0.0).sum()
    emb = -A * np.sqrt(rho)
    return dict(rho=rho, embed=emb, pair_half=0.5 * pair, energy=emb + 0.5 * pair, neighbours_in_rc=int(r.size))


def main():
    out = dict(potential="Finnis-Sinclair Fe, Phil. Mag. A 50 (1984) 45",
               parameters=dict(d=D, A=A, beta=BETA, c=C, c0=C0, c1=C1, c2=C2), cutoff=RC, cases=[])
    for name, a, strain in CASES:
        e = energy(a, strain)
        out["cases"].append(dict(name=name, lattice_const=a, strain=list(strain), **e))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fs_bcc_lattice.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
