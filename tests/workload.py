"""Shared workload builders for the tests (lattice, tables, stencils).

Everything here calls the product's host helpers (stencil generator, synthetic
tables, initial conditions) and numpy; device calls live in the GPU tests.
"""
import numpy as np

import mmd

A = 2.855
RC = 5.6
SKIN = 0.3
HARTREE_BOHR = 27.2 * 0.529  # r phi_ij = 27.2 * 0.529 * Z_i Z_j in eV * angstrom (LAMMPS pair_style eam)


def write_funcfl(path, z_scale=1.0, nr=600, dr=0.01, nrho=500, drho=0.05, C=0.01, mass=55.845):
    """One-element funcfl file: F(rho) = -sqrt(rho) + C rho^2, Z(r) = 0.1 z_scale
    (rc - r)^2 exp(-r/2), rho(r) = (rc - r)^2 exp(-r), cut at rc (values
    printed with 17 significant digits, so they read back exactly)."""
    r = np.arange(nr) * dr
    rho = np.arange(nrho) * drho
    F = -np.sqrt(rho) + C * rho ** 2
    Z = np.where(r < RC, 0.1 * z_scale * (RC - r) ** 2 * np.exp(-0.5 * r), 0.0)
    rr = np.where(r < RC, (RC - r) ** 2 * np.exp(-r), 0.0)
    lines = ["synthetic funcfl", f"26 {mass!r} 2.855 BCC", f"{nrho} {drho!r} {nr} {dr!r} {RC!r}"]
    for a in (F, Z, rr):
        for k in range(0, len(a), 5):
            lines.append(" ".join("%.17g" % v for v in a[k:k + 5]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return dict(F=F, Z=Z, rho=rr)


def _distinct_rho(tab, step=0.15):
    """The tables with rho_t scaled by 1 + step * t: the named synthetic and the
    Finnis-Sinclair forms share one rho(r) across types, which would leave the
    rule "rho by the neighbour's type" (src/kernels/kernel_pairs.hpp:26)
    untested."""
    out = dict(tab)
    out["rho_r"] = np.asarray(tab["rho_r"], float) * (1.0 + step * np.arange(tab["n_types"]))[:, None]
    return out


def case(cells=(6, 6, 6), n_types=1, sigma=0.03, temperature=600.0, vacancy_frac=0.0, interstitial_frac=0.0,
         seed=7, ghost_cells=3, jitter=None, n_r=5000, n_rho=5000, n_blobs=0, blob_sigma=8.0, skin=SKIN,
         distinct_rho=False):
    spec = mmd.lattice_spec(cells, ghost_cells=ghost_cells, a=A, n_types=n_types, sigma=sigma,
                            temperature=temperature, vacancy_frac=vacancy_frac, interstitial_frac=interstitial_frac,
                            n_blobs=n_blobs, blob_sigma=blob_sigma, seed=seed)
    cap = int(interstitial_frac * 2 * np.prod(cells)) + 16
    dom, box, st, itl = mmd.lattice_init(spec, itl_cap=cap)
    tab = mmd.fs_tables(n_types=n_types, rc=RC, jitter=(0.1 if n_types > 1 else 0.0) if jitter is None else jitter,
                        seed=seed, n_r=n_r, n_rho=n_rho)
    if distinct_rho:
        tab = _distinct_rho(tab)
    st3 = mmd.stencil_build(A, RC + skin, 0)
    even = mmd.stencil_linear(dom, st3)
    odd = mmd.stencil_linear(dom, -st3)
    wide3 = mmd.wide_stencil(A, RC, skin)
    return dict(dom=dom, box=box, st=st, itl=itl, tab=tab, st3=st3, even=even, odd=odd, wide3=wide3,
                masses=mmd.DEFAULT_MASSES, spec=spec)


def local_vectors(st, idx=None):
    li = st.local_index() if idx is None else idx
    return li, np.stack([st.fx[li], st.fy[li], st.fz[li]], 1)


# -------------------------------------------------- numpy restatement of the spline

def np_spline(f, dx):
    """Cubic Hermite with 4th-order finite-difference slopes (LAMMPS pair_eam
    interpolate), vectorised independently of the C oracle."""
    f = np.asarray(f, float)
    n = len(f)
    s = np.empty(n)
    s[0] = f[1] - f[0]
    s[1] = 0.5 * (f[2] - f[0])
    s[n - 2] = 0.5 * (f[n - 1] - f[n - 3])
    s[n - 1] = f[n - 1] - f[n - 2]
    s[2:n - 2] = ((f[0:n - 4] - f[4:n]) + 8.0 * (f[3:n - 1] - f[1:n - 3])) / 12.0
    d = f[1:] - f[:-1]
    c4 = np.zeros(n)
    c3 = np.zeros(n)
    c4[:-1] = 3.0 * d - 2.0 * s[:-1] - s[1:]
    c3[:-1] = s[:-1] + s[1:] - 2.0 * d
    return dict(f=f, s=s, c4=c4, c3=c3, d2=s / dx, d1=2.0 * c4 / dx, d0=3.0 * c3 / dx, n=n)


def np_seg(sp, x0, inv_dx, x):
    p = (np.asarray(x, float) - x0) * inv_dx
    p = np.maximum(p, 0.0)
    m = np.minimum(p.astype(np.int64), sp["n"] - 2)
    p = np.minimum(p - m, 1.0)
    return m, p


def np_value(sp, x0, inv_dx, x):
    m, p = np_seg(sp, x0, inv_dx, x)
    return ((sp["c3"][m] * p + sp["c4"][m]) * p + sp["s"][m]) * p + sp["f"][m]


def np_deriv(sp, x0, inv_dx, x):
    m, p = np_seg(sp, x0, inv_dx, x)
    return (sp["d0"][m] * p + sp["d1"][m]) * p + sp["d2"][m]


class NpTables:
    def __init__(self, tab):
        nt = tab["n_types"]
        self.nt = nt
        self.r0, self.idr = tab["r0"], 1.0 / tab["dr"]
        self.p0, self.idp = tab["rho0"], 1.0 / tab["drho"]
        self.rho = [np_spline(tab["rho_r"][t], tab["dr"]) for t in range(nt)]
        self.emb = [np_spline(tab["emb"][t], tab["drho"]) for t in range(nt)]
        self.phi = [[np_spline(tab["phi_r"][a][b], tab["dr"]) for b in range(nt)] for a in range(nt)]


def brute_force(pos, types, L, nt_tab, rc, slip=None):
    """O(N^2) minimum-image EAM: rho, df, forces, total potential energy.

    slip plants one plausible mistake (to show the comparison with the oracle
    catches it): "rho_centre_type" (rho by the centre's type), "swap_df"
    (df_i and df_j transposed), "swap_rho_types" (rho'_ti and rho'_tj
    transposed in the force), "sign" (fpair sign flipped), "drop_phi" (pair
    term left out of the force), "energy_double" (pair energy counted twice)."""
    T = nt_tab
    pos = np.asarray(pos, float)
    types = np.asarray(types, int)
    N = len(pos)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    mask = (r2 < rc * rc) & ~np.eye(N, dtype=bool)
    r = np.sqrt(np.where(mask, r2, 1.0))
    rho = np.zeros(N)
    for t in range(T.nt):
        sel = mask & ((types[:, None] if slip == "rho_centre_type" else types[None, :]) == t)
        rho += np.where(sel, np_value(T.rho[t], T.r0, T.idr, r), 0.0).sum(1)
    df = np.zeros(N)
    Femb = np.zeros(N)
    for t in range(T.nt):
        s = types == t
        df[s] = np_deriv(T.emb[t], T.p0, T.idp, rho[s])
        Femb[s] = np_value(T.emb[t], T.p0, T.idp, rho[s])
    rhop = np.zeros((T.nt, N, N))
    for t in range(T.nt):
        rhop[t] = np_deriv(T.rho[t], T.r0, T.idr, r)
    ti, tj = types[:, None] * np.ones((1, N), int), np.ones((N, 1), int) * types[None, :]
    phip = np.zeros((N, N))
    phiv = np.zeros((N, N))
    for a in range(T.nt):
        for b in range(T.nt):
            sel = (ti == a) & (tj == b)
            phip = np.where(sel, np_deriv(T.phi[a][b], T.r0, T.idr, r), phip)
            phiv = np.where(sel, np_value(T.phi[a][b], T.r0, T.idr, r), phiv)
    rhojp = np.take_along_axis(rhop.transpose(1, 2, 0), tj[:, :, None], 2)[:, :, 0]
    rhoip = np.take_along_axis(rhop.transpose(1, 2, 0), ti[:, :, None], 2)[:, :, 0]
    dfi, dfj = (df[None, :], df[:, None]) if slip == "swap_df" else (df[:, None], df[None, :])
    if slip == "swap_rho_types":
        rhojp, rhoip = rhoip, rhojp
    psip = dfi * rhojp + dfj * rhoip + (0.0 if slip == "drop_phi" else phip)
    fpair = np.where(mask, (psip if slip == "sign" else -psip) / r, 0.0)
    F = (d * fpair[:, :, None]).sum(1)
    E = Femb.sum() + (1.0 if slip == "energy_double" else 0.5) * np.where(mask, phiv, 0.0).sum()
    return rho, df, F, E


def golden_fs_cases():
    """Closed-form Finnis-Sinclair lattice energies (tests/golden/fs_bcc_lattice.json)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fs_bcc_lattice.json")) as f:
        return json.load(f)["cases"]


def strained_lattice(a, strain, cells=(6, 6, 6)):
    """Perfect single-type BCC lattice of constant a, every ext position and the
    box scaled by (1 + strain) per axis, with Finnis-Sinclair tables."""
    spec = mmd.lattice_spec(cells, ghost_cells=3, a=a, n_types=1, sigma=0.0, temperature=0.0, seed=1)
    dom, box, st, itl = mmd.lattice_init(spec)
    s = 1.0 + np.asarray(strain, float)
    st.x *= s[0]
    st.y *= s[1]
    st.z *= s[2]
    st3 = mmd.stencil_build(a, RC + SKIN, 0)
    return dict(dom=dom, box=box * s, st=st, itl=itl, tab=mmd.fs_tables(rc=RC), st3=st3,
                even=mmd.stencil_linear(dom, st3), odd=mmd.stencil_linear(dom, -st3), masses=mmd.DEFAULT_MASSES)


def write_fs_setfl(path, n_types=2, jitter=0.1, seed=3, nr=3000, dr=0.002, nrho=3000, drho=0.01):
    """A DYNAMO setfl file of the Finnis-Sinclair Fe functions (embedding
    strength per type and pair term per type pair scaled by 1 +- jitter),
    values from the analytic forms; r * phi is tabulated, as the format asks."""
    d, a_fs, beta, c = 3.569745, 1.828905, 1.8, 3.40
    c0, c1, c2 = 1.2371147, -0.3592185, -0.0385607
    rng = np.random.default_rng(seed)
    amp = 1.0 + jitter * rng.uniform(-1, 1, n_types)
    spair = 1.0 + jitter * rng.uniform(-1, 1, (n_types, n_types))
    r = np.arange(nr) * dr
    rho = np.arange(nrho) * drho
    q = r - d
    dens = np.where(q < 0, q * q * (1 + beta * q / d), 0.0)
    p = r - c
    pair = np.where(p < 0, p * p * (c0 + c1 * r + c2 * r * r), 0.0)
    lines = ["Finnis-Sinclair Fe (Phil. Mag. A 50 (1984) 45)", "test fixture", "scaled per type",
             f"{n_types} " + " ".join(f"E{t}" for t in range(n_types)), f"{nrho} {drho!r} {nr} {dr!r} {RC!r}"]

    def emit(v):
        for k in range(0, len(v), 5):
            lines.append(" ".join("%.17g" % x for x in v[k:k + 5]))

    for t in range(n_types):
        lines.append(f"{26 + t} {mmd.DEFAULT_MASSES[t]!r} 2.855 bcc")
        emit(-a_fs * amp[t] * np.sqrt(rho))
        emit(dens)
    for i in range(n_types):
        for j in range(i + 1):
            emit(r * pair * 0.5 * (spair[i, j] + spair[j, i]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def ghost_sources(dom):
    """Reference-order index of each ext record's periodic source (-1 for local records)."""
    ex, ey, ez = dom.ext
    gx, gy, gz = dom.ghost
    bx, by, bz = dom.box
    z, y, x = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
    sx, sy, sz = (x - gx) % bx + gx, (y - gy) % by + gy, (z - gz) % bz + gz
    src = ((sz * ey + sy) * ex + sx).ravel().astype(np.int64)
    idx = np.arange(src.size)
    return np.where(src == idx, -1, src)
