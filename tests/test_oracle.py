"""Pins the CPU oracle (oracle/eam_oracle.c) against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures for this path, and the
spline arithmetic lives in hip-potential v0.2.0, absent from the reference
tree.  So the oracle is pinned here by closed forms, symmetry, brute force and
finite differences -- each chosen so that a plausible slip (a dropped term, a
wrong sign, the centre's type used instead of the neighbour's, a transposed
df_i/df_j) fails one of them.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

A, RC = W.A, W.RC


# ------------------------------------------------------------------ splines

def test_spline_knots_exact_and_c1():
    tab = mmd.synth_tables()
    T = O.Tables(tab, mmd.DEFAULT_MASSES)
    r = tab["r0"] + tab["dr"] * np.arange(0, tab["n_r"])
    vals = np.array([T.value("rho", 0, x) for x in r[10:200]])
    # knots are hit up to the rounding of r0 + i*dr (p may land at m-1, frac 1)
    assert np.allclose(vals, tab["rho_r"][0][10:200], rtol=4e-15, atol=0)
    # C1 across knots: derivative from the left == from the right
    for m in (50, 777, 3000):
        x = r[m]
        eps = 1e-9 * tab["dr"]
        dl, dr_ = T.deriv("phi", 0, x - eps), T.deriv("phi", 0, x + eps)
        assert abs(dl - dr_) < 1e-8 * abs(dl)


def test_spline_reproduces_cubic_in_interior():
    # the 5-point slope estimate is exact for polynomials up to degree 4, so an
    # interior segment reproduces a cubic exactly (up to rounding)
    x0, dx, n = 1.0, 0.01, 200
    x = x0 + dx * np.arange(n)
    f = 0.3 * x ** 3 - 2.0 * x ** 2 + x - 5.0
    c = O.spline_coeffs(f, dx)
    sp = O.OrcSpline(n, x0, dx, 1.0 / dx, c.ctypes.data_as(O._d))
    for xq in (1.234, 1.5555, 2.71):
        v = O.lib().orc_spline_value(sp, xq)
        d = O.lib().orc_spline_deriv(sp, xq)
        assert abs(v - (0.3 * xq ** 3 - 2 * xq ** 2 + xq - 5)) < 1e-12
        assert abs(d - (0.9 * xq ** 2 - 4 * xq + 1)) < 1e-10


def test_spline_matches_closed_forms():
    tab = mmd.synth_tables()
    T = O.Tables(tab, mmd.DEFAULT_MASSES)
    rc = RC
    for r in (1.6, 2.4725, 2.855, 4.03, 4.94, 5.5):
        rho = (rc - r) ** 2 * np.exp(-r)
        drho = (-2 * (rc - r) - (rc - r) ** 2) * np.exp(-r)
        phi = 1.2 * (rc - r) ** 2 * (1 + 0.3 / r)
        dphi = 1.2 * (-2 * (rc - r) * (1 + 0.3 / r) - (rc - r) ** 2 * 0.3 / r ** 2)
        assert abs(T.value("rho", 0, r) - rho) < 1e-11 * max(abs(rho), 1e-3)
        assert abs(T.deriv("rho", 0, r) - drho) < 1e-7 * max(abs(drho), 1e-3)
        assert abs(T.value("phi", 0, r) - phi) < 1e-11 * max(abs(phi), 1e-3)
        assert abs(T.deriv("phi", 0, r) - dphi) < 1e-7 * max(abs(dphi), 1e-3)
    for rho in (5.0, 9.9035, 18.4):
        F = -np.sqrt(rho) + 0.01 * rho ** 2
        dF = -0.5 / np.sqrt(rho) + 0.02 * rho
        assert abs(T.value("emb", 0, rho) - F) < 1e-10
        assert abs(T.deriv("emb", 0, rho) - dF) < 1e-7


def test_spline_numpy_restatement_bit_exact():
    tab = mmd.synth_tables()
    T = O.Tables(tab, mmd.DEFAULT_MASSES)
    sp = W.np_spline(tab["phi_r"][0][0], tab["dr"])
    xs = np.linspace(0.3, 5.7, 997)
    v_np = W.np_value(sp, tab["r0"], 1.0 / tab["dr"], xs)
    d_np = W.np_deriv(sp, tab["r0"], 1.0 / tab["dr"], xs)
    v_o = np.array([T.value("phi", 0, x) for x in xs])
    d_o = np.array([T.deriv("phi", 0, x) for x in xs])
    assert np.array_equal(v_np, v_o)
    assert np.array_equal(d_np, d_o)


# ------------------------------------------------------------------ stencil

def _brute_stencil(a, radius, parity):
    # explicit Cartesian positions: site (X, Y, Z) of parity s = X mod 2 sits at
    # (X a/2, (Y + s/2) a, (Z + s/2) a); the centre is (parity, 0, 0)
    out = set()
    R = int(radius / a) + 2
    centre = np.array([parity * a / 2, parity * a / 2, parity * a / 2])
    for dX in range(-2 * R, 2 * R + 1):
        for dY in range(-R, R + 1):
            for dZ in range(-R, R + 1):
                X = parity + dX
                s = X % 2
                q = np.array([X * a / 2, (dY + s / 2) * a, (dZ + s / 2) * a])
                if (dX, dY, dZ) != (0, 0, 0) and np.linalg.norm(q - centre) < radius:
                    out.add((dX, dY, dZ))
    return out


@pytest.mark.parametrize("radius,count", [(5.6, 58), (5.9, 64), (6.6, 112), (7.5, 168)])
def test_stencil_sets_and_counts(radius, count):
    for parity in (0, 1):
        o = O.stencil3(A, radius, parity)
        p = mmd.stencil_build(A, radius, parity)
        assert len(o) == count
        assert np.array_equal(o, p)  # product generator == oracle generator, entry by entry
        assert set(map(tuple, o)) == _brute_stencil(A, radius, parity)
    # entry k has the same ideal distance for both parities (negation)
    assert np.array_equal(O.stencil3(A, radius, 1), -O.stencil3(A, radius, 0))


def test_stencil_linear_offsets_match_reference_formula():
    c = W.case(cells=(4, 5, 6))
    d = c["dom"]
    st3 = c["st3"]
    ref = (st3[:, 2] * d.ext[1] + st3[:, 1]) * d.ext[0] + st3[:, 0]
    assert np.array_equal(c["even"], ref)
    with pytest.raises(mmd.MmdError):
        mmd.stencil_linear(d, np.array([[0, 0, d.ghost[2] + 1]], np.int32))


# ------------------------------------------------------------- EAM passes

def _run_oracle(c, want_e=False):
    st, T = c["st"], O.Tables(c["tab"], c["masses"])
    ee = O.rho_df(c["dom"], c["even"], c["odd"], T, RC, st, want_e=want_e)
    O.ghost_df(c["dom"], st)
    ep = O.force(c["dom"], c["even"], c["odd"], T, RC, st, want_e=want_e)
    return T, ee + ep


def test_perfect_lattice_closed_form_and_zero_force():
    c = W.case(cells=(5, 5, 5), sigma=0.0, temperature=0.0)
    c["tab"] = mmd.synth_tables(rc=RC)  # phi = 1.2 (rc-r)^2 (1 + 0.3/r), rho = (rc-r)^2 e^-r, F = -sqrt + 0.01 rho^2
    _run_oracle(c)
    st = c["st"]
    li = st.local_index()
    shells = [(8, np.sqrt(3) / 2 * A), (6, A), (12, np.sqrt(2) * A), (24, np.sqrt(11) / 2 * A), (8, np.sqrt(3) * A)]
    rho_cf = sum(n * (RC - r) ** 2 * np.exp(-r) for n, r in shells)
    assert np.allclose(st.rho[li], rho_cf, rtol=1e-12, atol=0)
    df_cf = -0.5 / np.sqrt(rho_cf) + 0.02 * rho_cf
    assert np.allclose(st.df[li], df_cf, rtol=1e-7)
    assert np.abs(np.stack([st.fx[li], st.fy[li], st.fz[li]])).max() < 1e-12


def test_finnis_sinclair_fe_cohesive_energy():
    """The Finnis-Sinclair Fe tables reproduce the published fit (Phil. Mag. A
    50 (1984) 45): cohesive energy 4.28 eV at a = 2.8665 A, a closed-form
    density over the two shells inside d, zero force on the perfect lattice,
    and a minimum of the energy per atom at that lattice constant."""
    d, beta = 3.569745, 1.8
    e = {}
    for a in (2.8465, 2.8665, 2.8865):
        spec = mmd.lattice_spec((5, 5, 5), ghost_cells=3, a=a, n_types=1, sigma=0.0, temperature=0.0, seed=1)
        dom, box, st, _ = mmd.lattice_init(spec)
        st3 = mmd.stencil_build(a, RC + W.SKIN, 0)
        c = dict(dom=dom, st=st, tab=mmd.fs_tables(rc=RC), masses=mmd.DEFAULT_MASSES,
                 even=mmd.stencil_linear(dom, st3), odd=mmd.stencil_linear(dom, -st3))
        _, E = _run_oracle(c, want_e=True)
        li = st.local_index()
        e[a] = E / li.size
        if a == 2.8665:
            q = np.array([np.sqrt(3) / 2 * a, a]) - d
            rho_cf = (np.array([8, 6]) * q * q * (1 + beta * q / d)).sum()
            assert np.allclose(st.rho[li], rho_cf, rtol=1e-12, atol=0)
            assert np.abs(np.stack([st.fx[li], st.fy[li], st.fz[li]])).max() < 1e-11
    assert abs(e[2.8665] + 4.28) < 2e-3
    assert e[2.8665] < e[2.8465] and e[2.8665] < e[2.8865]


@pytest.mark.parametrize("n_types,vac", [(1, 0.0), (3, 0.05)])
def test_oracle_matches_bruteforce(n_types, vac):
    c = W.case(cells=(4, 4, 4), n_types=n_types, sigma=0.05, temperature=800.0, vacancy_frac=vac, seed=11,
               distinct_rho=True)
    _, E = _run_oracle(c, want_e=True)
    st = c["st"]
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    pos = np.stack([st.x[occ], st.y[occ], st.z[occ]], 1)
    rho, df, F, Ebf = W.brute_force(pos, st.type[occ], c["box"], W.NpTables(c["tab"]), RC)
    assert np.allclose(st.rho[occ], rho, rtol=1e-12, atol=0)
    assert np.allclose(st.df[occ], df, rtol=1e-12, atol=1e-15)
    Fo = np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1)
    assert np.abs(Fo - F).max() < 1e-11 * np.abs(F).max()
    assert abs(E - Ebf) < 1e-11 * abs(Ebf)
    if vac:
        assert (st.type[li] < 0).sum() > 0


@pytest.mark.parametrize("slip", ["rho_centre_type", "swap_df", "swap_rho_types", "sign", "drop_phi",
                                  "energy_double"])
def test_bruteforce_comparison_catches_planted_slips(slip):
    """The brute-force comparison above has teeth: each plausible mistake,
    planted in the numpy restatement, breaks the agreement with the oracle
    (3-type alloy with vacancies, so type and df slips change the result)."""
    c = W.case(cells=(4, 4, 4), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.05, seed=11,
               distinct_rho=True)
    _, E = _run_oracle(c, want_e=True)
    st = c["st"]
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    pos = np.stack([st.x[occ], st.y[occ], st.z[occ]], 1)
    rho, df, F, Ebf = W.brute_force(pos, st.type[occ], c["box"], W.NpTables(c["tab"]), RC, slip=slip)
    Fo = np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1)
    agree = (np.allclose(st.rho[occ], rho, rtol=1e-12, atol=0) and np.abs(Fo - F).max() < 1e-11 * np.abs(F).max()
             and abs(E - Ebf) < 1e-11 * abs(Ebf))
    assert not agree


def test_force_is_minus_gradient_and_momentum_conserved():
    c = W.case(cells=(4, 4, 4), n_types=3, sigma=0.08, temperature=0.0, vacancy_frac=0.03, seed=5)
    T, _ = _run_oracle(c)
    st, dom, box = c["st"], c["dom"], c["box"]
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    F = np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1)
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).sum()  # symmetric phi: Newton's third law

    def energy(s):
        O.ghost_xt(dom, box, s)
        e1 = O.rho_df(dom, c["even"], c["odd"], T, RC, s, want_e=True)
        O.ghost_df(dom, s)
        return e1 + O.force(dom, c["even"], c["odd"], T, RC, s, want_e=True)

    h = 1e-5
    for k in (3, 17, 40):
        i = occ[k]
        for ax, fld in enumerate(("x", "y", "z")):
            sp, sm = st.copy(), st.copy()
            getattr(sp, fld)[i] += h
            getattr(sm, fld)[i] -= h
            g = (energy(sp) - energy(sm)) / (2 * h)
            assert abs(-g - F[k, ax]) < 2e-6 * max(1.0, abs(F[k, ax]))


def test_translation_invariance():
    c = W.case(cells=(4, 4, 4), n_types=3, sigma=0.05, seed=9)
    _run_oracle(c)
    st = c["st"]
    li = st.local_index()
    F0 = np.stack([st.fx[li], st.fy[li], st.fz[li]], 1)
    s2 = st.copy()
    s2.x += 0.37
    s2.y -= 1.1
    s2.z += 2.0
    T = O.Tables(c["tab"], c["masses"])
    O.rho_df(c["dom"], c["even"], c["odd"], T, RC, s2)
    O.ghost_df(c["dom"], s2)
    O.force(c["dom"], c["even"], c["odd"], T, RC, s2)
    F1 = np.stack([s2.fx[li], s2.fy[li], s2.fz[li]], 1)
    assert np.abs(F1 - F0).max() < 1e-10


def test_ghost_refresh_periodic_images():
    c = W.case(cells=(4, 5, 6), sigma=0.05)
    st, dom, box = c["st"], c["dom"], c["box"]
    s2 = st.copy()
    s2.x[:] = 0
    s2.type[:] = -1
    li = st.local_index()
    s2.x[li], s2.y[li], s2.z[li], s2.type[li] = st.x[li], st.y[li], st.z[li], st.type[li]
    O.ghost_xt(dom, box, s2)
    assert np.array_equal(s2.type, st.type)
    assert np.allclose(s2.x, st.x, rtol=0, atol=1e-12)
    assert np.allclose(s2.z, st.z, rtol=0, atol=1e-12)


def test_velocity_verlet_harmonic_dimer():
    # two atoms in an otherwise empty lattice box; phi harmonic, rho = 0, F = 0:
    # relative coordinate oscillates with omega = sqrt(2k/m)
    c = W.case(cells=(5, 5, 5), sigma=0.0, temperature=0.0)
    st, dom, box = c["st"], c["dom"], c["box"]
    st.type[:] = -1
    d = dom
    i0 = ((d.ghost[2] + 2) * d.ext[1] + d.ghost[1] + 2) * d.ext[0] + d.ghost[0] + 4
    i1 = i0 + 1  # body-centre neighbour at sqrt(3)/2 a
    st.type[i0] = st.type[i1] = 0
    r0 = np.sqrt(3) / 2 * A
    k = 2.0
    tab = dict(c["tab"])
    r = tab["r0"] + tab["dr"] * np.arange(tab["n_r"])
    tab["phi_r"] = (0.5 * k * (r - r0) ** 2 - 0.5 * k * (RC - r0) ** 2)[None, None, :].copy()
    tab["rho_r"] = np.zeros_like(tab["rho_r"])
    tab["emb"] = np.zeros_like(tab["emb"])
    T = O.Tables(tab, c["masses"])
    amp = 0.05
    u = np.array([1.0, 1.0, 1.0]) / np.sqrt(3)
    st.x[i1] += amp * u[0]
    st.y[i1] += amp * u[1]
    st.z[i1] += amp * u[2]
    st.vx[:] = st.vy[:] = st.vz[:] = 0.0
    O.ghost_xt(dom, box, st)
    O.rho_df(dom, c["even"], c["odd"], T, RC, st)
    O.ghost_df(dom, st)
    O.force(dom, c["even"], c["odd"], T, RC, st)
    m = c["masses"][0]
    omega = np.sqrt(2 * k / m * 9.64853321e-3)
    dt, n = 0.5, 400
    O.md_run(dom, c["even"], c["odd"], T, RC, box, dt, n, st, energies=False)
    sep = np.sqrt((st.x[i1] - st.x[i0]) ** 2 + (st.y[i1] - st.y[i0]) ** 2 + (st.z[i1] - st.z[i0]) ** 2)
    expect = r0 + amp * np.cos(omega * dt * n)
    assert abs(sep - expect) < 2e-4 * amp / 0.05 + (omega * dt) ** 2 * amp


def test_md_energy_conservation():
    c = W.case(cells=(4, 4, 4), n_types=3, sigma=0.03, temperature=300.0, seed=3)
    T, _ = _run_oracle(c)
    en = O.md_run(c["dom"], c["even"], c["odd"], T, RC, c["box"], 0.5, 40, c["st"])
    tot = en[:, 0] + en[:, 1]
    n_atoms = (c["st"].type[c["st"].local_index()] >= 0).sum()
    assert np.ptp(tot) / n_atoms < 2e-4
    assert np.ptp(en[:, 1]) > 10 * np.ptp(tot)  # kinetic energy does move


# ------------------------------------------------------------ interstitials

def test_anchor_is_nearest_site():
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = rng.uniform(-3, 20, 3)
        X, Y, Z = O.anchor(A, p)
        s = X & 1
        site = np.array([X * A / 2, (Y + s / 2) * A, (Z + s / 2) * A])
        best = np.inf
        for dX in range(-3, 4):
            for dY in range(-2, 3):
                for dZ in range(-2, 3):
                    X2 = X + dX
                    s2 = X2 & 1
                    q = np.array([X2 * A / 2, (Y + dY + s2 / 2) * A, (Z + dZ + s2 / 2) * A])
                    best = min(best, np.linalg.norm(q - p))
        assert np.linalg.norm(site - p) <= best + 1e-12


@pytest.mark.parametrize("n_types", [1, 3])
def test_oracle_interstitials_match_bruteforce(n_types):
    c = W.case(cells=(6, 6, 6), n_types=n_types, sigma=0.04, temperature=600.0, vacancy_frac=0.01,
               interstitial_frac=0.01, seed=21, distinct_rho=True)
    st, itl, dom, box = c["st"], c["itl"], c["dom"], c["box"]
    assert itl.n >= 3
    T = O.Tables(c["tab"], c["masses"])
    ee = O.rho_df_itl(dom, c["even"], c["odd"], c["wide3"], A, box, T, RC, st, itl, want_e=True)
    O.ghost_df(dom, st)
    ep = O.force_itl(dom, c["even"], c["odd"], c["wide3"], A, box, T, RC, st, itl, want_e=True)
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    pos = np.concatenate([np.stack([st.x[occ], st.y[occ], st.z[occ]], 1), np.stack([itl.x, itl.y, itl.z], 1)])
    types = np.concatenate([st.type[occ], itl.type])
    rho, df, F, E = W.brute_force(pos, types, box, W.NpTables(c["tab"]), RC)
    n = len(occ)
    assert np.allclose(st.rho[occ], rho[:n], rtol=1e-12)
    assert np.allclose(itl.rho, rho[n:], rtol=1e-12)
    assert np.allclose(itl.df, df[n:], rtol=1e-11)
    Fl = np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1)
    Fi = np.stack([itl.fx, itl.fy, itl.fz], 1)
    scale = np.abs(F).max()
    assert np.abs(Fl - F[:n]).max() < 1e-11 * scale
    assert np.abs(Fi - F[n:]).max() < 1e-11 * scale
    assert abs(ee + ep - E) < 1e-11 * abs(E)


def test_oracle_interstitial_md_conserves_energy():
    c = W.case(cells=(6, 6, 6), n_types=3, sigma=0.04, temperature=600.0, vacancy_frac=0.01,
               interstitial_frac=0.01, seed=21)
    st, itl, dom, box = c["st"], c["itl"], c["dom"], c["box"]
    T = O.Tables(c["tab"], c["masses"])
    O.rho_df_itl(dom, c["even"], c["odd"], c["wide3"], A, box, T, RC, st, itl)
    O.ghost_df(dom, st)
    O.force_itl(dom, c["even"], c["odd"], c["wide3"], A, box, T, RC, st, itl)
    en = O.md_run_itl(dom, c["even"], c["odd"], c["wide3"], A, T, RC, box, 0.5, 30, st, itl)
    tot = en.sum(1)
    n_at = (st.type[st.local_index()] >= 0).sum() + itl.n
    assert np.ptp(tot) / n_at < 5e-4
    assert np.ptp(en[:, 1]) > 10 * np.ptp(tot)
