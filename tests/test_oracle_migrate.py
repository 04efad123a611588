"""Interstitial migration in the oracle (orc_migrate): known answers and
invariants that pin its rules.

Rules under test: a lattice atom outside its own Wigner-Seitz cell joins the
end of the interstitial list, in local site order, and leaves a vacancy; a list
entry whose Wigner-Seitz site is vacant moves into it (periodic image removed),
the lowest index winning a contested site; the rest of the list keeps its
order; every field travels with its atom.  The nearest-site rule itself is
checked against a brute-force search over neighbouring sites.
"""
import numpy as np
import pytest

import mmd
import oracle as O

A = 2.855
GHOST = 3
FIELDS = mmd.HostState.FIELDS


def site_pos(X, Y, Z):
    s = X & 1
    return np.array([X * A / 2, (Y + s / 2) * A, (Z + s / 2) * A])


def lattice(cells=(4, 4, 4), sigma=0.0, temperature=300.0, vacancy_frac=0.0, interstitial_frac=0.0, seed=5,
            cap_extra=16):
    spec = mmd.lattice_spec(cells, ghost_cells=GHOST, a=A, n_types=3, sigma=sigma, temperature=temperature,
                            vacancy_frac=vacancy_frac, interstitial_frac=interstitial_frac, seed=seed)
    cap = int(interstitial_frac * 2 * np.prod(cells)) + cap_extra
    dom, box, st, itl = mmd.lattice_init(spec, itl_cap=cap)
    # distinct values in every carried field, so a mix-up of atoms or fields shows
    rng = np.random.default_rng(seed)
    li = st.local_index()
    for f in ("fx", "fy", "fz", "rho", "df"):
        getattr(st, f)[li] = rng.normal(size=li.size)
        getattr(itl, f)[:] = rng.normal(size=itl.n)
    O.ghost_xt(dom, box, st)
    O.ghost_df(dom, st)
    return dom, box, st, itl, cap


def ext_of(st, xl, yl, zl):
    bx, by = st.dom.box[0], st.dom.box[1]
    return st.local_index()[(zl * by + yl) * bx + xl]


def with_list(itl, entries):
    """A list holding the given (position, type) entries plus distinct fields."""
    o = mmd.Interstitials(len(entries))
    rng = np.random.default_rng(len(entries))
    for k, (p, t) in enumerate(entries):
        o.x[k], o.y[k], o.z[k] = p
        o.type[k] = t
    for f in ("vx", "vy", "vz", "fx", "fy", "fz", "rho", "df"):
        getattr(o, f)[:] = rng.normal(size=len(entries))
    return o


def atom(st, i):
    return tuple(getattr(st, f)[i] for f in FIELDS) + (st.type[i],)


def entry(itl, k):
    return tuple(getattr(itl, f)[k] for f in FIELDS) + (itl.type[k],)


# ------------------------------------------------------------ nearest site

def test_nearest_site_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        p = rng.uniform(-5, 25, 3)
        X, Y, Z = O.anchor(A, p)
        d_own = np.linalg.norm(site_pos(X, Y, Z) - p)
        best = np.inf
        for dX in range(-3, 4):
            for dY in range(-2, 3):
                for dZ in range(-2, 3):
                    if (dX, dY, dZ) != (0, 0, 0):
                        best = min(best, np.linalg.norm(site_pos(X + dX, Y + dY, Z + dZ) - p))
        assert d_own <= best + 1e-12


def test_nearest_site_ties_and_sites():
    # every ideal site is its own nearest site
    for X in range(-4, 5):
        for Y in (-1, 0, 3):
            for Z in (-2, 0, 1):
                assert O.anchor(A, site_pos(X, Y, Z)) == (X, Y, Z)
    # the corner/body midpoint (a/4)(1,1,1) ties: the corner sub-lattice wins
    assert O.anchor(A, np.array([0.25, 0.25, 0.25]) * 2.0) == (0, 0, 0)
    # a tetrahedral site (a/2, a/4, 0) is equidistant from four sites; the
    # corner candidate rounds halves up
    assert O.anchor(2.0, np.array([1.0, 0.5, 0.0])) == (2, 0, 0)


def test_ideal_lattice_positions():
    dom, box, st, itl, cap = lattice()
    bx, by, bz = dom.box
    for (xl, yl, zl) in ((0, 0, 0), (1, 0, 0), (5, 2, 3), (bx - 1, by - 1, bz - 1)):
        i = ext_of(st, xl, yl, zl)
        assert np.array_equal([st.x[i], st.y[i], st.z[i]], site_pos(xl, yl, zl))


# ------------------------------------------------------------ known answers

def test_no_defects_no_change():
    dom, box, st, itl, cap = lattice(cells=(5, 4, 6), sigma=0.08, interstitial_frac=0.01)
    assert itl.n >= 1
    st0, itl0 = st.copy(), itl.take(itl.n)
    new, out, fill = O.migrate(dom, A, box, st, itl, cap)
    assert (out, fill) == (0, 0)
    for f in FIELDS + ("type",):
        assert np.array_equal(getattr(st, f), getattr(st0, f))
        assert np.array_equal(getattr(new, f), getattr(itl0, f))


def test_hop_into_neighbouring_vacancy():
    dom, box, st, itl, cap = lattice()
    v, b = ext_of(st, 4, 2, 2), ext_of(st, 5, 2, 2)  # corner site and a nearest-neighbour body site
    st.type[v] = -1
    pv, pb = site_pos(4, 2, 2), site_pos(5, 2, 2)
    st.x[b], st.y[b], st.z[b] = pb + 0.8 * (pv - pb)
    O.ghost_xt(dom, box, st)
    moved = atom(st, b)
    new, out, fill = O.migrate(dom, A, box, st, itl.take(0), cap)
    assert (out, fill, new.n) == (1, 1, 0)
    assert atom(st, v) == moved
    assert st.type[b] == -1
    for f in ("vx", "vy", "vz", "fx", "fy", "fz", "rho", "df"):
        assert getattr(st, f)[b] == 0.0
    assert st.x[b] == moved[0]  # a new vacancy keeps its last position


def test_leaver_joins_list_in_site_order():
    dom, box, st, itl, cap = lattice()
    old = with_list(itl, [(site_pos(3, 1, 1) + [0.1, 0.7, 0.0], 1)])
    # three atoms pushed into occupied neighbours' cells; site order is z, y, doubled x
    moves = [((6, 3, 3), (7, 3, 3)), ((2, 0, 0), (3, 0, 0)), ((1, 2, 0), (0, 2, 0))]
    expect = []
    for (s_from, s_to) in sorted(moves, key=lambda m: (m[0][2], m[0][1], m[0][0])):
        i = ext_of(st, *s_from)
        st.x[i], st.y[i], st.z[i] = site_pos(*s_to) + [0.05, -0.02, 0.03]
        expect.append(i)
    expect_atoms = [atom(st, i) for i in expect]
    O.ghost_xt(dom, box, st)
    new, out, fill = O.migrate(dom, A, box, st, old, cap)
    assert (out, fill, new.n) == (3, 0, 4)
    assert entry(new, 0) == entry(old, 0)
    for k, a in enumerate(expect_atoms):
        assert entry(new, 1 + k) == a
    for i in expect:
        assert st.type[i] == -1


def test_lowest_index_wins_contested_vacancy():
    dom, box, st, itl, cap = lattice()
    v = ext_of(st, 2, 2, 2)
    st.type[v] = -1
    O.ghost_xt(dom, box, st)
    pv = site_pos(2, 2, 2)
    old = with_list(itl, [(site_pos(6, 1, 1) + 0.6, 2), (pv + [0.3, 0.1, 0.0], 1), (pv - [0.2, 0.2, 0.2], 0),
                          (site_pos(0, 3, 3) + [0.0, 0.6, 0.0], 2)])
    new, out, fill = O.migrate(dom, A, box, st, old, cap)
    assert (out, fill, new.n) == (0, 1, 3)
    assert st.type[v] == 1 and st.vx[v] == old.vx[1] and st.rho[v] == old.rho[1]
    assert [entry(new, k) for k in range(3)] == [entry(old, k) for k in (0, 2, 3)]


def test_fill_removes_periodic_image():
    dom, box, st, itl, cap = lattice(cells=(4, 5, 3))
    v = ext_of(st, 0, 0, 0)
    st.type[v] = -1
    O.ghost_xt(dom, box, st)
    p = site_pos(0, 0, 0) + np.array([0.2, -0.3, 0.1]) + np.array([box[0], -box[1], 2 * box[2]])
    old = with_list(itl, [(p, 2)])
    new, out, fill = O.migrate(dom, A, box, st, old, cap)
    assert (out, fill, new.n) == (0, 1, 0)
    # one rounding each (image times box length is exact for images 1 and 2)
    assert (st.x[v], st.y[v], st.z[v]) == (p[0] - box[0], p[1] + box[1], p[2] - 2 * box[2])
    assert np.allclose([st.x[v], st.y[v], st.z[v]], site_pos(0, 0, 0) + [0.2, -0.3, 0.1], atol=1e-12)
    # the ghost images of the filled site follow it
    gi = (st.type == 2) & (np.abs(st.x - st.x[v]) + np.abs(st.y - st.y[v]) + np.abs(st.z - st.z[v]) > 1.0)
    assert gi.sum() >= 1


def test_over_capacity_changes_nothing():
    dom, box, st, itl, cap = lattice(cap_extra=2)
    old = with_list(itl, [(site_pos(3, 1, 1) + [0.1, 0.7, 0.0], 1), (site_pos(5, 1, 1) + [0.1, 0.7, 0.0], 1)])
    for s_from, s_to in (((6, 3, 3), (7, 3, 3)), ((2, 0, 0), (3, 0, 0))):
        i = ext_of(st, *s_from)
        st.x[i], st.y[i], st.z[i] = site_pos(*s_to) + 0.05
    O.ghost_xt(dom, box, st)
    st0 = st.copy()
    with pytest.raises(OverflowError):
        O.migrate(dom, A, box, st, old, 3)
    for f in FIELDS + ("type",):
        assert np.array_equal(getattr(st, f), getattr(st0, f))
    new, out, fill = O.migrate(dom, A, box, st, old, 4)
    assert (out, fill, new.n) == (2, 0, 4)


# ------------------------------------------------------------ invariants

def wrapped_site(dom, p):
    X, Y, Z = O.anchor(A, p)
    return X % dom.box[0], Y % dom.box[1], Z % dom.box[2]


@pytest.mark.parametrize("seed", [1, 2])
def test_invariants_on_disordered_state(seed):
    dom, box, st, itl, cap = lattice(cells=(8, 7, 6), sigma=0.6, temperature=900.0, vacancy_frac=0.1,
                                     interstitial_frac=0.02, seed=seed, cap_extra=400)
    li = st.local_index()
    before = {a[3:6]: a for a in [atom(st, i) for i in li if st.type[i] >= 0] +
              [entry(itl, k) for k in range(itl.n)]}
    n0 = len(before)
    new, out, fill = O.migrate(dom, A, box, st, itl, cap)
    assert out > 10 and fill > 10, (out, fill)
    after = [atom(st, i) for i in li if st.type[i] >= 0] + [entry(new, k) for k in range(new.n)]
    assert len(after) == n0 and new.n - itl.n == out - fill
    # every atom survives once with all its fields; positions only by whole periods
    for a in after:
        b = before.pop(a[3:6])
        assert a[3:] == b[3:]
        shift = (np.array(a[:3]) - np.array(b[:3])) / box
        assert np.allclose(shift, np.round(shift), atol=1e-12)
    assert not before
    # lattice atoms sit in their own cells; entries' cells are occupied
    bx, by = dom.box[0], dom.box[1]
    for n, i in enumerate(li):
        if st.type[i] >= 0:
            xl, yl, zl = n % bx, (n // bx) % by, n // (bx * by)
            assert O.anchor(A, (st.x[i], st.y[i], st.z[i])) == (xl, yl, zl)
    for k in range(new.n):
        s = wrapped_site(dom, (new.x[k], new.y[k], new.z[k]))
        assert st.type[ext_of(st, *s)] >= 0
    # idempotent
    st1 = st.copy()
    new2, out2, fill2 = O.migrate(dom, A, box, st, new, cap)
    assert (out2, fill2, new2.n) == (0, 0, new.n)
    for f in FIELDS + ("type",):
        assert np.array_equal(getattr(st, f), getattr(st1, f))
