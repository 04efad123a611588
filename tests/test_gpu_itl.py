"""GPU parity of the interstitial (off-lattice atom) passes against the CPU
oracle, through the C ABI.

The oracle's interstitial passes are pinned against an O(N^2) brute force in
tests/test_oracle.py.  Tolerances as for the lattice path: FP64 results agree
to rounding (rho relative <= 1e-12, forces <= 1e-11 of the largest force);
integer bookkeeping (types, vacancy flags, interstitial count) bit-exact.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC

ITL_CASES = {
    "bcc8_1type": dict(cells=(8, 8, 8), n_types=1, sigma=0.04, temperature=600.0, vacancy_frac=0.01,
                       interstitial_frac=0.01, seed=21),
    "bcc10_3type": dict(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.005,
                        interstitial_frac=0.006, seed=22),
    "blobs_12": dict(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01,
                     interstitial_frac=0.004, n_blobs=2, blob_sigma=6.0, seed=23),
    # rho_t differs per type: each side of a lattice-interstitial pair takes the other's rho
    "bcc10_3type_distinct_rho": dict(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=1200.0,
                                     vacancy_frac=0.005, interstitial_frac=0.006, seed=24, distinct_rho=True),
    # one blob of sigma 8 A, as the large-box configuration places them, filled to the initial
    # conditions' packing limit (at most one interstitial per 3x3x3 cells around each):
    # 829 interstitials, up to 22 within 2 rc of one and 4 within rc
    "blob_packed_24": dict(cells=(24, 24, 24), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01,
                           interstitial_frac=0.03, n_blobs=1, blob_sigma=8.0, seed=25),
}


def test_packed_blob_case_is_dense():
    """The packed-blob case holds the densest interstitial neighbourhoods the
    initial conditions make (the bins and both pass sides see them)."""
    c = W.case(**ITL_CASES["blob_packed_24"])
    it, L = c["itl"], c["box"]
    P = np.stack([it.x, it.y, it.z], 1)
    d = P[:, None, :] - P[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    assert it.n > 800
    assert ((r2 < (2 * RC) ** 2).sum(1) - 1).max() >= 20
    assert ((r2 < RC ** 2).sum(1) - 1).max() >= 3


def _setup(c, prec=mmd.PREC_FP64):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.set_precision(prec)
    ctx.upload(c["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=1)
    ctx.itl_set(c["itl"], c["wide3"])
    return ctx


def _oracle(c):
    st, itl = c["st"].copy(), c["itl"].take(c["itl"].n)
    T = O.Tables(c["tab"], c["masses"])
    e1 = O.rho_df_itl(c["dom"], c["even"], c["odd"], c["wide3"], A, c["box"], T, RC, st, itl, want_e=True)
    O.ghost_df(c["dom"], st)
    e2 = O.force_itl(c["dom"], c["even"], c["odd"], c["wide3"], A, c["box"], T, RC, st, itl, want_e=True)
    return st, itl, T, e1 + e2


@pytest.mark.parametrize("name", list(ITL_CASES))
def test_interstitial_passes_match_oracle(name):
    c = W.case(**ITL_CASES[name])
    assert c["itl"].n >= 2
    ref, ritl, T, E = _oracle(c)
    with _setup(c) as ctx:
        ctx.rho_df(RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        out = c["st"].copy()
        ctx.download(out, mmd.F_RHO | mmd.F_DF | mmd.F_F | mmd.F_TYPE)
        gi = ctx.itl_get(c["itl"].n)
        ep, ek = ctx.energy()
    assert gi.n == c["itl"].n and np.array_equal(gi.type, c["itl"].type)
    assert np.array_equal(out.type, c["st"].type)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-12)
    assert np.allclose(out.df[occ], ref.df[occ], rtol=1e-12, atol=1e-15)
    F = np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1)
    Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
    scale = max(np.abs(Fr).max(), np.abs(np.stack([ritl.fx, ritl.fy, ritl.fz])).max())
    assert np.abs(F - Fr).max() <= 1e-11 * scale
    assert np.allclose(gi.rho, ritl.rho, rtol=1e-12)
    assert np.allclose(gi.df, ritl.df, rtol=1e-12)
    Fi = np.stack([gi.fx, gi.fy, gi.fz], 1)
    Fir = np.stack([ritl.fx, ritl.fy, ritl.fz], 1)
    assert np.abs(Fi - Fir).max() <= 1e-11 * scale
    assert abs(ep - E) <= 1e-12 * abs(E)
    # momentum: lattice + interstitial forces cancel (symmetric pair tables)
    tot = F.sum(0) + Fi.sum(0)
    assert np.abs(tot).max() < 1e-9 * np.abs(F).sum()


@pytest.mark.parametrize("name", ["bcc10_3type", "blob_packed_24"])
def test_interstitial_md_matches_oracle(name):
    c = W.case(**ITL_CASES[name])
    ref, ritl, T, _ = _oracle(c)
    nsteps = 8
    en_ref = O.md_run_itl(c["dom"], c["even"], c["odd"], c["wide3"], A, T, RC, c["box"], 1.0, nsteps, ref, ritl)
    with _setup(c) as ctx:
        ctx.md_prime()
        en = np.zeros((nsteps, 2))
        ctx.md_run(nsteps, en)
        out = c["st"].copy()
        ctx.download(out)
        gi = ctx.itl_get(c["itl"].n)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.abs(getattr(out, f)[occ] - getattr(ref, f)[occ]).max() < 1e-10, f
        assert np.abs(getattr(gi, f) - getattr(ritl, f)).max() < 1e-10, f
    assert np.allclose(en, en_ref, rtol=1e-12)


def test_interstitial_graph_equals_eager_and_energy_conserved():
    c = W.case(**ITL_CASES["blobs_12"])
    res = []
    for use_graph in (False, True):
        with _setup(c) as ctx:
            ctx.md_setup(c["box"], A, 0.5, RC, energy_every=1)
            ctx.md_prime()
            en = np.zeros((30, 2))
            ctx.md_run(15, en[:15], use_graph=use_graph)
            ctx.md_run(15, en[15:], use_graph=use_graph)
            o = c["st"].copy()
            ctx.download(o)
            res.append((o, ctx.itl_get(c["itl"].n), en))
    (a, ia, ea), (b, ib, eb) = res
    assert np.array_equal(a.x, b.x) and np.array_equal(ia.x, ib.x) and np.array_equal(ea, eb)
    tot = ea.sum(1)
    n_at = (a.type[a.local_index()] >= 0).sum() + ia.n
    assert np.ptp(tot) / n_at < 5e-4
    assert np.ptp(ea[:, 1]) > 10 * np.ptp(tot)


def test_interstitial_graph_without_prime_and_after_new_box():
    """The interstitial bins are sized by mmd_itl_set and mmd_md_setup, never
    inside a captured step: a graph-replayed run straight after set-up (no
    mmd_md_prime), and again on one context after a new box of another size,
    equals the eager run bit for bit."""
    cases = [W.case(**ITL_CASES["blobs_12"]), W.case(**ITL_CASES["bcc10_3type"])]
    outs = {}
    for use_graph in (False, True):
        ctx = mmd.Context(0)
        res = []
        for c in cases:
            ctx.set_domain(c["dom"])
            ctx.set_stencil(c["even"], c["odd"])
            ctx.set_tables(c["tab"], c["masses"])
            ctx.upload(c["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
            ctx.itl_set(c["itl"], c["wide3"])
            ctx.md_setup(c["box"], A, 0.5, RC, energy_every=2)
            en = np.zeros((6, 2))  # [nsteps][2]: rows of the recorded steps are written
            ctx.md_run(6, en, use_graph=use_graph)
            o = c["st"].copy()
            ctx.download(o)
            res.append((o, ctx.itl_get(c["itl"].n), en))
        ctx.close()
        outs[use_graph] = res
    for (a, ia, ea), (b, ib, eb) in zip(outs[False], outs[True]):
        for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
            assert np.array_equal(getattr(ia, f), getattr(ib, f)), f
        assert np.array_equal(ea, eb)
        assert np.isfinite(ea).all() and (ea[1::2] != 0).all()


def test_interstitials_mixed_precision_within_target():
    c = W.case(**ITL_CASES["bcc10_3type"])
    res = {}
    for prec in (mmd.PREC_FP64, mmd.PREC_MIXED):
        with _setup(c, prec) as ctx:
            ctx.rho_df(RC)
            ctx.ghost_refresh(mmd.F_DF)
            ctx.force(RC)
            o = c["st"].copy()
            ctx.download(o, mmd.F_F)
            res[prec] = (o, ctx.itl_get(c["itl"].n))
    a, ia = res[mmd.PREC_FP64]
    b, ib = res[mmd.PREC_MIXED]
    li = a.local_index()
    occ = li[a.type[li] >= 0]
    Fa = np.concatenate([np.stack([a.fx[occ], a.fy[occ], a.fz[occ]], 1), np.stack([ia.fx, ia.fy, ia.fz], 1)])
    Fb = np.concatenate([np.stack([b.fx[occ], b.fy[occ], b.fz[occ]], 1), np.stack([ib.fx, ib.fy, ib.fz], 1)])
    rms = np.sqrt((np.linalg.norm(Fa, axis=1) ** 2).mean())
    assert np.linalg.norm(Fb - Fa) / np.linalg.norm(Fa) <= 1e-6
    assert np.linalg.norm(Fb - Fa, axis=1).max() / rms <= 1e-6


def test_interstitial_errors():
    c = W.case(**ITL_CASES["bcc8_1type"])
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.upload(c["st"], mmd.F_X | mmd.F_TYPE)
    bad = c["itl"].take(c["itl"].n)
    bad.type[0] = 5
    with pytest.raises(mmd.MmdError) as e:
        ctx.itl_set(bad, c["wide3"])
    assert e.value.code == mmd.E_INVALID
    ctx.itl_set(c["itl"], c["wide3"])
    with pytest.raises(mmd.MmdError) as e:
        ctx.rho_df(RC)  # interstitial passes need the MD box
    assert e.value.code == mmd.E_STATE
    ctx.close()
