"""Newton's-third-law half stencil of the periodic step (mmd_set_half_stencil)
against the CPU oracle, which walks the full stencil as the reference does.

Each pair's force is evaluated once and its reaction pulled by the neighbour,
so a site's force is the same sum in another order: FP64 forces agree with the
oracle to the lattice path's bar (1e-11 of the largest force) and with the
full-stencil device pass to 1e-13; mixed precision to its 1e-6 bound.  The
net force of a periodic box stays zero up to rounding.  Pair energies count
each pair once in both modes.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC

CASES = {
    "bcc12_3type_vac": dict(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01,
                            seed=2),
    "ragged_7x9x5": dict(cells=(7, 9, 5), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.02, seed=3),
    "wide_row_70x3x4": dict(cells=(70, 3, 4), n_types=2, sigma=0.05, temperature=800.0, seed=4, ghost_cells=2),
    "wide_stencil_168": dict(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=900.0, vacancy_frac=0.01, seed=5,
                             skin=1.9),
}


def _ctx(c, prec, half, energy_every=0):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.set_precision(prec)
    ctx.upload(c["st"], mmd.F_X | mmd.F_TYPE | mmd.F_V)
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=energy_every)
    if half:
        ctx.set_half_stencil(True)
    return ctx


def _primed_forces(c, prec, half):
    with _ctx(c, prec, half) as ctx:
        ctx.md_prime()
        out = c["st"].copy()
        ctx.download(out, mmd.F_F | mmd.F_RHO | mmd.F_DF)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    return out, occ, np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1)


def _oracle_forces(c, occ):
    ref = c["st"].copy()
    T = O.Tables(c["tab"], c["masses"])
    O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref)
    O.ghost_df(c["dom"], ref)
    O.force(c["dom"], c["even"], c["odd"], T, RC, ref)
    return np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
@pytest.mark.parametrize("name", list(CASES))
def test_half_stencil_forces_match_oracle(name, prec):
    c = W.case(**CASES[name])
    out, occ, F = _primed_forces(c, prec, True)
    _, occ_f, Ff = _primed_forces(c, prec, False)
    assert np.array_equal(occ, occ_f)
    Fr = _oracle_forces(c, occ)
    li = out.local_index()
    vac = li[out.type[li] < 0]
    assert np.all(out.fx[vac] == 0)  # vacancies untouched
    if prec == mmd.PREC_FP64:
        assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
        assert np.abs(F - Ff).max() <= 1e-13 * np.abs(Ff).max()
    else:
        rms = np.sqrt((Fr ** 2).sum(1).mean())
        assert np.linalg.norm(F - Fr, axis=1).max() <= 1e-6 * rms
        assert np.abs(F - Ff).max() <= 1e-12 * np.abs(Ff).max()
    assert np.abs(F.sum(0)).max() <= 1e-13 * np.abs(F).sum()


def test_half_stencil_md_run_and_energies_match_full():
    """Ten steps with energies every step, the half-stencil run replayed from
    captured graphs: trajectories and energies agree with the full stencil."""
    c = W.case(cells=(10, 10, 10), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01, seed=9)
    res = {}
    for half in (False, True):
        with _ctx(c, mmd.PREC_FP64, half, energy_every=1) as ctx:
            ctx.md_prime()
            en = np.zeros((10, 2))
            ctx.md_run(5, en[:5], use_graph=half)
            ctx.sync()
            ctx.md_run(5, en[5:], use_graph=half)
            ep, ek = ctx.energy()
            out = c["st"].copy()
            ctx.download(out)
        res[half] = (out, en, ep, ek)
    (a, ea, epa, eka), (b, eb, epb, ekb) = res[False], res[True]
    li = a.local_index()
    occ = li[a.type[li] >= 0]
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.abs(getattr(a, f)[occ] - getattr(b, f)[occ]).max() < 1e-12, f
    n_at = len(occ)
    assert np.abs(ea - eb).max() / n_at < 1e-12
    assert abs(epa - epb) / n_at < 1e-12 and abs(eka - ekb) / n_at < 1e-12


def test_half_stencil_graph_replay_equals_eager():
    c = W.case(cells=(8, 9, 10), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01, seed=12)
    outs = []
    for use_graph in (False, True):
        with _ctx(c, mmd.PREC_MIXED, True, energy_every=2) as ctx:
            ctx.md_prime()
            en = np.zeros((6, 2))
            ctx.md_run(3, en[:3], use_graph=use_graph)
            ctx.sync()
            ctx.md_run(3, en[3:], use_graph=use_graph)
            o = c["st"].copy()
            ctx.download(o)
            outs.append((o, en))
    (a, ea), (b, eb) = outs
    for f in ("x", "vx", "fx", "rho", "df"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(ea, eb)


def test_half_stencil_with_interstitials_matches_oracle():
    c = W.case(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.005,
               interstitial_frac=0.006, seed=22)
    assert c["itl"].n >= 2
    st, itl = c["st"].copy(), c["itl"].take(c["itl"].n)
    T = O.Tables(c["tab"], c["masses"])
    O.rho_df_itl(c["dom"], c["even"], c["odd"], c["wide3"], A, c["box"], T, RC, st, itl)
    O.ghost_df(c["dom"], st)
    O.force_itl(c["dom"], c["even"], c["odd"], c["wide3"], A, c["box"], T, RC, st, itl)
    with _ctx(c, mmd.PREC_FP64, True) as ctx:
        ctx.itl_set(c["itl"], c["wide3"])
        ctx.md_prime()
        out = c["st"].copy()
        ctx.download(out, mmd.F_F)
        gi = ctx.itl_get(c["itl"].n)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    F = np.concatenate([np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1), np.stack([gi.fx, gi.fy, gi.fz], 1)])
    Fr = np.concatenate([np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1), np.stack([itl.fx, itl.fy, itl.fz], 1)])
    assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
    assert np.abs(F.sum(0)).max() <= 1e-13 * np.abs(F).sum()


def test_half_stencil_state_rules():
    """An asymmetric stencil refuses the mode; a new stencil turns it off."""
    c = W.case(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01, seed=2)
    with mmd.Context(0) as ctx:
        ctx.set_domain(c["dom"])
        with pytest.raises(mmd.MmdError) as e:
            ctx.set_half_stencil(True)
        assert e.value.code == mmd.E_STATE
        ctx.set_stencil(c["even"], c["odd"][:-1])
        with pytest.raises(mmd.MmdError) as e:
            ctx.set_half_stencil(True)
        assert e.value.code == mmd.E_INVALID
        ctx.set_stencil(c["even"], c["odd"])
        ctx.set_half_stencil(True)
        ctx.set_half_stencil(False)
        ctx.set_half_stencil(True)
