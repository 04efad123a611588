"""Interstitial migration on the device (mmd_itl_migrate) against the oracle's
sequential restatement (orc_migrate), through the C ABI.

Migration is integer bookkeeping on positions read as stored, so the bar is
bit-identity: every field of every ext site and of the interstitial list after
the call, and the transition counts.  Cases: a strongly disordered state (many
leavers, contested vacancies, periodic images) and the states a hot defect run
on the device reaches, migrating between chunks of steps (FP64 and mixed).
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC
FIELDS = mmd.HostState.FIELDS


def _ctx(c, cap, prec=mmd.PREC_FP64, fields=mmd.F_ALL):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.set_precision(prec)
    ctx.upload(c["st"], fields)
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=1)
    ctx.itl_set(c["itl"], c["wide3"], cap)
    return ctx


def _state(ctx, dom, cap):
    st = mmd.HostState(dom)
    ctx.download(st, mmd.F_ALL)
    return st, ctx.itl_get(cap)


def _same(st, itl, st_ref, itl_ref, where):
    for f in FIELDS + ("type",):
        a, b = getattr(st, f), getattr(st_ref, f)
        assert np.array_equal(a, b), f"{where}: lattice field {f} differs at {np.flatnonzero(a != b)[:8]}"
    assert itl.n == itl_ref.n, f"{where}: list length {itl.n} vs {itl_ref.n}"
    for f in FIELDS + ("type",):
        a, b = getattr(itl, f), getattr(itl_ref, f)
        assert np.array_equal(a, b), f"{where}: list field {f} differs at {np.flatnonzero(a != b)[:8]}"


def _n_atoms(st, itl):
    return int((st.type[st.local_index()] >= 0).sum()) + itl.n


@pytest.mark.parametrize("seed", [1, 2])
def test_migrate_disordered_state_matches_oracle(seed):
    c = W.case(cells=(10, 9, 8), n_types=3, sigma=0.6, temperature=900.0, vacancy_frac=0.1,
               interstitial_frac=0.02, seed=seed)
    st = c["st"]
    rng = np.random.default_rng(seed)
    for f in ("fx", "fy", "fz", "rho", "df"):
        getattr(st, f)[:] = rng.normal(size=st.x.size)
    cap = c["itl"].n + 800
    print(f"migrate: disordered 10x9x8, seed {seed}, {c['itl'].n} interstitials", flush=True)
    with _ctx(c, cap) as ctx:
        st0, itl0 = _state(ctx, c["dom"], cap)  # the list's f, rho, df start at zero on the device
        n0 = _n_atoms(st0, itl0)
        out, fill = ctx.itl_migrate()
        st1, itl1 = _state(ctx, c["dom"], cap)
        again = ctx.itl_migrate()
        st2, itl2 = _state(ctx, c["dom"], cap)
    ref_itl, rout, rfill = O.migrate(c["dom"], A, c["box"], st0, itl0, cap)
    assert (out, fill) == (rout, rfill)
    assert out > 50 and fill > 10, (out, fill)
    _same(st1, itl1, st0, ref_itl, "after migration")
    assert _n_atoms(st1, itl1) == n0
    assert again == (0, 0)
    _same(st2, itl2, st1, itl1, "second call")


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
def test_migrate_between_hot_md_chunks(prec):
    """A 1500 K defect box on the device, migrating every 25 steps: each call
    matches the oracle applied to the state the device held just before it;
    atoms are conserved and the total energy stays flat across migrations."""
    c = W.case(cells=(10, 10, 10), n_types=3, sigma=0.05, temperature=1500.0, vacancy_frac=0.03,
               interstitial_frac=0.01, seed=31)
    cap = c["itl"].n + 400
    moves = [0, 0]
    energies = []
    print(f"migrate: hot MD 10^3, precision {prec}", flush=True)
    with _ctx(c, cap, prec, fields=mmd.F_X | mmd.F_V | mmd.F_TYPE) as ctx:
        ctx.md_prime()
        st, itl = _state(ctx, c["dom"], cap)
        n0 = _n_atoms(st, itl)
        for chunk in range(8):
            ctx.md_run(25, use_graph=True)
            energies.append(sum(ctx.energy()))
            st, itl = _state(ctx, c["dom"], cap)
            out, fill = ctx.itl_migrate()
            ref_itl, rout, rfill = O.migrate(c["dom"], A, c["box"], st, itl, cap)
            assert (out, fill) == (rout, rfill), chunk
            st1, itl1 = _state(ctx, c["dom"], cap)
            _same(st1, itl1, st, ref_itl, f"chunk {chunk}")
            assert _n_atoms(st1, itl1) == n0
            moves[0] += out
            moves[1] += fill
    assert moves[0] > 0 and moves[1] > 0, moves
    assert np.ptp(energies) / n0 < 1e-4, energies


def test_migrate_over_capacity_changes_nothing():
    c = W.case(cells=(8, 8, 8), n_types=1, sigma=0.6, temperature=600.0, vacancy_frac=0.05, seed=3)
    cap = c["itl"].n + 2
    with _ctx(c, cap) as ctx:
        st0, itl0 = _state(ctx, c["dom"], cap)
        with pytest.raises(mmd.MmdError) as e:
            ctx.itl_migrate()
        assert e.value.code == mmd.E_RANGE
        st1, itl1 = _state(ctx, c["dom"], cap)
    _same(st1, itl1, st0, itl0, "after a refused migration")
