"""mmd_md_run_noalloc (the MD run that allocates nothing, its scratch in a
library-owned device array) against mmd_md_run through the C ABI: the same
kernels in the same order, so the bar is bit-identity of the whole
ghost-extended state, the interstitial list and the recorded energies, on two
calls in a row (the second on a smaller box, which must not see the first
call's scratch contents), and the refusal of a call whose partials and
records exceed the scratch array."""
import numpy as np
import pytest

import mmd
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC


def _open(c, energy_every):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.set_precision(mmd.PREC_MIXED)
    ctx.upload(c["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=energy_every)
    if c["itl"].n:
        ctx.itl_set(c["itl"], c["wide3"], cap=c["itl"].n + 8)
    ctx.md_prime()
    return ctx


def _run(c, nsteps, every, noalloc):
    with _open(c, every) as ctx:
        en = np.full((nsteps, 2), -7.0)
        if noalloc:
            ctx.md_run_noalloc(nsteps, en)
        else:
            ctx.md_run(nsteps, en)
        out = c["st"].copy()
        ctx.download(out)
        itl = ctx.itl_get(c["itl"].n) if c["itl"].n else None
    return out, itl, en


def _same(a, b):
    (sa, ia, ea), (sb, ib, eb) = a, b
    for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df", "type"):
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
    if ia is not None:
        for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df"):
            assert np.array_equal(getattr(ia, f), getattr(ib, f)), f
    assert np.array_equal(ea, eb)


def test_noalloc_equals_md_run_on_two_calls():
    big = W.case(cells=(12, 11, 10), n_types=3, sigma=0.05, temperature=900.0, vacancy_frac=0.01,
                 interstitial_frac=0.002, seed=21)
    small = W.case(cells=(7, 6, 6), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01, seed=22)
    assert big["itl"].n > 0
    _same(_run(big, 12, 3, True), _run(big, 12, 3, False))
    # the smaller box right after: fewer blocks and records, no interstitials
    r = _run(small, 7, 2, True)
    _same(r, _run(small, 7, 2, False))
    assert np.all(r[2][[1, 3, 5]] != -7.0) and np.all(r[2][[0, 2, 4, 6]] == -7.0)  # unrecorded rows untouched


def test_noalloc_refuses_beyond_its_scratch():
    c = W.case(cells=(6, 6, 6), n_types=1, sigma=0.03, temperature=300.0, seed=23)
    with _open(c, 1) as ctx:
        before = c["st"].copy()
        ctx.download(before)
        with pytest.raises(mmd.MmdError) as e:
            ctx.md_run_noalloc(600000)  # 600,000 energy records need 1.2 M doubles
        assert e.value.code == mmd.E_RANGE
        after = c["st"].copy()
        ctx.download(after)
        assert np.array_equal(before.x, after.x) and np.array_equal(before.vx, after.vx)
