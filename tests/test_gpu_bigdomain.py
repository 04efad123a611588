"""Domains past 2^27 ext sites: every position, df and rho record is addressed
with a 64-bit offset (a 32-bit byte offset of a 32-byte position record would
wrap at 2^27 sites and read another site's record).

* 828 x 414 x 406 ext sites (139.2 M, about 14 GB of device state): a perfect
  single-element BCC box of 408 x 408 x 400 cells with the named synthetic
  tables.  A record read from the wrong site would sit at another lattice
  position and change that atom's density and force, so every atom must have
  the density the oracle gives a small perfect box (closed-form shell sum of
  the same spline), the df of that density, and zero force.
* A domain of 2^31 ext sites or more is refused with MMD_E_RANGE before any
  allocation, and the context stays usable.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC


def _perfect_state(dom, a):
    """Ideal BCC positions over the ext lattice, reference layout (x doubled):
    site X even at (X/2, Y, Z) a, X odd at (X/2 + 1/2, Y + 1/2, Z + 1/2) a,
    ghost sites at their unwrapped positions (the periodic images)."""
    st = mmd.HostState(dom)
    ex, ey, ez = dom.ext
    gx, gy, gz = dom.ghost
    X = np.arange(ex) - gx
    half = 0.5 * a * (X & 1)
    xs = a * (X >> 1) + half
    ys = (a * (np.arange(ey) - gy))[:, None] + half[None, :]
    plane = ex * ey
    xp = np.tile(xs, ey)
    for z in range(ez):
        sl = slice(z * plane, (z + 1) * plane)
        st.x[sl] = xp
        st.y[sl] = ys.ravel()
        st.z[sl] = np.tile(a * (z - gz) + half, ey)
    st.type[:] = 0
    return st


def _local(arr, dom):
    ex, ey, ez = dom.ext
    gx, gy, gz = dom.ghost
    bx, by, bz = dom.box
    return arr.reshape(ez, ey, ex)[gz:gz + bz, gy:gy + by, gx:gx + bx]


def test_domain_past_2e27_sites_perfect_lattice():
    tab = mmd.synth_tables(n_types=1, rc=RC)
    st3 = mmd.stencil_build(A, RC + W.SKIN, 0)
    # the oracle's density and df on a small perfect box of the same lattice
    small = W.case(cells=(8, 8, 8), n_types=1, sigma=0.0, temperature=0.0, seed=1)
    ref = small["st"].copy()
    T = O.Tables(tab, mmd.DEFAULT_MASSES)
    O.rho_df(small["dom"], small["even"], small["odd"], T, RC, ref)
    li = ref.local_index()
    rho0, df0 = ref.rho[li[0]], ref.df[li[0]]
    assert np.allclose(ref.rho[li], rho0, rtol=1e-13)

    cells = (408, 408, 400)
    print("big-domain case: 408 x 408 x 400 cells, 139.2 M ext sites", flush=True)
    spec = mmd.lattice_spec(cells, ghost_cells=3, a=A)
    dom = mmd.Domain()
    box = np.zeros(3)
    mmd._check(mmd.lib().mmd_lattice_domain(mmd.C.byref(spec), mmd.C.byref(dom), mmd.ptr(box)))
    assert dom.n_ext() > 2 ** 27
    st = _perfect_state(dom, A)
    with mmd.Context(0) as ctx:
        ctx.set_domain(dom)
        ctx.set_stencil(mmd.stencil_linear(dom, st3), mmd.stencil_linear(dom, -st3))
        ctx.set_tables(tab, mmd.DEFAULT_MASSES)
        ctx.upload(st, mmd.F_X | mmd.F_TYPE)
        ctx.md_setup(box, A, 1.0, RC)
        ctx.rho_df(RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        ctx.download(st, mmd.F_RHO | mmd.F_DF | mmd.F_F)
    rho = _local(st.rho, dom)
    assert np.abs(rho - rho0).max() <= 1e-12 * rho0
    # df = -1/(2 sqrt(rho)) + 2 C rho: two terms of ~0.2 cancel to ~0.04, so
    # rho's last-bit differences (these positions round differently from the
    # small box's) reach df at a few 1e-13 relative
    assert np.abs(_local(st.df, dom) - df0).max() <= 1e-11 * abs(df0)
    # single pair forces are ~1 eV/A here; the net force vanishes by symmetry
    for f in ("fx", "fy", "fz"):
        assert np.abs(_local(getattr(st, f), dom)).max() < 1e-9, f


def test_domain_of_2e31_sites_refused():
    d = mmd.Domain()
    for k, b in enumerate((2048, 1024, 1024)):
        d.box[k], d.ghost[k], d.ext[k], d.start[k] = b, 0, b, 0
    assert d.n_ext() == 2 ** 31
    with mmd.Context(0) as ctx:
        with pytest.raises(mmd.MmdError) as e:
            ctx.set_domain(d)
        assert e.value.code == mmd.E_RANGE
        c = W.case(cells=(4, 4, 4), n_types=1, seed=1)
        ctx.set_domain(c["dom"])  # the context still works after the refusal
