"""GPU parity on the named benchmark workloads that the other GPU tests do not
run as named, and the state rules of the MD entry points.

* The plain CPU reference configuration: a single-element BCC lattice of
  20^3 cells (16,000 atoms), a = 2.855 A, the named synthetic tables (5000
  knots on [0.5, 5.6] A, A = 1.2, B = 0.3, C = 0.01, no jitter), Gaussian
  displacement 0.03 A, Maxwell velocities at 600 K, 10 velocity-Verlet steps
  of 1 fs, through mmd_md_run (eager and graph replay) against the oracle's
  md_run step by step.
* A serving replica (3-type alloy of the named synthetic tables, 10 cells, 10
  steps) against the oracle's md_run on the same initial conditions: FP64 to
  the oracle bar, mixed precision to 1e-6 of the FP64 oracle's forces on the
  replica's own final positions.

Bars (the oracle's arithmetic, FMAs contracted on the device): positions and
velocities within 1e-10 A (A/fs) after 10 steps, forces within 1e-11 of the
largest force, per-step energies within 1e-12 relative; mixed precision
forces within 1e-6 of the RMS force (max per atom) and normwise.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC, SKIN = W.A, W.RC, W.SKIN


def _named(cells, n_types, sigma, temperature, seed, ghost_cells=3, type_frac=(0.9, 0.05, 0.05)):
    spec = mmd.lattice_spec(cells, ghost_cells=ghost_cells, a=A, n_types=n_types, type_frac=type_frac, sigma=sigma,
                            temperature=temperature, seed=seed)
    dom, box, st, _ = mmd.lattice_init(spec)
    tab = mmd.synth_tables(n_types=n_types, rc=RC, jitter=0.1 if n_types > 1 else 0.0, seed=seed)
    st3 = mmd.stencil_build(A, RC + SKIN, 0)
    return dict(dom=dom, box=box, st=st, tab=tab, st3=st3, even=mmd.stencil_linear(dom, st3),
                odd=mmd.stencil_linear(dom, -st3))


def _open(c, prec=mmd.PREC_FP64, energy_every=1):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], mmd.DEFAULT_MASSES)
    ctx.set_precision(prec)
    ctx.upload(c["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=energy_every)
    ctx.md_prime()
    return ctx


def _oracle_md(c, nsteps):
    ref = c["st"].copy()
    T = O.Tables(c["tab"], mmd.DEFAULT_MASSES)
    O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref)
    O.ghost_df(c["dom"], ref)
    O.force(c["dom"], c["even"], c["odd"], T, RC, ref)
    en = O.md_run(c["dom"], c["even"], c["odd"], T, RC, c["box"], 1.0, nsteps, ref)
    return ref, en, T


def _occupied(st):
    li = st.local_index()
    return li[st.type[li] >= 0]


def _forces(st, idx):
    return np.stack([st.fx[idx], st.fy[idx], st.fz[idx]], 1)


@pytest.mark.parametrize("use_graph", [False, True])
def test_cpu_reference_configuration_md_matches_oracle(use_graph):
    c = _named((20, 20, 20), n_types=1, sigma=0.03, temperature=600.0, seed=1)
    occ = _occupied(c["st"])
    assert occ.size == 16000
    nsteps = 10
    ref, en_ref, _ = _oracle_md(c, nsteps)
    with _open(c) as ctx:
        en = np.zeros((nsteps, 2))
        ctx.md_run(nsteps, en, use_graph=use_graph)
        out = c["st"].copy()
        ctx.download(out)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.abs(getattr(out, f)[occ] - getattr(ref, f)[occ]).max() < 1e-10, f
    F, Fr = _forces(out, occ), _forces(ref, occ)
    assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
    assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-12, atol=0)
    assert np.allclose(en, en_ref, rtol=1e-12, atol=0)
    assert np.array_equal(out.type, ref.type)


def test_cpu_reference_configuration_mixed_within_bound():
    """Mixed precision on the same run: forces within 1e-6 of the FP64 oracle's
    on the mixed run's own final positions, energies within 1e-9 eV/atom."""
    c = _named((20, 20, 20), n_types=1, sigma=0.03, temperature=600.0, seed=1)
    occ = _occupied(c["st"])
    with _open(c, prec=mmd.PREC_MIXED) as ctx:
        en = np.zeros((10, 2))
        ctx.md_run(10, en)
        out = c["st"].copy()
        ctx.download(out)
    ref = out.copy()
    T = O.Tables(c["tab"], mmd.DEFAULT_MASSES)
    e_rho = O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
    O.ghost_df(c["dom"], ref)
    e_f = O.force(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
    F, Fr = _forces(out, occ), _forces(ref, occ)
    rms = np.sqrt((np.linalg.norm(Fr, axis=1) ** 2).mean())
    assert np.linalg.norm(F - Fr, axis=1).max() <= 1e-6 * rms
    assert np.linalg.norm(F - Fr) <= 1e-6 * np.linalg.norm(Fr)
    assert abs(en[-1, 0] - (e_rho + e_f)) / occ.size < 1e-9


GHOST = 2  # the serving replicas' shell: the stencil's reach
TYPE_FRAC = (0.9, 0.05, 0.05)


def _keep(cells):
    spec = mmd.lattice_spec((cells,) * 3, ghost_cells=GHOST, a=A)
    dom = mmd.Domain()
    box = np.zeros(3)
    mmd._check(mmd.lib().mmd_lattice_domain(mmd.C.byref(spec), mmd.C.byref(dom), mmd.ptr(box)))
    return mmd.HostState(dom)


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
def test_serving_replica_matches_oracle(prec):
    cells, temp, seed, nsteps = 10, 900.0, 31, 10
    c = _named((cells,) * 3, n_types=3, sigma=0.05, temperature=temp, seed=seed, ghost_cells=GHOST,
               type_frac=TYPE_FRAC)
    tab = c["tab"]
    keep = _keep(cells)
    with mmd.Server(tab, precision=prec, a=A, cutoff=RC, skin=SKIN, dt=1.0, ghost_cells=GHOST,
                    type_frac=TYPE_FRAC, sigma=0.05, max_replicas=4, hbm_fraction=0.1, chunk_steps=4,
                    workers=2) as srv:
        # another replica shares the batch for the whole run
        res, st = srv.run([dict(cells=cells, temperature=temp, steps=nsteps, seed=seed, arrival_s=0.0, keep=keep),
                           dict(cells=7, temperature=1200.0, steps=3 * nsteps, seed=5, arrival_s=0.0)])
    assert all(r["status"] == 0 for r in res), res
    occ = _occupied(c["st"])
    assert np.array_equal(keep.type, c["st"].type)
    if prec == mmd.PREC_FP64:
        ref, en_ref, _ = _oracle_md(c, nsteps)
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.abs(getattr(keep, f)[occ] - getattr(ref, f)[occ]).max() < 1e-10, f
        F, Fr = _forces(keep, occ), _forces(ref, occ)
        assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
        assert abs(res[0]["epot"] - en_ref[-1, 0]) <= 1e-12 * abs(en_ref[-1, 0])
        assert abs(res[0]["ekin"] - en_ref[-1, 1]) <= 1e-12 * abs(en_ref[-1, 1])
    else:
        ref = keep.copy()
        T = O.Tables(tab, mmd.DEFAULT_MASSES)
        e_rho = O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
        O.ghost_df(c["dom"], ref)
        e_f = O.force(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
        F, Fr = _forces(keep, occ), _forces(ref, occ)
        rms = np.sqrt((np.linalg.norm(Fr, axis=1) ** 2).mean())
        assert np.linalg.norm(F - Fr, axis=1).max() <= 1e-6 * rms
        assert abs(res[0]["epot"] - (e_rho + e_f)) / occ.size < 1e-9
        assert abs(res[0]["ekin"] - O.kinetic(c["dom"], T, keep)) <= 1e-12 * res[0]["ekin"]


def test_md_calls_refuse_a_stale_setup():
    """A new domain needs mmd_md_setup again; tables with fewer types than the
    atoms use are refused before any launch."""
    c = _named((6, 6, 6), n_types=3, sigma=0.05, temperature=600.0, seed=3)
    with _open(c, energy_every=0) as ctx:
        ctx.md_run(2)
        # tables of one type while types 1 and 2 are uploaded
        ctx.set_tables(mmd.synth_tables(n_types=1, rc=RC), mmd.DEFAULT_MASSES)
        for call in (lambda: ctx.md_run(1), ctx.md_prime, ctx.energy):
            with pytest.raises(mmd.MmdError) as e:
                call()
            assert e.value.code == mmd.E_INVALID
        ctx.set_tables(c["tab"], mmd.DEFAULT_MASSES)
        ctx.md_run(1)
        # a new domain: the MD parameters belong to the old one
        c2 = _named((7, 6, 5), n_types=3, sigma=0.05, temperature=600.0, seed=4)
        ctx.set_domain(c2["dom"])
        ctx.set_stencil(c2["even"], c2["odd"])
        ctx.upload(c2["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
        for call in (lambda: ctx.md_run(1), ctx.md_prime, ctx.energy, lambda: ctx.ghost_refresh(mmd.F_X)):
            with pytest.raises(mmd.MmdError) as e:
                call()
            assert e.value.code == mmd.E_STATE
        ctx.md_setup(c2["box"], A, 1.0, RC)
        ctx.md_prime()
        ctx.md_run(1)
