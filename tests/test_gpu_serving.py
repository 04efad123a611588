"""GPU serving loop: a replica's result is the same alone, in a mix and on the
single-domain device path.

The batched kernels run a replica's blocks with the single-domain per-block
code and numbering, and its energies are reduced in the same fixed order, so
the bar is bit-identity (FP64 and mixed): final positions, velocities,
forces, rho and df over the whole ghost-extended lattice, and the admission
and final energies.
"""
import numpy as np
import pytest

import mmd
import workload as W

pytestmark = pytest.mark.gpu

A, RC, SKIN, GHOST = W.A, W.RC, W.SKIN, 3
SIGMA = 0.05
TYPE_FRAC = (0.9, 0.05, 0.05)


@pytest.fixture(scope="module")
def tab():
    return mmd.fs_tables(n_types=3, rc=RC, jitter=0.1, seed=11)


def _server(tab, prec, **kw):
    opts = dict(precision=prec, a=A, cutoff=RC, skin=SKIN, dt=1.0, ghost_cells=GHOST, type_frac=TYPE_FRAC,
                sigma=SIGMA, max_replicas=8, hbm_fraction=0.2, chunk_steps=4, workers=4)
    opts.update(kw)
    return mmd.Server(tab, **opts)


def _keep(cells):
    dom = mmd.Domain()
    spec = mmd.lattice_spec((cells,) * 3, ghost_cells=GHOST, a=A)
    box = np.zeros(3)
    mmd._check(mmd.lib().mmd_lattice_domain(mmd.C.byref(spec), mmd.C.byref(dom), mmd.ptr(box)))
    return mmd.HostState(dom)


def _single_domain(tab, prec, cells, temperature, seed, steps):
    """The same replica on the single-domain device path (mmd_ctx)."""
    spec = mmd.lattice_spec((cells,) * 3, ghost_cells=GHOST, a=A, n_types=3, type_frac=TYPE_FRAC, sigma=SIGMA,
                            temperature=temperature, seed=seed)
    dom, box, st, _ = mmd.lattice_init(spec)
    st3 = mmd.stencil_build(A, RC + SKIN, 0)
    with mmd.Context(0) as ctx:
        ctx.set_domain(dom)
        ctx.set_stencil(mmd.stencil_linear(dom, st3), mmd.stencil_linear(dom, -st3))
        ctx.set_tables(tab, mmd.DEFAULT_MASSES)
        ctx.set_precision(prec)
        ctx.upload(st, mmd.F_X | mmd.F_TYPE | mmd.F_V)
        ctx.md_setup(box, A, 1.0, RC, energy_every=0)
        ctx.md_prime()
        e0 = ctx.energy()
        if steps:
            ctx.md_run(steps, use_graph=True)
        e1 = ctx.energy()
        out = st.copy()
        ctx.download(out, mmd.F_ALL)
    return out, e0, e1


FIELDS = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df", "type")


def _same(a, b, where):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(x, y), f"{where}: field {f} differs (max |d| = {np.max(np.abs(x - y))})"


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
def test_alone_mix_and_single_domain_identical(tab, prec):
    probe = dict(id=7, cells=9, temperature=900.0, steps=10, seed=42, arrival_s=0.0)
    # alone
    alone = _keep(9)
    with _server(tab, prec) as srv:
        res, st = srv.run([dict(probe, keep=alone)])
        assert res[0]["status"] == 0 and st["completed"] == 1
        r_alone = res[0]
        # in a mix: others of other sizes, running before, during and after the probe
        mixed = _keep(9)
        # a long-running replica is resident when the probe arrives, others come and go around it
        others = [dict(cells=c, temperature=t, steps=s, seed=100 + k, arrival_s=a)
                  for k, (c, t, s, a) in enumerate([(14, 300.0, 4000, 0.0), (6, 1500.0, 5, 0.0), (11, 1200.0, 17, 0.0),
                                                    (8, 600.0, 9, 0.05), (12, 800.0, 30, 0.08)])]
        res, st = srv.run(others[:3] + [dict(probe, arrival_s=0.02, keep=mixed)] + others[3:])
        assert all(r["status"] == 0 for r in res), res
        r_mix, long_ = res[3], res[0]
        assert long_["t_ready"] < r_mix["t_done"] and r_mix["t_ready"] < long_["t_done"], (long_, r_mix)
        assert st["max_active"] >= 2
    _same(alone, mixed, "alone vs mix")
    for k in ("epot0", "ekin0", "epot", "ekin"):
        assert r_alone[k] == r_mix[k], k
    # single-domain path
    ref, e0, e1 = _single_domain(tab, prec, 9, 900.0, 42, 10)
    _same(alone, ref, "server vs single-domain")
    assert (r_alone["epot0"], r_alone["ekin0"]) == e0
    assert (r_alone["epot"], r_alone["ekin"]) == e1


def test_graph_and_direct_launch_identical(tab):
    reps = [dict(cells=c, temperature=800.0, steps=12, seed=s, arrival_s=0.0) for c, s in ((7, 1), (10, 2))]
    outs = []
    for use_graph in (True, False):
        keeps = [_keep(7), _keep(10)]
        with _server(tab, mmd.PREC_MIXED, use_graph=use_graph) as srv:
            res, _ = srv.run([dict(r, keep=k) for r, k in zip(reps, keeps)])
        assert all(r["status"] == 0 for r in res)
        outs.append((keeps, [(r["epot"], r["ekin"]) for r in res]))
    for a, b in zip(outs[0][0], outs[1][0]):
        _same(a, b, "graph vs direct")
    assert outs[0][1] == outs[1][1]


def test_memory_bounded_admission_on_device(tab):
    """An arena with room for two 12-cell replicas: all eight complete, at most
    two resident at a time, energies conserved over the run."""
    one = mmd.replica_bytes(12, GHOST)
    reps = [dict(cells=12, temperature=600.0, steps=8, seed=s, arrival_s=0.0) for s in range(8)]
    with _server(tab, mmd.PREC_MIXED, hbm_limit=2 * one + 1024) as srv:
        res, st = srv.run(reps)
    assert st["completed"] == 8 and st["failed"] == 0
    assert st["max_active"] <= 2
    for r in res:
        n = r["n_atoms"]
        drift = abs((r["epot"] + r["ekin"]) - (r["epot0"] + r["ekin0"])) / n
        assert drift < 5e-5, r


def test_too_large_replica_rejected_on_device(tab):
    one = mmd.replica_bytes(8, GHOST)
    reps = [dict(cells=8, temperature=600.0, steps=3, seed=1, arrival_s=0.0),
            dict(cells=30, temperature=600.0, steps=3, seed=2, arrival_s=0.0)]
    with _server(tab, mmd.PREC_FP64, hbm_limit=2 * one) as srv:
        res, st = srv.run(reps)
    assert res[0]["status"] == 0 and res[1]["status"] == mmd.E_NOMEM
