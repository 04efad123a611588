"""GPU parity of the EAM hot path (rho + fused df, df, force), the drop-in
host-array hooks, the ghost refresh and the device MD step against the CPU
oracle, through the C ABI.

Tolerances: the FP64 path reproduces the oracle's arithmetic (same spline,
same stencil order); the device contracts multiply-adds into FMAs, so results
agree to rounding: rho, df relative <= 1e-12, forces absolute <= 1e-11 of the
largest force component.  The mixed path is held to the stated target: force
error <= 1e-6 of the RMS force (per atom, max) and normwise.  Integer fields
(types, vacancy flags) are compared bit-exactly.
"""
import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

pytestmark = pytest.mark.gpu

A, RC = W.A, W.RC


def _ctx(c, prec=mmd.PREC_FP64, fields=mmd.F_X | mmd.F_TYPE | mmd.F_V):
    ctx = mmd.Context(0)
    ctx.set_domain(c["dom"])
    ctx.set_stencil(c["even"], c["odd"])
    ctx.set_tables(c["tab"], c["masses"])
    ctx.set_precision(prec)
    ctx.upload(c["st"], fields)
    return ctx


def _oracle_pass(c):
    ref = c["st"].copy()
    T = O.Tables(c["tab"], c["masses"])
    O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref)
    O.ghost_df(c["dom"], ref)
    O.force(c["dom"], c["even"], c["odd"], T, RC, ref)
    return ref, T


CASES = {
    "bcc20_1type": dict(cells=(20, 20, 20), n_types=1, sigma=0.03, temperature=600.0, seed=1),
    "bcc12_3type_vac": dict(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01,
                            seed=2),
    "ragged_7x9x5": dict(cells=(7, 9, 5), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.02, seed=3),
    "wide_row_70x3x4": dict(cells=(70, 3, 4), n_types=2, sigma=0.05, temperature=800.0, seed=4, ghost_cells=2),
    # rho_t differs per type, so rho by the neighbour's type is tested
    "bcc12_3type_distinct_rho": dict(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01,
                                     seed=6, distinct_rho=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_rho_df_force_fp64_match_oracle(name):
    c = W.case(**CASES[name])
    ref, _ = _oracle_pass(c)
    with _ctx(c) as ctx:
        ctx.rho_df(RC)
        out = c["st"].copy()
        ctx.download(out, mmd.F_RHO | mmd.F_DF | mmd.F_TYPE)
        # force from the device's own df with a device ghost refresh of df
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        ctx.download(out, mmd.F_F | mmd.F_DF)
    li = out.local_index()
    assert np.array_equal(out.type, c["st"].type)
    occ = li[out.type[li] >= 0]
    assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-12, atol=0)
    assert np.allclose(out.df[occ], ref.df[occ], rtol=1e-12, atol=1e-15)
    assert np.allclose(out.df, ref.df, rtol=1e-12, atol=1e-15)  # ghost df refreshed like the oracle's
    F = np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1)
    Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
    assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
    vac = li[out.type[li] < 0]
    assert np.all(out.rho[vac] == 0) and np.all(out.fx[vac] == 0)  # vacancies untouched


def test_funcfl_tables_match_oracle(tmp_path):
    """Tables read from two funcfl files (mmd_funcfl_read: F, rho and the
    effective charges, phi_ij = 27.2 * 0.529 * Z_i Z_j / r) drive the device
    passes on a 2-type box with vacancies, to the oracle bar."""
    fa, fb = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    W.write_funcfl(fa)
    W.write_funcfl(fb, z_scale=1.3, C=0.02, mass=58.693)
    tab, mass = mmd.funcfl_read([fa, fb])
    c = W.case(cells=(7, 6, 5), n_types=2, sigma=0.05, temperature=900.0, vacancy_frac=0.02, seed=11)
    c["tab"], c["masses"] = tab, mass
    ref, _ = _oracle_pass(c)
    with _ctx(c) as ctx:
        ctx.rho_df(RC)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        out = c["st"].copy()
        ctx.download(out, mmd.F_RHO | mmd.F_DF | mmd.F_F)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    assert set(np.unique(out.type[occ])) == {0, 1}
    assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-12, atol=0)
    assert np.allclose(out.df[occ], ref.df[occ], rtol=1e-12, atol=1e-15)
    F = np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1)
    Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
    assert np.abs(Fr).max() > 0.1
    assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
def test_wide_stencil_large_lds_tile(prec):
    """A 7.5 A stencil (168 entries, reach 3 cells each way) needs a 92 KB
    LDS tile per force block, beyond the default 64 KB: exercises the opt-in
    LDS path.  FP64 is held to the oracle bar, mixed to its force bound."""
    c = W.case(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=900.0, vacancy_frac=0.01, seed=5, skin=1.9)
    assert len(c["st3"]) == 168
    ref, _ = _oracle_pass(c)
    with _ctx(c, prec) as ctx:
        ctx.rho_df(RC)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        out = c["st"].copy()
        ctx.download(out, mmd.F_RHO | mmd.F_DF | mmd.F_F)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    F = np.stack([out.fx[occ], out.fy[occ], out.fz[occ]], 1)
    Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
    if prec == mmd.PREC_FP64:
        assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-12, atol=0)
        assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
    else:
        assert np.allclose(out.rho[occ], ref.rho[occ], rtol=1e-8, atol=0)
        rms = np.sqrt((Fr ** 2).sum(1).mean())
        assert np.linalg.norm(F - Fr, axis=1).max() <= 1e-6 * rms


def test_standalone_df_matches_oracle():
    c = W.case(**CASES["bcc12_3type_vac"])
    ref, T = _oracle_pass(c)
    st = c["st"].copy()
    st.rho[:] = ref.rho
    with _ctx(c, fields=mmd.F_TYPE | mmd.F_RHO) as ctx:
        ctx.upload(st, mmd.F_RHO)
        ctx.df()
        out = st.copy()
        ctx.download(out, mmd.F_DF)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    assert np.allclose(out.df[occ], ref.df[occ], rtol=1e-13, atol=1e-15)


def test_perfect_lattice_on_device():
    c = W.case(cells=(16, 16, 16), sigma=0.0, temperature=0.0)
    # the synthetic forms are non-zero out to rc, so every shell of the stencil counts
    c["tab"] = mmd.synth_tables(rc=RC)
    with _ctx(c) as ctx:
        ctx.rho_df(RC)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        out = c["st"].copy()
        ctx.download(out, mmd.F_RHO | mmd.F_F)
    li = out.local_index()
    shells = [(8, np.sqrt(3) / 2 * A), (6, A), (12, np.sqrt(2) * A), (24, np.sqrt(11) / 2 * A), (8, np.sqrt(3) * A)]
    rho_cf = sum(n * (RC - r) ** 2 * np.exp(-r) for n, r in shells)
    assert np.allclose(out.rho[li], rho_cf, rtol=1e-12)
    assert np.abs(np.stack([out.fx[li], out.fy[li], out.fz[li]])).max() < 1e-12


def test_results_bit_reproducible():
    c = W.case(**CASES["bcc12_3type_vac"])
    outs = []
    with _ctx(c) as ctx:
        for _ in range(2):
            ctx.rho_df(RC)
            ctx.force(RC)
            o = c["st"].copy()
            ctx.download(o, mmd.F_RHO | mmd.F_F)
            outs.append(o)
    assert np.array_equal(outs[0].rho, outs[1].rho) and np.array_equal(outs[0].fz, outs[1].fz)


class _PinnedRecords:
    """n _cuAtomElement records in pinned host memory (mmd_host_alloc, the
    allocation hip_create_atoms_mem makes), as a numpy view."""

    def __init__(self, n):
        import ctypes as C
        self.p = C.c_void_p()
        nbytes = n * mmd.REF_AOS_DTYPE.itemsize
        mmd._check(mmd.lib().mmd_host_alloc(C.byref(self.p), nbytes))
        self.recs = np.frombuffer((C.c_char * nbytes).from_address(self.p.value), dtype=mmd.REF_AOS_DTYPE)
        self.recs[...] = np.zeros(1, mmd.REF_AOS_DTYPE)

    def __enter__(self):
        return self.recs

    def __exit__(self, *a):
        del self.recs
        mmd.lib().mmd_host_free(self.p)


@pytest.mark.parametrize("memory", ["pageable", "pinned"])
def test_dropin_aos_hooks_match_oracle(memory):
    """hip_eam_rho_calc / hip_eam_force_calc contract: host AoS records in,
    rho/df (or f) of local records out, everything else untouched.  Pageable
    records take the whole-array copies; pinned ones are copied in by
    z-chunks while earlier chunks are computed, and the outputs are written
    straight into the records.  The 7x9x5 box's 11 ext z-layers make six
    chunks of two layers."""
    c = W.case(**CASES["ragged_7x9x5"])
    n = c["st"].x.size
    if memory == "pinned":
        with _PinnedRecords(n) as recs:
            _dropin_checks(c, recs)
    else:
        _dropin_checks(c, np.zeros(n, mmd.REF_AOS_DTYPE))


def _dropin_checks(c, recs):
    ref, _ = _oracle_pass(c)
    st = c["st"]
    n = st.x.size
    recs["id"] = np.arange(n)
    recs["type"] = st.type
    recs["x"] = np.stack([st.x, st.y, st.z], 1)
    recs["v"] = np.stack([st.vx, st.vy, st.vz], 1)
    recs["rho"] = -7.0
    recs["df"] = -7.0
    before = recs.copy()
    with _ctx(c) as ctx:
        ctx.rho_calc_aos(recs, RC)
        li = st.local_index()
        occ = li[st.type[li] >= 0]
        assert np.allclose(recs["rho"][occ], ref.rho[occ], rtol=1e-12)
        assert np.allclose(recs["df"][occ], ref.df[occ], rtol=1e-12)
        mask = np.ones(n, bool)
        mask[occ] = False
        assert np.all(recs["rho"][mask] == -7.0)  # ghosts and vacancies untouched
        for f in ("id", "type", "x", "v", "f"):
            assert np.array_equal(recs[f], before[f])
        # the caller refreshes ghost df between the hooks (MISA-MD core does it over MPI)
        recs["df"] = ref.df
        ctx.force_calc_aos(recs, RC)
        F = recs["f"][occ]
        Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
        assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
        # standalone df hook on the records' rho
        recs["df"] = 0.0
        ctx.df_calc_aos(recs)
        assert np.allclose(recs["df"][occ], ref.df[occ], rtol=1e-13)


def test_dropin_rejects_type_without_table():
    """A record whose type has no table fails the hook with MMD_E_INVALID and
    leaves the host records as they were; the next valid call computes as
    before (the device never indexed a table with that type)."""
    c = W.case(**CASES["ragged_7x9x5"])
    st = c["st"]
    n = st.x.size
    recs = np.zeros(n, mmd.REF_AOS_DTYPE)
    recs["type"] = st.type
    recs["x"] = np.stack([st.x, st.y, st.z], 1)
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    n_types = int(np.asarray(c["tab"]["rho_r"]).shape[0])
    with _ctx(c) as ctx:
        ctx.rho_calc_aos(recs, RC)
        good = recs["rho"].copy()
        bad = recs.copy()
        bad["type"][occ[len(occ) // 2]] = n_types
        bad["rho"] = -3.0
        with pytest.raises(mmd.MmdError) as e:
            ctx.rho_calc_aos(bad, RC)
        assert e.value.code == mmd.E_INVALID
        assert np.all(bad["rho"] == -3.0)
        recs["rho"] = 0.0
        ctx.rho_calc_aos(recs, RC)
        assert np.array_equal(recs["rho"][occ], good[occ])


PIPE_CASES = {
    "ragged_7x9x5": CASES["ragged_7x9x5"],
    # 168-entry stencil: z reach 3, a 92 KB LDS tile
    "wide_stencil": dict(cells=(10, 9, 8), n_types=3, sigma=0.05, temperature=900.0, vacancy_frac=0.01, seed=5,
                         skin=1.9),
}


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
@pytest.mark.parametrize("name", list(PIPE_CASES))
def test_dropin_pinned_chunks_equal_resident_passes(name, prec):
    """Pinned records take the z-chunk pipeline (records copied in by
    chunks, the pass run over the tile rows whose halo has arrived, outputs
    written as their rows finish).  Its rho, df and f equal the
    device-resident passes' bit for bit, for the default stencil (z reach 2)
    and a 168-entry one (z reach 3), in both precisions."""
    c = W.case(**PIPE_CASES[name])
    st = c["st"]
    n = st.x.size
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    with _ctx(c, prec) as ctx:
        ctx.rho_df(RC)
        res = st.copy()
        ctx.download(res, mmd.F_RHO | mmd.F_DF)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_DF)
        ctx.force(RC)
        ctx.download(res, mmd.F_F | mmd.F_DF)
    with _PinnedRecords(n) as recs, _ctx(c, prec) as ctx:
        recs["id"] = np.arange(n)
        recs["type"] = st.type
        recs["x"] = np.stack([st.x, st.y, st.z], 1)
        recs["v"] = np.stack([st.vx, st.vy, st.vz], 1)
        ctx.rho_calc_aos(recs, RC)
        assert np.array_equal(recs["rho"][occ], res.rho[occ])
        assert np.array_equal(recs["df"][occ], res.df[occ])
        recs["df"] = res.df  # every ext site's df, as the core's ghost exchange leaves it
        ctx.force_calc_aos(recs, RC)
        assert np.array_equal(recs["f"][occ], np.stack([res.fx[occ], res.fy[occ], res.fz[occ]], 1))


def test_dropin_pinned_rejects_type_without_table():
    """Pinned records run the pass pipelined by z-chunks, so a type with no
    table is reported after the last chunk: the hook fails with
    MMD_E_INVALID, only output fields may have changed, and the next valid
    call computes as before."""
    c = W.case(**CASES["ragged_7x9x5"])
    st = c["st"]
    n = st.x.size
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    n_types = int(np.asarray(c["tab"]["rho_r"]).shape[0])
    with _PinnedRecords(n) as recs, _ctx(c) as ctx:
        recs["id"] = np.arange(n)
        recs["type"] = st.type
        recs["x"] = np.stack([st.x, st.y, st.z], 1)
        ctx.rho_calc_aos(recs, RC)
        good = recs["rho"].copy()
        recs["type"][occ[len(occ) // 2]] = n_types
        recs["rho"] = -3.0
        before = recs.copy()
        with pytest.raises(mmd.MmdError) as e:
            ctx.rho_calc_aos(recs, RC)
        assert e.value.code == mmd.E_INVALID
        for f in ("id", "type", "x", "v", "f"):
            assert np.array_equal(recs[f], before[f])
        recs["type"] = st.type
        recs["rho"] = 0.0
        ctx.rho_calc_aos(recs, RC)
        assert np.array_equal(recs["rho"][occ], good[occ])


def test_set_tables_rejects_pair_tables_beyond_record_limit():
    """Pair-table records are addressed with 24-bit multiply-adds: 8 types x
    262,144 r knots (2^24 records) is refused with MMD_E_RANGE before anything
    is built, and the context keeps working with valid tables."""
    c = W.case(**CASES["ragged_7x9x5"])
    nt, nr = 8, 1 << 18
    big = dict(n_types=nt, n_r=nr, r0=0.5, dr=1e-5, n_rho=5, rho0=0.0, drho=1.0, rho_r=np.zeros((nt, nr)),
               emb=np.zeros((nt, 5)), phi_r=np.zeros((nt, nt, nr)), rc=RC)
    with _ctx(c) as ctx:
        with pytest.raises(mmd.MmdError) as e:
            ctx.set_tables(big, (55.845,) * nt)
        assert e.value.code == mmd.E_RANGE
        ctx.set_tables(c["tab"], c["masses"])
        ctx.rho_df(RC)


def test_ghost_refresh_matches_oracle():
    c = W.case(**CASES["ragged_7x9x5"])
    st = c["st"].copy()
    li = st.local_index()
    scrambled = st.copy()
    mask = np.ones(st.x.size, bool)
    mask[li] = False
    scrambled.x[mask] = 99.0
    scrambled.type[mask] = -1
    with _ctx(c) as ctx:
        ctx.upload(scrambled, mmd.F_X | mmd.F_TYPE)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ctx.ghost_refresh(mmd.F_X)
        out = st.copy()
        ctx.download(out, mmd.F_X | mmd.F_TYPE)
    ref = scrambled.copy()
    O.ghost_xt(c["dom"], c["box"], ref)
    assert np.array_equal(out.type, ref.type)
    assert np.array_equal(out.x, ref.x) and np.array_equal(out.y, ref.y) and np.array_equal(out.z, ref.z)


def _prime(ctx, c, energy_every=0):
    ctx.md_setup(c["box"], A, 1.0, RC, energy_every=energy_every)
    ctx.md_prime()


def test_md_steps_match_oracle():
    c = W.case(cells=(10, 10, 10), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.01, seed=9)
    ref, T = _oracle_pass(c)
    nsteps = 10
    en_ref = O.md_run(c["dom"], c["even"], c["odd"], T, RC, c["box"], 1.0, nsteps, ref)
    with _ctx(c) as ctx:
        _prime(ctx, c, energy_every=1)
        en = np.zeros((nsteps, 2))
        ctx.md_run(nsteps, en)
        out = c["st"].copy()
        ctx.download(out)
    li = out.local_index()
    occ = li[out.type[li] >= 0]
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.abs(getattr(out, f)[occ] - getattr(ref, f)[occ]).max() < 1e-10, f
    assert np.allclose(en, en_ref, rtol=1e-12, atol=0)
    n_at = len(occ)
    assert np.abs(en - en_ref).max() / n_at < 1e-12


def test_graph_replay_equals_eager():
    c = W.case(cells=(8, 9, 10), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01, seed=12)
    outs = []
    for use_graph in (False, True):
        with _ctx(c) as ctx:
            _prime(ctx, c, energy_every=2)
            en = np.zeros((6, 2))
            ctx.md_run(3, en[:3], use_graph=use_graph)
            ctx.sync()
            ctx.md_run(3, en[3:], use_graph=use_graph)  # second call replays the captured graphs
            o = c["st"].copy()
            ctx.download(o)
            outs.append((o, en))
    (a, ea), (b, eb) = outs
    for f in ("x", "vx", "fx", "rho", "df"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(ea, eb)


def test_energy_entry_point_matches_oracle():
    c = W.case(**CASES["bcc12_3type_vac"])
    st = c["st"]
    T = O.Tables(c["tab"], c["masses"])
    ref = st.copy()
    e1 = O.rho_df(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
    O.ghost_df(c["dom"], ref)
    e2 = O.force(c["dom"], c["even"], c["odd"], T, RC, ref, want_e=True)
    ke = O.kinetic(c["dom"], T, ref)
    with _ctx(c) as ctx:
        ctx.md_setup(c["box"], A, 1.0, RC)
        ep, ek = ctx.energy()
    assert abs(ep - (e1 + e2)) < 1e-12 * abs(e1 + e2)
    assert abs(ek - ke) < 1e-12 * ke


def _force_err(F, Fr):
    d = np.linalg.norm(F - Fr, axis=1)
    rms = np.sqrt((np.linalg.norm(Fr, axis=1) ** 2).mean())
    return np.linalg.norm(F - Fr) / np.linalg.norm(Fr), d.max() / rms


def _fallback_fraction(c, ranges):
    """Share of the in-cutoff pairs (stencil pairs between occupied local
    sites and their neighbours) whose squared distance lies outside the mixed
    tables' ranges."""
    st = c["st"]
    li = st.local_index()
    i = li[st.type[li] >= 0]
    par = ((i % st.dom.ext[0]) - st.dom.ghost[0]) & 1
    offs = np.where(par[:, None] == 0, c["even"][None, :], c["odd"][None, :])
    j = i[:, None] + offs
    d2 = (st.x[i, None] - st.x[j]) ** 2 + (st.y[i, None] - st.y[j]) ** 2 + (st.z[i, None] - st.z[j]) ** 2
    s = d2[(d2 < RC * RC) & (st.type[j] >= 0)]
    (head, tail), holes = ranges
    inside = (s >= head) & (s < tail)
    for lo, hi in holes:
        inside &= ~((s >= lo) & (s < hi))
    return 1.0 - inside.mean()


@pytest.mark.parametrize("name", ["bcc12_3type_vac", "ragged_7x9x5", "bcc12_3type_distinct_rho"])
def test_mixed_precision_within_target(name):
    c = W.case(**CASES[name])
    res = {}
    for prec in (mmd.PREC_FP64, mmd.PREC_MIXED):
        with _ctx(c, prec=prec) as ctx:
            ctx.md_setup(c["box"], A, 1.0, RC)
            ctx.rho_df(RC)
            ctx.ghost_refresh(mmd.F_DF)
            ctx.force(RC)
            o = c["st"].copy()
            ctx.download(o, mmd.F_RHO | mmd.F_DF | mmd.F_F)
            res[prec] = (o, ctx.energy())
            ranges = ctx.mixed_ranges()
    a, b = res[mmd.PREC_FP64][0], res[mmd.PREC_MIXED][0]
    li = a.local_index()
    occ = li[a.type[li] >= 0]
    # the fast path covers the pairs: at most 2% of in-cutoff pairs fall back to FP64
    assert _fallback_fraction(c, ranges) < 0.02, ranges
    assert np.allclose(b.rho[occ], a.rho[occ], rtol=1e-9)
    Fa = np.stack([a.fx[occ], a.fy[occ], a.fz[occ]], 1)
    Fb = np.stack([b.fx[occ], b.fy[occ], b.fz[occ]], 1)
    normwise, per_atom = _force_err(Fb, Fa)
    assert 0.0 < normwise <= 1e-6 and per_atom <= 1e-6, (normwise, per_atom)
    (ea, _), (eb, _) = res[mmd.PREC_FP64][1], res[mmd.PREC_MIXED][1]
    assert abs(ea - eb) / len(occ) < 1e-9


def test_error_paths():
    c = W.case(cells=(4, 4, 4))
    ctx = mmd.Context(0)
    with pytest.raises(mmd.MmdError) as e:
        ctx.set_stencil(c["even"], c["odd"])
    assert e.value.code == mmd.E_STATE
    ctx.set_domain(c["dom"])
    with pytest.raises(mmd.MmdError) as e:
        ctx.rho_df(RC)
    assert e.value.code == mmd.E_STATE
    bad = c["even"].copy()
    bad[0] = 10 ** 7
    with pytest.raises(mmd.MmdError) as e:
        ctx.set_stencil(bad, c["odd"])
    assert e.value.code == mmd.E_RANGE
    ctx.set_stencil(c["even"], c["odd"])
    tab = dict(c["tab"])
    tab["phi_r"] = tab["phi_r"].copy()
    ctx.set_tables(tab, c["masses"])
    with pytest.raises(mmd.MmdError) as e:
        ctx.rho_df(7.0)  # beyond the tabulated range
    assert e.value.code == mmd.E_INVALID
    t3 = mmd.synth_tables(n_types=2, jitter=0.1)
    t3["phi_r"] = t3["phi_r"].copy()
    t3["phi_r"][0, 1, 10] += 1.0  # asymmetric pair table
    with pytest.raises(mmd.MmdError) as e:
        ctx.set_tables(t3, c["masses"])
    assert e.value.code == mmd.E_INVALID
    ctx.close()


def test_mixed_precision_close_contacts():
    """Colliding lattice atoms (pairs at 1.0-1.6 A, as in hot defect runs) stay
    within the mixed-precision target: the r^2 tables hand such pairs to the
    FP64 spline."""
    c = W.case(cells=(8, 8, 8), n_types=3, sigma=0.05, temperature=1200.0, vacancy_frac=0.01, seed=31)
    st = c["st"]
    li = st.local_index()
    occ = li[st.type[li] >= 0]
    rng = np.random.default_rng(5)
    for k, i in enumerate(rng.choice(occ, 12, replace=False)):
        j = i + c["even"][0] if ((i % c["dom"].ext[0]) % 2 == 0) else i + c["odd"][0]
        if st.type[j] < 0:
            continue
        d = np.array([st.x[j] - st.x[i], st.y[j] - st.y[i], st.z[j] - st.z[i]])
        target = 1.0 + 0.05 * k
        shift = d * (1.0 - target / np.linalg.norm(d))
        st.x[i] += shift[0]
        st.y[i] += shift[1]
        st.z[i] += shift[2]
    res = {}
    for prec in (mmd.PREC_FP64, mmd.PREC_MIXED):
        with _ctx(c, prec=prec) as ctx:
            ctx.md_setup(c["box"], A, 1.0, RC)
            ctx.ghost_refresh(mmd.F_X)
            ctx.rho_df(RC)
            ctx.ghost_refresh(mmd.F_DF)
            ctx.force(RC)
            o = st.copy()
            ctx.download(o, mmd.F_F)
            res[prec] = o
    a, b = res[mmd.PREC_FP64], res[mmd.PREC_MIXED]
    Fa = np.stack([a.fx[occ], a.fy[occ], a.fz[occ]], 1)
    Fb = np.stack([b.fx[occ], b.fy[occ], b.fz[occ]], 1)
    normwise, per_atom = _force_err(Fb, Fa)
    assert normwise <= 1e-6 and per_atom <= 1e-6, (normwise, per_atom)


@pytest.mark.parametrize("prec", [mmd.PREC_FP64, mmd.PREC_MIXED])
def test_golden_fs_lattices_on_device(prec):
    """Closed-form golden energies and densities (tests/golden/fs_bcc_lattice.json:
    direct analytic shell sums of the Finnis-Sinclair Fe functions) through the
    C ABI: FP64 to rounding; mixed energies and the rho of the standalone pass
    within the 1e-9 value bound, and the rho the MD calls leave (mmd_energy,
    whose dynamics take rho from the lean quadratic records) within their 2e-8
    bound."""
    tol = 1e-11 if prec == mmd.PREC_FP64 else 1e-9
    tol_md = 1e-11 if prec == mmd.PREC_FP64 else 2e-8
    for case in W.golden_fs_cases():
        c = W.strained_lattice(case["lattice_const"], case["strain"])
        with _ctx(c, prec=prec) as ctx:
            ctx.md_setup(c["box"], A, 1.0, RC)
            ep, _ = ctx.energy()
            out = c["st"].copy()
            ctx.download(out, mmd.F_RHO | mmd.F_F)
            ctx.rho_df(RC)
            std = c["st"].copy()
            ctx.download(std, mmd.F_RHO)
        li = out.local_index()
        n = li.size
        assert np.allclose(std.rho[li], case["rho"], rtol=tol, atol=0), case["name"]
        assert np.allclose(out.rho[li], case["rho"], rtol=tol_md, atol=0), case["name"]
        assert abs(ep / n - case["energy"]) < tol * abs(case["energy"]), (case["name"], ep / n)
        assert np.abs(np.stack([out.fx[li], out.fy[li], out.fz[li]])).max() < 1e3 * tol, case["name"]


def test_energy_with_asymmetric_stencil():
    """A stencil whose parity-1 half lacks an entry (not symmetric): the pair
    energy falls back to half of every visited pair, as the oracle sums it."""
    c = W.case(**CASES["bcc12_3type_vac"])
    odd = c["odd"][:-1].copy()  # drop the last parity-1 entry
    T = O.Tables(c["tab"], c["masses"])
    ref = c["st"].copy()
    e1 = O.rho_df(c["dom"], c["even"], odd, T, RC, ref, want_e=True)
    O.ghost_df(c["dom"], ref)
    e2 = O.force(c["dom"], c["even"], odd, T, RC, ref, want_e=True)
    with mmd.Context(0) as ctx:
        ctx.set_domain(c["dom"])
        ctx.set_stencil(c["even"], odd)
        ctx.set_tables(c["tab"], c["masses"])
        ctx.upload(c["st"], mmd.F_X | mmd.F_TYPE | mmd.F_V)
        ctx.md_setup(c["box"], A, 1.0, RC)
        ep, _ = ctx.energy()
    assert abs(ep - (e1 + e2)) < 1e-12 * abs(e1 + e2)
