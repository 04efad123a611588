"""Host logic of the library: synthetic tables and the initial-condition
generator (no device calls)."""
import numpy as np
import pytest

import mmd
import workload as W


def test_synth_tables_closed_forms_and_symmetry():
    t = mmd.synth_tables(n_types=3, jitter=0.1, seed=4)
    r = t["r0"] + t["dr"] * np.arange(t["n_r"])
    assert abs(t["dr"] - 5.1 / 4999) < 1e-15
    rc = 5.6
    assert np.allclose(t["rho_r"][0], (rc - r) ** 2 * np.exp(-r), rtol=1e-14, atol=1e-300)
    for a in range(3):
        for b in range(3):
            assert np.array_equal(t["phi_r"][a][b], t["phi_r"][b][a])
            ratio = t["phi_r"][a][b][:100] / ((rc - r[:100]) ** 2)
            # A (1 + B/r) with A, B within +-10% of 1.2, 0.3
            A_est = (ratio[0] * r[0] - ratio[99] * r[99]) / (r[0] - r[99])
            assert 1.08 - 1e-9 <= A_est <= 1.32 + 1e-9
    rho = t["rho0"] + t["drho"] * np.arange(t["n_rho"])
    t1 = mmd.synth_tables(n_types=1)
    assert np.allclose(t1["emb"][0], -np.sqrt(rho) + 0.01 * rho ** 2, rtol=1e-14, atol=1e-14)
    assert np.allclose(t1["phi_r"][0][0], 1.2 * (rc - r) ** 2 * (1 + 0.3 / r), rtol=1e-14, atol=1e-300)


def test_lattice_init_deterministic_and_statistics():
    c1 = W.case(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.005, seed=3)
    c2 = W.case(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.005, seed=3)
    c3 = W.case(cells=(12, 12, 12), n_types=3, sigma=0.05, temperature=800.0, vacancy_frac=0.005, seed=4)
    s1, s2, s3 = c1["st"], c2["st"], c3["st"]
    assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.type, s2.type) and np.array_equal(s1.vz, s2.vz)
    assert not np.array_equal(s1.x, s3.x)
    li = s1.local_index()
    t = s1.type[li]
    N = len(li)
    occ = li[t >= 0]
    frac = np.array([(t == k).sum() for k in range(3)]) / (t >= 0).sum()
    assert abs(frac[0] - 0.9) < 4 * np.sqrt(0.09 / N) and abs(frac[1] - 0.05) < 4 * np.sqrt(0.05 / N)
    vac = (t < 0).mean()
    assert abs(vac - 0.005) < 4 * np.sqrt(0.005 / N)
    # displacement sigma from the ideal sites
    d = c1["dom"]
    X = (li % d.ext[0]) - d.ghost[0]
    ideal_x = X * W.A / 2
    dx = s1.x[li] - ideal_x
    assert abs(np.std(dx) - 0.05) < 0.003
    # temperature from kinetic energy, COM momentum removed
    m = np.array(mmd.DEFAULT_MASSES)[s1.type[occ]]
    v2 = s1.vx[occ] ** 2 + s1.vy[occ] ** 2 + s1.vz[occ] ** 2
    ke = 0.5 * (m * v2).sum() / 9.64853321e-3
    T = 2 * ke / (3 * len(occ) * 8.617333262e-5)
    assert abs(T - 800.0) < 800.0 * 4 * np.sqrt(2 / (3 * len(occ)))
    P = np.array([(m * s1.vx[occ]).sum(), (m * s1.vy[occ]).sum(), (m * s1.vz[occ]).sum()])
    assert np.abs(P).max() < 1e-9


def test_interstitials_at_tetrahedral_sites():
    c = W.case(cells=(10, 10, 10), sigma=0.0, temperature=300.0, interstitial_frac=0.002, seed=8)
    itl = c["itl"]
    assert itl.n == round(0.002 * 2000)
    a = W.A
    for k in range(itl.n):
        p = np.array([itl.x[k], itl.y[k], itl.z[k]])
        f = (p / a) % 1.0
        vals = sorted(np.round(f * 4) / 4 % 1.0)
        assert vals == [0.0, 0.25, 0.5] or vals == [0.0, 0.5, 0.75]
    cells = np.floor(np.stack([itl.x, itl.y, itl.z], 1) / a).astype(int)
    for i in range(itl.n):
        for j in range(i + 1, itl.n):
            dd = (cells[i] - cells[j] + 5) % 10 - 5
            assert np.abs(dd).max() >= 2  # no two interstitials in neighbouring cells


def test_blob_defects_cluster():
    c = W.case(cells=(24, 24, 24), sigma=0.0, temperature=0.0, vacancy_frac=0.01, interstitial_frac=0.001,
               n_blobs=2, seed=5)
    st = c["st"]
    li = st.local_index()
    vac = li[st.type[li] < 0]
    assert len(vac) == round(0.01 * 2 * 24 ** 3)
    assert c["itl"].n == round(0.001 * 2 * 24 ** 3)


def test_lattice_init_rejects_bad_specs():
    s = mmd.lattice_spec((4, 4, 4), ghost_cells=5)
    with pytest.raises(mmd.MmdError):
        mmd.lattice_init(s)
    s = mmd.lattice_spec((4, 4, 4), vacancy_frac=0.9)
    with pytest.raises(mmd.MmdError):
        mmd.lattice_init(s)
