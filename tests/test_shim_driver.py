"""The drop-in boundary end to end from C++: tests/mini_misa/driver.cpp is a
stand-in for MISA-MD core's EAM step that calls the real hook file
(misa-md-hip_amd/shim/misa_md_hooks.cpp, the reference's exact hip_*
signatures) on pinned 104-byte atom records, with the potential read from a
setfl file through MMD_SETFL.  Its rho, df and forces must match the oracle on
the same records and tables."""
import os
import subprocess

import numpy as np
import pytest

import mmd
import oracle as O
import workload as W

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "mini_misa", "mini_misa")


def _driver():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.dirname(DRIVER)], check=True)
    return DRIVER


def test_driver_links_the_library():
    out = subprocess.run(["ldd", _driver()], capture_output=True, text=True, check=True).stdout
    line = [s for s in out.splitlines() if "libmmd_eam.so" in s]
    assert line and "not found" not in line[0], out


def _write_inputs(d, c):
    dom, st = c["dom"], c["st"]
    d.mkdir()
    vals = list(dom.ghost) + list(dom.box) + list(dom.ext) + list(dom.start)
    (d / "domain.txt").write_text(" ".join(str(int(v)) for v in vals) + f" {W.RC!r} {len(c['even'])} {len(c['odd'])}\n")
    np.asarray(c["even"], np.int64).tofile(d / "even.bin")
    np.asarray(c["odd"], np.int64).tofile(d / "odd.bin")
    st.type.astype(np.int32).tofile(d / "type.bin")
    np.stack([st.x, st.y, st.z], 1).tofile(d / "x.bin")
    W.ghost_sources(dom).tofile(d / "ghost_src.bin")


@pytest.mark.gpu
def test_mini_misa_hook_sequence_matches_oracle(tmp_path):
    c = W.case(cells=(10, 9, 8), n_types=2, sigma=0.05, temperature=800.0, vacancy_frac=0.01, seed=5)
    setfl = tmp_path / "fs2.setfl"
    W.write_fs_setfl(setfl, n_types=2)
    tab, masses = mmd.setfl_read(setfl)
    _write_inputs(tmp_path / "in", c)
    out = tmp_path / "out.bin"
    env = dict(os.environ, MMD_SETFL=str(setfl), MMD_DEVICE="0")
    print("mini_misa: 10x9x8 cells, 2 types, two steps through the hook file", flush=True)
    r = subprocess.run([_driver(), str(tmp_path / "in"), str(out), "2"], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    res = np.fromfile(out).reshape(-1, 5)
    ref = c["st"].copy()
    T = O.Tables(tab, masses)
    O.rho_df(c["dom"], c["even"], c["odd"], T, W.RC, ref)
    O.ghost_df(c["dom"], ref)
    O.force(c["dom"], c["even"], c["odd"], T, W.RC, ref)
    li = ref.local_index()
    occ = li[ref.type[li] >= 0]
    assert np.allclose(res[occ, 0], ref.rho[occ], rtol=1e-12, atol=0)
    assert np.allclose(res[occ, 1], ref.df[occ], rtol=1e-12, atol=1e-15)
    F = res[occ, 2:5]
    Fr = np.stack([ref.fx[occ], ref.fy[occ], ref.fz[occ]], 1)
    assert np.abs(F - Fr).max() <= 1e-11 * np.abs(Fr).max()
    assert np.abs(Fr).max() > 1e-3  # a thermal state, not a trivially zero force field
