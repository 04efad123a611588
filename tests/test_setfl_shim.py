"""setfl reader and the MISA-MD hook adapter (CPU).

The setfl file is written here from the closed forms phi(r) = A (rc-r)^2 (1+B/r),
rho(r) = (rc-r)^2 exp(-r), F(rho) = -sqrt(rho) + C rho^2 (values printed with 17
significant digits, so they read back exactly); r*phi is what the file holds.
The adapter is compiled with g++ against stand-in types carrying the member
names the reference plugin reads, and must reproduce the reference's domain
mapping (src/hip_arch_acc_imp.cpp:61-72) and _cuAtomElement byte layout
(src/kernel_types.h:40-48).
"""
import os
import subprocess

import numpy as np
import pytest

import mmd
from workload import HARTREE_BOHR, write_funcfl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RC = 5.6


def write_setfl(path, nt=2, nr=600, dr=0.01, nrho=500, drho=0.05, params=None):
    params = params or [(1.2 + 0.1 * t, 0.3 - 0.05 * t, 0.01 + 0.002 * t) for t in range(nt)]
    r = np.arange(nr) * dr
    rho = np.arange(nrho) * drho
    lines = ["synthetic EAM", "closed forms", "test", f"{nt} " + " ".join(f"E{t}" for t in range(nt)),
             f"{nrho} {drho!r} {nr} {dr!r} {RC!r}"]
    vals = {}

    def emit(a):
        for k in range(0, len(a), 5):
            lines.append(" ".join("%.17g" % v for v in a[k:k + 5]))

    for t in range(nt):
        A, B, C = params[t]
        lines.append(f"{26 + t} {55.845 + t} 2.855 bcc")
        F = -np.sqrt(rho) + C * rho ** 2
        rr = np.where(r < RC, (RC - r) ** 2 * np.exp(-r), 0.0)
        vals[("F", t)], vals[("rho", t)] = F, rr
        emit(F)
        emit(rr)
    for i in range(nt):
        for j in range(i + 1):
            A = 0.5 * (params[i][0] + params[j][0])
            B = 0.5 * (params[i][1] + params[j][1])
            z = np.where(r < RC, A * (RC - r) ** 2 * (r + B), 0.0)  # r * phi
            vals[("z", i, j)] = z
            emit(z)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return vals, r


def test_funcfl_read_values(tmp_path):
    """One file: F and rho as written, phi = 27.2 * 0.529 * Z^2 / r; two files:
    the cross term from the product of the two charges."""
    a, b = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    va = write_funcfl(a)
    vb = write_funcfl(b, z_scale=1.3, C=0.02, mass=58.693)
    rk = np.arange(1, 600) * 0.01
    tab, mass = mmd.funcfl_read([a])
    assert (tab["n_types"], tab["n_r"], tab["n_rho"]) == (1, 600, 500) and mass == (55.845,)
    assert tab["r0"] == 0.0 and tab["dr"] == 0.01 and tab["drho"] == 0.05 and tab["rc"] == RC
    assert np.array_equal(tab["emb"][0], va["F"]) and np.array_equal(tab["rho_r"][0], va["rho"])
    phi = tab["phi_r"][0, 0]
    assert np.array_equal(phi[1:], HARTREE_BOHR * va["Z"][1:] * va["Z"][1:] / rk)
    assert phi[0] == 2 * phi[1] - phi[2]
    tab, mass = mmd.funcfl_read([a, b])
    assert tab["n_types"] == 2 and mass == (55.845, 58.693)
    assert np.array_equal(tab["emb"][1], vb["F"]) and np.array_equal(tab["rho_r"][1], vb["rho"])
    for i, zi in enumerate((va["Z"], vb["Z"])):
        for j, zj in enumerate((va["Z"], vb["Z"])):
            z2r = HARTREE_BOHR * zi * zj if i >= j else HARTREE_BOHR * zj * zi
            assert np.array_equal(tab["phi_r"][i, j][1:], z2r[1:] / rk), (i, j)
    # a physical check of the conversion: at r = 1 A a unit charge product
    # gives 27.2 * 0.529 = 14.39 eV, the Coulomb energy of two elementary
    # charges one angstrom apart (14.40 eV) to the constants' rounding
    assert abs(HARTREE_BOHR - 14.39964) < 0.02


def test_funcfl_errors(tmp_path):
    a, b = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    write_funcfl(a)
    write_funcfl(b, dr=0.011)  # another grid
    txt = a.read_text().splitlines()
    (tmp_path / "trunc.funcfl").write_text("\n".join(txt[: len(txt) // 2]) + "\n")
    (tmp_path / "grid.funcfl").write_text("\n".join(txt[:2] + ["500 0.05 600 0 5.6"] + txt[3:]) + "\n")
    for files in ([a, b], [tmp_path / "trunc.funcfl"], [tmp_path / "grid.funcfl"], [tmp_path / "missing.funcfl"]):
        with pytest.raises(mmd.MmdError) as e:
            mmd.funcfl_read(files)
        assert e.value.code == mmd.E_INVALID


def test_setfl_read_values(tmp_path):
    p = tmp_path / "a.setfl"
    vals, r = write_setfl(p)
    tab, mass = mmd.setfl_read(p)
    assert (tab["n_types"], tab["n_r"], tab["n_rho"]) == (2, 600, 500)
    assert tab["r0"] == 0.0 and tab["dr"] == 0.01 and tab["drho"] == 0.05 and tab["rc"] == RC
    assert mass == (55.845, 56.845)
    for t in range(2):
        assert np.array_equal(tab["emb"][t], vals[("F", t)])
        assert np.array_equal(tab["rho_r"][t], vals[("rho", t)])
    for i in range(2):
        for j in range(i + 1):
            z = vals[("z", i, j)]
            phi = tab["phi_r"][i, j]
            assert np.array_equal(phi[1:], z[1:] / (np.arange(1, 600) * 0.01))
            assert phi[0] == 2 * phi[1] - phi[2]
            assert np.array_equal(tab["phi_r"][j, i], phi)


def test_setfl_errors(tmp_path):
    p = tmp_path / "bad.setfl"
    write_setfl(p)
    txt = p.read_text().splitlines()
    (tmp_path / "trunc.setfl").write_text("\n".join(txt[: len(txt) // 2]) + "\n")
    (tmp_path / "grid.setfl").write_text("\n".join(txt[:4] + ["500 0.05 600 -0.01 5.6"] + txt[5:]) + "\n")
    for name in ("trunc.setfl", "grid.setfl", "missing.setfl"):
        with pytest.raises(mmd.MmdError) as e:
            mmd.setfl_read(tmp_path / name)
        assert e.value.code == mmd.E_INVALID


def test_hook_adapter_conversions(tmp_path):
    exe = tmp_path / "probe"
    lib = os.path.join(ROOT, "misa-md-hip_amd", "mmd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "misa-md-hip_amd", "shim"), os.path.join(ROOT, "tests", "shim_probe.cpp"),
                    "-o", str(exe), "-L" + lib, "-lmmd_eam", "-Wl,-rpath," + lib, "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True)
    p = tmp_path / "a.setfl"
    write_setfl(p)
    fa, fb = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    write_funcfl(fa)
    write_funcfl(fb, z_scale=1.3, mass=58.693)
    out = subprocess.run([str(exe), str(p), f"{fa}:{fb}"], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    # ghost = local low - ghost low; box, ext as given; start = local low
    assert out[0] == "domain 6 3 3 | 100 50 60 | 112 56 66 | 40 10 20"
    assert out[1] == "stencil 4 2 -5000 -2"
    # _cuAtomElement: id@0 type@8 x@16 v@40 f@64 rho@88 df@96, 104 bytes, 4-byte type
    assert out[2] == "layout 104 8 16 40 64 88 96 4"
    tab, _ = mmd.setfl_read(p)
    f = out[3].split()
    assert f[:4] == ["tables", "2", "600", "500"]
    assert float(f[6]) == tab["phi_r"].reshape(-1)[600 + 7]
    f = out[4].split()
    assert f[:4] == ["funcfl", "2", "600", "500"]
    tab2, _ = mmd.funcfl_read([fa, fb])
    assert float(f[5]) == tab2["phi_r"].reshape(-1)[600 + 7]
    assert float(f[6]) == 58.693
