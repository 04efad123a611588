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
    out = subprocess.run([str(exe), str(p)], check=True, capture_output=True, text=True).stdout.splitlines()
    # ghost = local low - ghost low; box, ext as given; start = local low
    assert out[0] == "domain 6 3 3 | 100 50 60 | 112 56 66 | 40 10 20"
    assert out[1] == "stencil 4 2 -5000 -2"
    # _cuAtomElement: id@0 type@8 x@16 v@40 f@64 rho@88 df@96, 104 bytes, 4-byte type
    assert out[2] == "layout 104 8 16 40 64 88 96 4"
    tab, _ = mmd.setfl_read(p)
    f = out[3].split()
    assert f[:4] == ["tables", "2", "600", "500"]
    assert float(f[6]) == tab["phi_r"].reshape(-1)[600 + 7]
