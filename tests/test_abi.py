"""The C-ABI library loads and exports every entry point include/*.h declares
(no device calls: this runs without a GPU)."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import mmd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\b(mmd_[a-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_hook_replacements():
    d = _declared()
    for n in ("mmd_ctx_create", "mmd_set_domain", "mmd_set_stencil", "mmd_set_tables", "mmd_eam_rho_df",
              "mmd_eam_df", "mmd_eam_force", "mmd_eam_rho_calc_aos", "mmd_eam_force_calc_aos", "mmd_host_alloc"):
        assert n in d


def test_library_exports_every_declared_symbol():
    L = mmd.lib()
    missing = [n for n in sorted(_declared()) if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mmd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mmd_[a-z0-9_]+)", out))
    assert _declared() <= exported


def test_library_is_built_for_gfx950_only():
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                          "--input=" + mmd.LIB_PATH], capture_output=True, text=True)
    if out.returncode == 0 and out.stdout.strip():
        targets = [t for t in out.stdout.split() if "amdgcn" in t]
        assert targets and all("gfx950" in t for t in targets)


def test_errors_are_status_codes_not_exceptions():
    # host-only entry points report bad arguments through status + message
    import ctypes as C
    L = mmd.lib()
    rc = L.mmd_stencil_linear(None, None, 1, None)
    assert rc == mmd.E_INVALID
    assert b"mmd_stencil_linear" in L.mmd_last_error()
    assert L.mmd_host_alloc(None, 8) == mmd.E_INVALID
    assert L.mmd_ctx_destroy(None) == mmd.OK
    h = C.c_void_p()
    # device count on a host without a GPU is 0 or an error status, never a crash
    rc = L.mmd_ctx_create(99, C.byref(h))
    assert rc in (mmd.E_INVALID, mmd.E_HIP)


def test_md_run_refuses_short_energy_arrays():
    """The MD runs write a [nsteps][2] energy array row by step: a smaller one
    is refused on the host before any library call (no context needed)."""
    ctx = mmd.Context.__new__(mmd.Context)
    for bad in (np.zeros((3, 2)), np.zeros((6, 2), np.float32), np.zeros((2, 6)).T):
        with pytest.raises(ValueError):
            ctx.md_run(6, bad)
        with pytest.raises(ValueError):
            ctx.md_run_noalloc(6, bad)
