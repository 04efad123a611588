"""ctypes binding of the CPU oracle (oracle/liborc_eam.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product never imports it.
Arrays use the reference plugin's layout (ghost-extended, x doubled).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc_eam.so")

_d = C.POINTER(C.c_double)
_i8 = C.POINTER(C.c_int8)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class OrcDomain(C.Structure):
    _fields_ = [("ghost", C.c_int32 * 3), ("box", C.c_int32 * 3), ("ext", C.c_int32 * 3),
                ("start", C.c_int64 * 3)]


class OrcSpline(C.Structure):
    _fields_ = [("n", C.c_int32), ("x0", C.c_double), ("dx", C.c_double), ("inv_dx", C.c_double),
                ("c", _d)]


class OrcTables(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("rho", C.POINTER(OrcSpline)), ("emb", C.POINTER(OrcSpline)),
                ("phi", C.POINTER(OrcSpline)), ("mass", _d)]


class OrcItl(C.Structure):
    _fields_ = [("n", C.c_int32)] + [(f, _d) for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz",
                                                         "rho", "df")] + [("type", _i8)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    D, T = C.POINTER(OrcDomain), C.POINTER(OrcTables)
    sig = {
        "orc_spline_build": (None, [_d, C.c_int32, C.c_double, _d]),
        "orc_spline_value": (C.c_double, [C.POINTER(OrcSpline), C.c_double]),
        "orc_spline_deriv": (C.c_double, [C.POINTER(OrcSpline), C.c_double]),
        "orc_stencil3": (C.c_int32, [C.c_double, C.c_double, C.c_int32, _i32, C.c_int32]),
        "orc_rho_df": (None, [D, _i64, C.c_int32, _i64, C.c_int32, T, C.c_double, _d, _d, _d, _i8, _d, _d, _d,
                              C.c_int32]),
        "orc_df": (None, [D, T, _i8, _d, _d]),
        "orc_force": (None, [D, _i64, C.c_int32, _i64, C.c_int32, T, C.c_double, _d, _d, _d, _i8, _d, _d, _d,
                             _d, _d, C.c_int32]),
        "orc_ghost_periodic_xt": (None, [D, _d, _d, _d, _d, _i8]),
        "orc_ghost_periodic_df": (None, [D, _d]),
        "orc_md_run": (None, [D, _i64, C.c_int32, _i64, C.c_int32, T, C.c_double, _d, C.c_double, C.c_int32,
                              _d, _d, _d, _d, _d, _d, _i8, _d, _d, _d, _d, _d, _d, C.c_int32]),
        "orc_kinetic": (C.c_double, [D, T, _d, _d, _d, _i8]),
        "orc_anchor": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _i64, _i64, _i64]),
        "orc_rho_df_itl": (None, [D, _i64, C.c_int32, _i64, C.c_int32, _i32, C.c_int32, C.c_double, _d, T,
                                  C.c_double, _d, _d, _d, _i8, _d, _d, C.POINTER(OrcItl), _d, C.c_int32]),
        "orc_force_itl": (None, [D, _i64, C.c_int32, _i64, C.c_int32, _i32, C.c_int32, C.c_double, _d, T,
                                 C.c_double, _d, _d, _d, _i8, _d, _d, _d, _d, C.POINTER(OrcItl), _d,
                                 C.c_int32]),
        "orc_md_run_itl": (None, [D, _i64, C.c_int32, _i64, C.c_int32, _i32, C.c_int32, C.c_double, T, C.c_double,
                                  _d, C.c_double, C.c_int32, _d, _d, _d, _d, _d, _d, _i8, _d, _d, _d, _d, _d,
                                  C.POINTER(OrcItl), _d, C.c_int32]),
        "orc_migrate": (C.c_int32, [D, C.c_double, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i8,
                                    C.POINTER(OrcItl), C.c_int32, _i32, _i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a, ct=C.c_double):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def domain(d):
    o = OrcDomain()
    for k in range(3):
        o.ghost[k], o.box[k], o.ext[k], o.start[k] = d.ghost[k], d.box[k], d.ext[k], d.start[k]
    return o


def spline_coeffs(f, dx):
    f = np.ascontiguousarray(f, float)
    c = np.zeros((len(f), 7))
    lib().orc_spline_build(_p(f), len(f), dx, _p(c))
    return c


class Tables:
    """Oracle spline tables built from knot values (same inputs the product takes)."""

    def __init__(self, tab, masses):
        nt = tab["n_types"]
        self.keep = []

        def mk(vals, x0, dx):
            c = spline_coeffs(vals, dx)
            self.keep.append(c)
            return OrcSpline(len(vals), x0, dx, 1.0 / dx, _p(c))

        self.rho = (OrcSpline * nt)(*[mk(tab["rho_r"][t], tab["r0"], tab["dr"]) for t in range(nt)])
        self.emb = (OrcSpline * nt)(*[mk(tab["emb"][t], tab["rho0"], tab["drho"]) for t in range(nt)])
        self.phi = (OrcSpline * (nt * nt))(*[mk(tab["phi_r"][a][b], tab["r0"], tab["dr"])
                                             for a in range(nt) for b in range(nt)])
        self.mass = np.ascontiguousarray(masses[:nt], float)
        self.s = OrcTables(nt, self.rho, self.emb, self.phi, _p(self.mass))

    def value(self, kind, idx, x):
        arr = {"rho": self.rho, "emb": self.emb, "phi": self.phi}[kind]
        return lib().orc_spline_value(C.byref(arr[idx]), x)

    def deriv(self, kind, idx, x):
        arr = {"rho": self.rho, "emb": self.emb, "phi": self.phi}[kind]
        return lib().orc_spline_deriv(C.byref(arr[idx]), x)


def stencil3(a, radius, parity, cap=8192):
    out = np.zeros((cap, 3), np.int32)
    n = lib().orc_stencil3(a, radius, parity, _p(out, C.c_int32), cap)
    assert n >= 0
    return out[:n].copy()


def linear(dom, st3):
    st3 = np.asarray(st3, np.int64)
    return np.ascontiguousarray((st3[:, 2] * dom.ext[1] + st3[:, 1]) * dom.ext[0] + st3[:, 0])


def rho_df(dom, even, odd, tabs, rc, st, nthreads=0, want_e=False):
    e = C.c_double(0.0)
    lib().orc_rho_df(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                     C.byref(tabs.s), rc, _p(st.x), _p(st.y), _p(st.z), _p(st.type, C.c_int8), _p(st.rho),
                     _p(st.df), C.byref(e) if want_e else None, nthreads)
    return e.value


def df(dom, tabs, st):
    lib().orc_df(C.byref(domain(dom)), C.byref(tabs.s), _p(st.type, C.c_int8), _p(st.rho), _p(st.df))


def force(dom, even, odd, tabs, rc, st, nthreads=0, want_e=False):
    e = C.c_double(0.0)
    lib().orc_force(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                    C.byref(tabs.s), rc, _p(st.x), _p(st.y), _p(st.z), _p(st.type, C.c_int8), _p(st.df),
                    _p(st.fx), _p(st.fy), _p(st.fz), C.byref(e) if want_e else None, nthreads)
    return e.value


def ghost_xt(dom, box_len, st):
    L = np.ascontiguousarray(box_len, float)
    lib().orc_ghost_periodic_xt(C.byref(domain(dom)), _p(L), _p(st.x), _p(st.y), _p(st.z),
                                _p(st.type, C.c_int8))


def ghost_df(dom, st):
    lib().orc_ghost_periodic_df(C.byref(domain(dom)), _p(st.df))


def md_run(dom, even, odd, tabs, rc, box_len, dt, nsteps, st, nthreads=0, energies=True):
    L = np.ascontiguousarray(box_len, float)
    en = np.zeros((nsteps, 2)) if energies else None
    lib().orc_md_run(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                     C.byref(tabs.s), rc, _p(L), dt, nsteps, _p(st.x), _p(st.y), _p(st.z), _p(st.vx), _p(st.vy),
                     _p(st.vz), _p(st.type, C.c_int8), _p(st.rho), _p(st.df), _p(st.fx), _p(st.fy), _p(st.fz),
                     _p(en) if energies else None, nthreads)
    return en


def kinetic(dom, tabs, st):
    return lib().orc_kinetic(C.byref(domain(dom)), C.byref(tabs.s), _p(st.vx), _p(st.vy), _p(st.vz),
                             _p(st.type, C.c_int8))


def anchor(a, p):
    X, Y, Z = C.c_int64(), C.c_int64(), C.c_int64()
    lib().orc_anchor(a, float(p[0]), float(p[1]), float(p[2]), C.byref(X), C.byref(Y), C.byref(Z))
    return X.value, Y.value, Z.value


def _itl_struct(itl):
    s = OrcItl()
    s.n = itl.n
    for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df"):
        setattr(s, f, _p(getattr(itl, f)))
    s.type = _p(itl.type, C.c_int8)
    return s


def rho_df_itl(dom, even, odd, wide3, a, box_len, tabs, rc, st, itl, nthreads=0, want_e=False):
    e = C.c_double(0.0)
    L = np.ascontiguousarray(box_len, float)
    w = np.ascontiguousarray(wide3, np.int32)
    s = _itl_struct(itl)
    lib().orc_rho_df_itl(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                         _p(w, C.c_int32), len(w), a, _p(L), C.byref(tabs.s), rc, _p(st.x), _p(st.y), _p(st.z),
                         _p(st.type, C.c_int8), _p(st.rho), _p(st.df), C.byref(s), C.byref(e) if want_e else None,
                         nthreads)
    return e.value


def force_itl(dom, even, odd, wide3, a, box_len, tabs, rc, st, itl, nthreads=0, want_e=False):
    e = C.c_double(0.0)
    L = np.ascontiguousarray(box_len, float)
    w = np.ascontiguousarray(wide3, np.int32)
    s = _itl_struct(itl)
    lib().orc_force_itl(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                        _p(w, C.c_int32), len(w), a, _p(L), C.byref(tabs.s), rc, _p(st.x), _p(st.y), _p(st.z),
                        _p(st.type, C.c_int8), _p(st.df), _p(st.fx), _p(st.fy), _p(st.fz), C.byref(s),
                        C.byref(e) if want_e else None, nthreads)
    return e.value


def md_run_itl(dom, even, odd, wide3, a, tabs, rc, box_len, dt, nsteps, st, itl, nthreads=0, energies=True):
    L = np.ascontiguousarray(box_len, float)
    w = np.ascontiguousarray(wide3, np.int32)
    s = _itl_struct(itl)
    en = np.zeros((nsteps, 2)) if energies else None
    lib().orc_md_run_itl(C.byref(domain(dom)), _p(even, C.c_int64), len(even), _p(odd, C.c_int64), len(odd),
                         _p(w, C.c_int32), len(w), a, C.byref(tabs.s), rc, _p(L), dt, nsteps, _p(st.x), _p(st.y),
                         _p(st.z), _p(st.vx), _p(st.vy), _p(st.vz), _p(st.type, C.c_int8), _p(st.rho), _p(st.df),
                         _p(st.fx), _p(st.fy), _p(st.fz), C.byref(s), _p(en) if energies else None, nthreads)
    return en


def migrate(dom, a, box_len, st, itl, cap):
    """Wigner-Seitz migration (orc_migrate) on the host state st (changed in
    place) and the list itl; returns (new list, leavers, fills).  Raises
    OverflowError, with nothing changed, when the list would exceed cap."""
    work = type(itl)(max(cap, 1))
    for f in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df", "type"):
        getattr(work, f)[: itl.n] = getattr(itl, f)[: itl.n]
    s = _itl_struct(work)
    s.n = itl.n
    L = np.ascontiguousarray(box_len, float)
    out, fil = C.c_int32(0), C.c_int32(0)
    rc = lib().orc_migrate(C.byref(domain(dom)), a, _p(L), _p(st.x), _p(st.y), _p(st.z), _p(st.vx), _p(st.vy),
                           _p(st.vz), _p(st.fx), _p(st.fy), _p(st.fz), _p(st.rho), _p(st.df),
                           _p(st.type, C.c_int8), C.byref(s), cap, C.byref(out), C.byref(fil))
    if rc == -1:
        raise OverflowError("interstitial list would exceed its capacity")
    if rc:
        raise ValueError("orc_migrate: domain with start != 0")
    return work.take(s.n), out.value, fil.value
