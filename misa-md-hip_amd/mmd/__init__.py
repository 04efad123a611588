"""ctypes binding of libmmd_eam.so (the C ABI declared in include/mmd_eam.h).

This is plumbing for tests and benchmarks: every computation runs in the
native library.  Loading fails loudly when the library has not been built.
Host arrays use the reference plugin's layout: one array over the
ghost-extended lattice, x doubled, linear index ((z*ext_y)+y)*ext_x+x.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmmd_eam.so")

OK, E_INVALID, E_STATE, E_HIP, E_NOMEM, E_RANGE, E_NONFINITE = range(7)
F_X, F_V, F_F, F_RHO, F_DF, F_TYPE = 0x01, 0x02, 0x04, 0x08, 0x10, 0x20
F_ALL = 0x3F
PREC_FP64, PREC_MIXED = 0, 1
K_RHO, K_FORCE, K_INTEGRATE, K_GHOST, K_ITL, K_COUNT = range(6)

_d = C.POINTER(C.c_double)
_i8 = C.POINTER(C.c_int8)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class MmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mmd status {code}: {msg}")
        self.code = code


class Domain(C.Structure):
    _fields_ = [("ghost", C.c_int32 * 3), ("box", C.c_int32 * 3), ("ext", C.c_int32 * 3),
                ("start", C.c_int64 * 3)]

    def n_ext(self):
        return int(self.ext[0]) * int(self.ext[1]) * int(self.ext[2])

    def as_dict(self):
        return {k: list(getattr(self, k)) for k in ("ghost", "box", "ext", "start")}


class EamTables(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("n_r", C.c_int32), ("r0", C.c_double), ("dr", C.c_double),
                ("n_rho", C.c_int32), ("rho0", C.c_double), ("drho", C.c_double),
                ("rho_r", _d), ("emb", _d), ("phi_r", _d), ("mass", _d)]


class HostSoa(C.Structure):
    _fields_ = [(n, _d) for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df")] + \
               [("type", _i8)]


class AosLayout(C.Structure):
    _fields_ = [("stride", C.c_size_t), ("off_type", C.c_size_t), ("off_x", C.c_size_t),
                ("off_v", C.c_size_t), ("off_f", C.c_size_t), ("off_rho", C.c_size_t),
                ("off_df", C.c_size_t), ("type_bytes", C.c_int32)]


class MdParams(C.Structure):
    _fields_ = [("box_len", C.c_double * 3), ("lattice_const", C.c_double), ("dt", C.c_double),
                ("cutoff", C.c_double), ("energy_every", C.c_int32)]


class SynthParams(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("rc", C.c_double), ("A", C.c_double), ("B", C.c_double),
                ("C", C.c_double), ("jitter", C.c_double), ("seed", C.c_uint64), ("n_r", C.c_int32),
                ("r0", C.c_double), ("n_rho", C.c_int32), ("rho_max", C.c_double)]


class FsParams(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("rc", C.c_double), ("jitter", C.c_double), ("seed", C.c_uint64),
                ("n_r", C.c_int32), ("r0", C.c_double), ("n_rho", C.c_int32), ("rho_max", C.c_double)]


class LatticeSpec(C.Structure):
    _fields_ = [("cells", C.c_int32 * 3), ("ghost_cells", C.c_int32), ("lattice_const", C.c_double),
                ("n_types", C.c_int32), ("type_cum", C.c_double * 8), ("mass", C.c_double * 8),
                ("sigma", C.c_double), ("temperature", C.c_double), ("vacancy_frac", C.c_double),
                ("interstitial_frac", C.c_double), ("n_blobs", C.c_int32), ("blob_sigma", C.c_double),
                ("seed", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("precision", C.c_int32), ("lattice_const", C.c_double),
                ("cutoff", C.c_double), ("skin", C.c_double), ("dt", C.c_double), ("ghost_cells", C.c_int32),
                ("type_cum", C.c_double * 8), ("sigma", C.c_double), ("max_replicas", C.c_int32),
                ("hbm_fraction", C.c_double), ("hbm_limit", C.c_size_t), ("chunk_steps", C.c_int32),
                ("workers", C.c_int32), ("use_graph", C.c_int32)]


class ReplicaSpec(C.Structure):
    _fields_ = [("id", C.c_uint64), ("cells", C.c_int32), ("temperature", C.c_double), ("steps", C.c_int32),
                ("seed", C.c_uint64), ("arrival_s", C.c_double), ("keep", C.POINTER(HostSoa))]


class ReplicaResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("n_atoms", C.c_int32), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_ready", C.c_double), ("t_done", C.c_double),
                ("epot0", C.c_double), ("ekin0", C.c_double), ("epot", C.c_double), ("ekin", C.c_double),
                ("slot", C.c_int32), ("arena_offset", C.c_uint64), ("arena_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("atom_steps", C.c_double), ("atom_steps_per_s", C.c_double),
                ("latency_p50_s", C.c_double), ("latency_p99_s", C.c_double), ("latency_mean_s", C.c_double),
                ("wait_mean_s", C.c_double), ("completed", C.c_int32), ("failed", C.c_int32),
                ("max_active", C.c_int32), ("batched_steps", C.c_int64), ("arena_bytes", C.c_double),
                ("graph_captures", C.c_int64)]


class SimParams(C.Structure):
    _fields_ = [("arena_bytes", C.c_size_t), ("max_replicas", C.c_int32), ("chunk_steps", C.c_int32),
                ("ghost_cells", C.c_int32), ("workers", C.c_int32), ("step_overhead_s", C.c_double),
                ("step_atom_s", C.c_double), ("load_atom_s", C.c_double)]


# reference record _cuAtomElement (src/kernel_types.h:40-48)
REF_AOS_DTYPE = np.dtype({"names": ["id", "type", "x", "v", "f", "rho", "df"],
                          "formats": [np.int64, np.int32, (np.float64, 3), (np.float64, 3), (np.float64, 3),
                                      np.float64, np.float64],
                          "offsets": [0, 8, 16, 40, 64, 88, 96], "itemsize": 104})

_lib = None


def lib():
    """Loads libmmd_eam.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "mmd_last_error": (C.c_char_p, []),
        "mmd_ctx_create": (C.c_int, [C.c_int32, C.POINTER(P)]),
        "mmd_ctx_destroy": (C.c_int, [P]),
        "mmd_device_count": (C.c_int, [_i32]),
        "mmd_set_domain": (C.c_int, [P, C.POINTER(Domain)]),
        "mmd_set_stencil": (C.c_int, [P, _i64, C.c_int32, _i64, C.c_int32]),
        "mmd_set_tables": (C.c_int, [P, C.POINTER(EamTables)]),
        "mmd_set_precision": (C.c_int, [P, C.c_int32]),
        "mmd_upload": (C.c_int, [P, C.c_uint32, C.POINTER(HostSoa)]),
        "mmd_download": (C.c_int, [P, C.c_uint32, C.POINTER(HostSoa)]),
        "mmd_eam_rho_df": (C.c_int, [P, C.c_double]),
        "mmd_eam_df": (C.c_int, [P]),
        "mmd_eam_force": (C.c_int, [P, C.c_double]),
        "mmd_sync": (C.c_int, [P]),
        "mmd_eam_rho_calc_aos": (C.c_int, [P, P, C.POINTER(AosLayout), C.c_double]),
        "mmd_eam_df_calc_aos": (C.c_int, [P, P, C.POINTER(AosLayout)]),
        "mmd_eam_force_calc_aos": (C.c_int, [P, P, C.POINTER(AosLayout), C.c_double]),
        "mmd_host_alloc": (C.c_int, [C.POINTER(P), C.c_size_t]),
        "mmd_host_free": (C.c_int, [P]),
        "mmd_md_setup": (C.c_int, [P, C.POINTER(MdParams)]),
        "mmd_set_half_stencil": (C.c_int, [P, C.c_int32]),
        "mmd_ghost_refresh": (C.c_int, [P, C.c_uint32]),
        "mmd_md_prime": (C.c_int, [P]),
        "mmd_md_run": (C.c_int, [P, C.c_int32, _d, C.c_int32]),
        "mmd_md_run_noalloc": (C.c_int, [P, C.c_int32, _d]),
        "mmd_energy": (C.c_int, [P, _d, _d]),
        "mmd_itl_set": (C.c_int, [P, C.c_int32, _d, _d, _d, _d, _d, _d, _i8, _i32, C.c_int32, C.c_int32]),
        "mmd_itl_get": (C.c_int, [P, _i32, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i8]),
        "mmd_itl_migrate": (C.c_int, [P, _i32, _i32]),
        "mmd_mixed_ranges": (C.c_int, [P, _d]),
        "mmd_timing_enable": (C.c_int, [P, C.c_int32]),
        "mmd_timing_get": (C.c_int, [P, _d, _i64]),
        "mmd_timing_reset": (C.c_int, [P]),
        "mmd_stream_copy": (C.c_int, [P, C.c_size_t, C.c_int32, _d]),
        "mmd_stencil_build": (C.c_int32, [C.c_double, C.c_double, C.c_int32, _i32, C.c_int32]),
        "mmd_stencil_linear": (C.c_int, [C.POINTER(Domain), _i32, C.c_int32, _i64]),
        "mmd_synth_tables": (C.c_int, [C.POINTER(SynthParams), _d, _d, _d, _d, _d, _d, _d]),
        "mmd_fs_tables": (C.c_int, [C.POINTER(FsParams), _d, _d, _d, _d, _d, _d, _d]),
        "mmd_lattice_domain": (C.c_int, [C.POINTER(LatticeSpec), C.POINTER(Domain), _d]),
        "mmd_lattice_init": (C.c_int, [C.POINTER(LatticeSpec), C.POINTER(HostSoa), C.c_int32, _i32,
                                       _d, _d, _d, _d, _d, _d, _i8]),
        "mmd_setfl_info": (C.c_int, [C.c_char_p, _i32, _i32, _i32, _d]),
        "mmd_setfl_read": (C.c_int, [C.c_char_p, _d, _d, _d, _d, _d, _d, _d, _d]),
        "mmd_funcfl_info": (C.c_int, [C.POINTER(C.c_char_p), C.c_int32, _i32, _i32, _d]),
        "mmd_funcfl_read": (C.c_int, [C.POINTER(C.c_char_p), C.c_int32, _d, _d, _d, _d, _d, _d, _d, _d]),
        "mmd_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(EamTables), C.POINTER(P)]),
        "mmd_server_destroy": (C.c_int, [P]),
        "mmd_server_run": (C.c_int, [P, C.POINTER(ReplicaSpec), C.c_int32, C.POINTER(ReplicaResult),
                                     C.POINTER(ServerStats)]),
        "mmd_server_replica_bytes": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_size_t)]),
        "mmd_server_simulate": (C.c_int, [C.POINTER(SimParams), C.POINTER(ReplicaSpec), C.c_int32,
                                          C.POINTER(ReplicaResult), C.POINTER(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def declared_symbols():
    return [n for n in lib().__dict__ if n.startswith("mmd_")]


def _check(rc):
    if rc != OK:
        msg = lib().mmd_last_error()
        raise MmdError(rc, msg.decode() if msg else "")


def ptr(a, ct=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def _energy_rows(energies, nsteps):
    """The MD runs write {potential, kinetic} into row `step` of a [nsteps][2]
    float64 array: refuse a smaller or other array before the library writes."""
    if energies is None:
        return
    if energies.dtype != np.float64 or not energies.flags["C_CONTIGUOUS"] or energies.size < 2 * max(nsteps, 0):
        raise ValueError(f"energies must be a C-contiguous float64 array of at least [{nsteps}][2], "
                         f"got {energies.dtype} {energies.shape}")


# ---------------------------------------------------------------- host helpers

def stencil_build(a, radius, parity, cap=4096):
    out = np.zeros((cap, 3), np.int32)
    n = lib().mmd_stencil_build(a, radius, parity, ptr(out, C.c_int32), cap)
    if n < 0:
        _check(E_RANGE)
    return out[:n].copy()


def stencil_linear(dom, st3):
    st3 = np.ascontiguousarray(st3, np.int32)
    out = np.zeros(len(st3), np.int64)
    _check(lib().mmd_stencil_linear(C.byref(dom), ptr(st3, C.c_int32), len(st3), ptr(out, C.c_int64)))
    return out


def synth_tables(n_types=1, rc=5.6, A=1.2, B=0.3, Cc=0.01, jitter=0.0, seed=1, n_r=5000, r0=0.5,
                 n_rho=5000, rho_max=30.0):
    p = SynthParams(n_types, rc, A, B, Cc, jitter, seed, n_r, r0, n_rho, rho_max)
    rho_r = np.zeros((n_types, n_r))
    emb = np.zeros((n_types, n_rho))
    phi = np.zeros((n_types, n_types, n_r))
    vals = [C.c_double() for _ in range(4)]
    _check(lib().mmd_synth_tables(C.byref(p), ptr(rho_r), ptr(emb), ptr(phi), *[C.byref(v) for v in vals]))
    return dict(n_types=n_types, n_r=n_r, r0=vals[0].value, dr=vals[1].value, n_rho=n_rho,
                rho0=vals[2].value, drho=vals[3].value, rho_r=rho_r, emb=emb, phi_r=phi, rc=rc)


def fs_tables(n_types=1, rc=5.6, jitter=0.0, seed=1, n_r=5000, r0=0.5, n_rho=5000, rho_max=30.0):
    """Finnis-Sinclair BCC Fe tables (mmd_fs_tables) in the form synth_tables returns."""
    p = FsParams(n_types, rc, jitter, seed, n_r, r0, n_rho, rho_max)
    rho_r = np.zeros((n_types, n_r))
    emb = np.zeros((n_types, n_rho))
    phi = np.zeros((n_types, n_types, n_r))
    vals = [C.c_double() for _ in range(4)]
    _check(lib().mmd_fs_tables(C.byref(p), ptr(rho_r), ptr(emb), ptr(phi), *[C.byref(v) for v in vals]))
    return dict(n_types=n_types, n_r=n_r, r0=vals[0].value, dr=vals[1].value, n_rho=n_rho,
                rho0=vals[2].value, drho=vals[3].value, rho_r=rho_r, emb=emb, phi_r=phi, rc=rc)


DEFAULT_MASSES = (55.845, 58.693, 63.546)


def setfl_read(path):
    """EAM tables (and masses) from a DYNAMO setfl file, in the form synth_tables returns."""
    nt, nr, nrho, rc = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
    p = str(path).encode()
    _check(lib().mmd_setfl_info(p, C.byref(nt), C.byref(nr), C.byref(nrho), C.byref(rc)))
    nt, nr, nrho = nt.value, nr.value, nrho.value
    rho_r, emb, phi = np.zeros((nt, nr)), np.zeros((nt, nrho)), np.zeros((nt, nt, nr))
    mass = np.zeros(nt)
    vals = [C.c_double() for _ in range(4)]
    _check(lib().mmd_setfl_read(p, ptr(rho_r), ptr(emb), ptr(phi), ptr(mass), *[C.byref(v) for v in vals]))
    tab = dict(n_types=nt, n_r=nr, r0=vals[0].value, dr=vals[1].value, n_rho=nrho, rho0=vals[2].value,
               drho=vals[3].value, rho_r=rho_r, emb=emb, phi_r=phi, rc=rc.value)
    return tab, tuple(mass)


def funcfl_read(paths):
    """EAM tables (and masses) from DYNAMO funcfl files, one per element type,
    in the form synth_tables returns."""
    paths = [str(p).encode() for p in paths]
    arr = (C.c_char_p * len(paths))(*paths)
    nr, nrho, rc = C.c_int32(), C.c_int32(), C.c_double()
    _check(lib().mmd_funcfl_info(arr, len(paths), C.byref(nr), C.byref(nrho), C.byref(rc)))
    nt, nr, nrho = len(paths), nr.value, nrho.value
    rho_r, emb, phi = np.zeros((nt, nr)), np.zeros((nt, nrho)), np.zeros((nt, nt, nr))
    mass = np.zeros(nt)
    vals = [C.c_double() for _ in range(4)]
    _check(lib().mmd_funcfl_read(arr, nt, ptr(rho_r), ptr(emb), ptr(phi), ptr(mass), *[C.byref(v) for v in vals]))
    tab = dict(n_types=nt, n_r=nr, r0=vals[0].value, dr=vals[1].value, n_rho=nrho, rho0=vals[2].value,
               drho=vals[3].value, rho_r=rho_r, emb=emb, phi_r=phi, rc=rc.value)
    return tab, tuple(mass)


def lattice_spec(cells, ghost_cells=3, a=2.855, n_types=1, type_frac=(0.9, 0.05, 0.05), masses=DEFAULT_MASSES,
                 sigma=0.03, temperature=600.0, vacancy_frac=0.0, interstitial_frac=0.0, n_blobs=0,
                 blob_sigma=8.0, seed=1):
    s = LatticeSpec()
    for k in range(3):
        s.cells[k] = int(cells[k])
    s.ghost_cells = ghost_cells
    s.lattice_const = a
    s.n_types = n_types
    fr = np.asarray(type_frac[:n_types], float) if n_types > 1 else np.ones(1)
    cum = np.cumsum(fr / fr.sum())
    for t in range(8):
        s.type_cum[t] = cum[t] if t < n_types else 1.0
        s.mass[t] = masses[t] if t < len(masses) else 1.0
    s.sigma, s.temperature = sigma, temperature
    s.vacancy_frac, s.interstitial_frac = vacancy_frac, interstitial_frac
    s.n_blobs, s.blob_sigma, s.seed = n_blobs, blob_sigma, seed
    return s


class HostState:
    """Host SoA arrays over the ghost-extended lattice in the reference layout."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df")

    def __init__(self, dom):
        self.dom = dom
        n = dom.n_ext()
        for f in self.FIELDS:
            setattr(self, f, np.zeros(n))
        self.type = np.full(n, -1, np.int8)

    def soa(self):
        h = HostSoa()
        for f in self.FIELDS:
            setattr(h, f, ptr(getattr(self, f)))
        h.type = ptr(self.type, C.c_int8)
        return h

    def local_index(self):
        """Linear indices of the local (non-ghost) sites, z-major."""
        d = self.dom
        gx, gy, gz = d.ghost
        bx, by, bz = d.box
        ex, ey = d.ext[0], d.ext[1]
        z, y, x = np.meshgrid(np.arange(bz) + gz, np.arange(by) + gy, np.arange(bx) + gx, indexing="ij")
        return ((z * ey + y) * ex + x).ravel()

    def copy(self):
        o = HostState(self.dom)
        for f in self.FIELDS + ("type",):
            setattr(o, f, getattr(self, f).copy())
        return o


class Interstitials:
    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "rho", "df")

    def __init__(self, n):
        for f in self.FIELDS:
            setattr(self, f, np.zeros(n))
        self.type = np.zeros(n, np.int8)
        self.n = n

    def take(self, n):
        o = Interstitials(n)
        for f in self.FIELDS + ("type",):
            setattr(o, f, getattr(self, f)[:n].copy())
        return o


def lattice_init(spec, itl_cap=0):
    dom = Domain()
    box = np.zeros(3)
    _check(lib().mmd_lattice_domain(C.byref(spec), C.byref(dom), ptr(box)))
    st = HostState(dom)
    itl = Interstitials(max(itl_cap, 1))
    n_itl = C.c_int32(0)
    h = st.soa()
    _check(lib().mmd_lattice_init(C.byref(spec), C.byref(h), itl_cap, C.byref(n_itl), ptr(itl.x), ptr(itl.y),
                                  ptr(itl.z), ptr(itl.vx), ptr(itl.vy), ptr(itl.vz), ptr(itl.type, C.c_int8)))
    return dom, box, st, itl.take(n_itl.value)


def tables_struct(tab, masses):
    keep = {k: np.ascontiguousarray(tab[k], float) for k in ("rho_r", "emb", "phi_r")}
    m = np.ascontiguousarray(masses[: tab["n_types"]], float)
    keep["mass"] = m
    t = EamTables(tab["n_types"], tab["n_r"], tab["r0"], tab["dr"], tab["n_rho"], tab["rho0"], tab["drho"],
                  ptr(keep["rho_r"]), ptr(keep["emb"]), ptr(keep["phi_r"]), ptr(m))
    t._keep = keep
    return t


# -------------------------------------------------------------------- context

class Context:
    """One device-resident sub-domain (mmd_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        _check(L.mmd_ctx_create(device, C.byref(h)))
        self.h = h
        self.dom = None

    def close(self):
        if self.h:
            _check(lib().mmd_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_domain(self, dom):
        _check(lib().mmd_set_domain(self.h, C.byref(dom)))
        self.dom = dom

    def set_stencil(self, even, odd):
        even = np.ascontiguousarray(even, np.int64)
        odd = np.ascontiguousarray(odd, np.int64)
        _check(lib().mmd_set_stencil(self.h, ptr(even, C.c_int64), len(even), ptr(odd, C.c_int64), len(odd)))

    def set_tables(self, tab, masses=DEFAULT_MASSES):
        t = tables_struct(tab, masses)
        _check(lib().mmd_set_tables(self.h, C.byref(t)))

    def set_precision(self, mode):
        _check(lib().mmd_set_precision(self.h, mode))

    def upload(self, st, fields=F_ALL):
        h = st.soa()
        _check(lib().mmd_upload(self.h, fields, C.byref(h)))

    def download(self, st, fields=F_ALL):
        h = st.soa()
        _check(lib().mmd_download(self.h, fields, C.byref(h)))

    def rho_df(self, rc):
        _check(lib().mmd_eam_rho_df(self.h, rc))

    def df(self):
        _check(lib().mmd_eam_df(self.h))

    def force(self, rc):
        _check(lib().mmd_eam_force(self.h, rc))

    def sync(self):
        _check(lib().mmd_sync(self.h))

    def rho_calc_aos(self, recs, rc, layout=None):
        lay = layout or ref_aos_layout()
        _check(lib().mmd_eam_rho_calc_aos(self.h, recs.ctypes.data_as(C.c_void_p), C.byref(lay), rc))

    def df_calc_aos(self, recs, layout=None):
        lay = layout or ref_aos_layout()
        _check(lib().mmd_eam_df_calc_aos(self.h, recs.ctypes.data_as(C.c_void_p), C.byref(lay)))

    def force_calc_aos(self, recs, rc, layout=None):
        lay = layout or ref_aos_layout()
        _check(lib().mmd_eam_force_calc_aos(self.h, recs.ctypes.data_as(C.c_void_p), C.byref(lay), rc))

    def md_setup(self, box_len, a, dt, rc, energy_every=0):
        p = MdParams((C.c_double * 3)(*box_len), a, dt, rc, energy_every)
        _check(lib().mmd_md_setup(self.h, C.byref(p)))

    def set_half_stencil(self, on=True):
        """Newton's-third-law half stencil for the periodic step (md_prime, md_run, energy)."""
        _check(lib().mmd_set_half_stencil(self.h, 1 if on else 0))

    def ghost_refresh(self, fields):
        _check(lib().mmd_ghost_refresh(self.h, fields))

    def md_prime(self):
        _check(lib().mmd_md_prime(self.h))

    def md_run(self, nsteps, energies=None, use_graph=False):
        _energy_rows(energies, nsteps)
        _check(lib().mmd_md_run(self.h, nsteps, ptr(energies), 1 if use_graph else 0))

    def md_run_noalloc(self, nsteps, energies=None):
        _energy_rows(energies, nsteps)
        _check(lib().mmd_md_run_noalloc(self.h, nsteps, ptr(energies)))

    def energy(self):
        ep, ek = C.c_double(), C.c_double()
        _check(lib().mmd_energy(self.h, C.byref(ep), C.byref(ek)))
        return ep.value, ek.value

    def itl_set(self, itl, wide3, cap=None):
        wide3 = np.ascontiguousarray(wide3, np.int32)
        cap = max(cap or itl.n, 1)
        _check(lib().mmd_itl_set(self.h, itl.n, ptr(itl.x), ptr(itl.y), ptr(itl.z), ptr(itl.vx), ptr(itl.vy),
                                 ptr(itl.vz), ptr(itl.type, C.c_int8), ptr(wide3, C.c_int32), len(wide3), cap))

    def itl_get(self, cap):
        o = Interstitials(max(cap, 1))
        n = C.c_int32(0)
        _check(lib().mmd_itl_get(self.h, C.byref(n), *[ptr(getattr(o, f)) for f in Interstitials.FIELDS],
                                 ptr(o.type, C.c_int8)))
        return o.take(n.value)

    def mixed_ranges(self):
        """Squared-distance ranges the mixed tables cover: (head, tail), [(lo, hi), (lo, hi)] holes."""
        o = np.zeros(6)
        _check(lib().mmd_mixed_ranges(self.h, ptr(o)))
        return (o[0], o[1]), [(o[2], o[3]), (o[4], o[5])]

    def itl_migrate(self):
        """Wigner-Seitz re-assignment (mmd_itl_migrate); returns the counts
        (lattice -> list, list -> lattice)."""
        a, b = C.c_int32(0), C.c_int32(0)
        _check(lib().mmd_itl_migrate(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def timing(self, on=True):
        _check(lib().mmd_timing_enable(self.h, 1 if on else 0))

    def timing_get(self):
        ms = np.zeros(K_COUNT)
        n = np.zeros(K_COUNT, np.int64)
        _check(lib().mmd_timing_get(self.h, ptr(ms), ptr(n, C.c_int64)))
        return ms, n

    def stream_copy(self, nbytes, repeats=10):
        ms = C.c_double()
        _check(lib().mmd_stream_copy(self.h, nbytes, repeats, C.byref(ms)))
        return ms.value

    def timing_reset(self):
        _check(lib().mmd_timing_reset(self.h))


def ref_aos_layout():
    return AosLayout(104, 8, 16, 40, 64, 88, 96, 4)


def wide_stencil(a, rc, skin, reach=1.6):
    """Interstitial neighbourhood W (parity 0): the anchor site itself plus every
    site within rc + skin + reach of it."""
    w = stencil_build(a, rc + skin + reach, 0)
    return np.vstack([np.zeros((1, 3), np.int32), w])


# -------------------------------------------------------------------- serving

def _struct_dict(o):
    return {k: getattr(o, k) for k, _ in o._fields_}


def replica_bytes(cells, ghost_cells=3):
    b = C.c_size_t(0)
    _check(lib().mmd_server_replica_bytes(int(cells), int(ghost_cells), C.byref(b)))
    return b.value


def _spec_array(reps):
    """reps: sequence of dicts with cells, temperature, steps, seed, arrival_s and optional id / keep
    (a HostState that receives the final ext state in the reference layout)."""
    arr = (ReplicaSpec * max(len(reps), 1))()
    keep = []
    for i, r in enumerate(reps):
        a = arr[i]
        a.id = int(r.get("id", i))
        a.cells, a.temperature, a.steps = int(r["cells"]), float(r["temperature"]), int(r["steps"])
        a.seed, a.arrival_s = int(r["seed"]), float(r.get("arrival_s", 0.0))
        if r.get("keep") is not None:
            h = r["keep"].soa()
            keep.append(h)
            a.keep = C.pointer(h)
    return arr, keep


def simulate(reps, arena_bytes, max_replicas=8, chunk_steps=10, ghost_cells=3, workers=4, step_overhead_s=1e-4,
             step_atom_s=1e-9, load_atom_s=1e-8):
    """Runs the serving scheduler against a simulated executor (virtual clock, fake memory budget)."""
    p = SimParams(int(arena_bytes), max_replicas, chunk_steps, ghost_cells, workers, step_overhead_s, step_atom_s,
                  load_atom_s)
    arr, _ = _spec_array(reps)
    res = (ReplicaResult * max(len(reps), 1))()
    st = ServerStats()
    _check(lib().mmd_server_simulate(C.byref(p), arr, len(reps), res, C.byref(st)))
    return [_struct_dict(res[i]) for i in range(len(reps))], _struct_dict(st)


class Server:
    """Continuous-batching MD server (mmd_server): replicas of a random alloy, one GPU."""

    def __init__(self, tab, masses=DEFAULT_MASSES, device=0, precision=PREC_MIXED, a=2.855, cutoff=5.6, skin=0.3,
                 dt=1.0, ghost_cells=3, type_frac=(0.9, 0.05, 0.05), sigma=0.05, max_replicas=64,
                 hbm_fraction=0.5, hbm_limit=0, chunk_steps=10, workers=8, use_graph=True):
        o = ServerOpts()
        o.device, o.precision, o.lattice_const, o.cutoff, o.skin, o.dt = device, precision, a, cutoff, skin, dt
        o.ghost_cells = ghost_cells
        nt = tab["n_types"]
        fr = np.asarray(type_frac[:nt], float) if nt > 1 else np.ones(1)
        cum = np.cumsum(fr / fr.sum())
        for t in range(8):
            o.type_cum[t] = cum[t] if t < nt else 1.0
        o.sigma, o.max_replicas, o.hbm_fraction, o.hbm_limit = sigma, max_replicas, hbm_fraction, int(hbm_limit)
        o.chunk_steps, o.workers, o.use_graph = chunk_steps, workers, int(bool(use_graph))
        self._tab = tables_struct(tab, masses)
        h = C.c_void_p()
        _check(lib().mmd_server_create(C.byref(o), C.byref(self._tab), C.byref(h)))
        self.h = h
        self.opts = o

    def run(self, reps):
        arr, keep = _spec_array(reps)
        res = (ReplicaResult * max(len(reps), 1))()
        st = ServerStats()
        _check(lib().mmd_server_run(self.h, arr, len(reps), res, C.byref(st)))
        del keep
        return [_struct_dict(res[i]) for i in range(len(reps))], _struct_dict(st)

    def close(self):
        if self.h:
            lib().mmd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
