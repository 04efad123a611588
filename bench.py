#!/usr/bin/env python3
"""MISA-MD EAM step throughput on one MI355X.

Headline workload (the FP64 metric): 128x128x128 BCC cells (4,194,304 atoms),
a = 2.855 A, 3 atom types drawn 90/5/5 from the seed, 9 synthetic pair tables
of the named functional forms (phi = A (rc-r)^2 (1 + B/r), rho = (rc-r)^2
exp(-r), F = -sqrt(rho) + C rho^2; A, B, C jittered +-10%, phi symmetric),
5000-point splines on r in [0.5, 5.6] A, Gaussian displacement 0.05 A,
Maxwell velocities at 800 K, velocity Verlet with dt = 1 fs.  Everything is
device resident: ghost refresh, rho + df, ghost df, force and integration run
on the GPU, energies are recorded every 10 steps.  One step = one full
velocity-Verlet step over every atom; `value` is atom-steps/s of the FP64 path.

The same run also measures
  * mixed precision against FP64 on the defect box (same 128^3 3-type alloy
    plus 0.5% vacancies and 0.2% interstitials at tetrahedral sites, 1200 K,
    200 steps): atom-steps/s of both, force error of mixed vs FP64 on
    identical states (start and end of the run, and every 10 steps along the
    mixed trajectory), the per-step energy difference per atom;
  * the Newton's-third-law half stencil against the full stencil on the
    headline box, FP64 and mixed (`half_stencil`);
  * the same comparison on the large box (256^3 cells, 33.4 M atoms, 1%
    vacancies and 0.5% interstitials clustered in 20 Gaussian blobs of
    sigma 8 A, 1800 K, 100 steps), reported as `large_box`;
  * the same FP64 step and the same mixed defect-box comparison with
    Finnis-Sinclair Fe tables (a mechanically stable lattice; the synthetic
    forms are far from equilibrium and disorder within ~100 steps), plus
    interstitial migration on that defect box's final state, reported apart
    as `stable_potential`;
  * the serving loop: 2,000 independent replicas (box edge log-uniform in
    10..64 BCC cells, 3-type alloy, 300-1500 K, 200 steps each, mixed
    precision) arriving as a Poisson stream at 75% of the capacity measured
    by a closed-loop burst just before; throughput over the whole trace and
    p50/p99 per-replica latency (arrival to results); the same burst served
    in FP64 gives the mixed-over-FP64 capacity ratio of the serving form;
  * the CPU oracle (FP64 C restatement of the reference's EAM algorithm,
    OpenMP on the host cores) on the headline box, one thread on the 20^3-cell
    single-element box, and the per-atom check of the GPU's final forces
    against the oracle on the headline box.

Inputs are resident in HBM before the timed region starts.  Per-kernel
durations come from HIP events recorded on the context's stream around every
launch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "misa-md-hip_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import mmd  # noqa: E402

A, RC, SKIN, DT = 2.855, 5.6, 0.3, 1.0
PEAK_HBM_GBPS = 8000.0
PEAK_FP64_TFLOPS = 78.6
ENERGY_EVERY = 10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(cells, seed=2024, temperature=800.0, vacancy_frac=0.0, interstitial_frac=0.0, potential="synthetic",
                  n_types=3, n_blobs=0, blob_sigma=8.0):
    """The benchmark box.  potential "synthetic": the functional forms the
    benchmark configurations name (phi = A (rc-r)^2 (1 + B/r), rho = (rc-r)^2
    exp(-r), F = -sqrt(rho) + C rho^2, A, B, C jittered +-10% per type / pair);
    "fs": Finnis-Sinclair Fe, a mechanically stable lattice (reported apart)."""
    spec = mmd.lattice_spec((cells, cells, cells), ghost_cells=3, a=A, n_types=n_types, sigma=0.05,
                            temperature=temperature, vacancy_frac=vacancy_frac,
                            interstitial_frac=interstitial_frac, n_blobs=n_blobs, blob_sigma=blob_sigma, seed=seed)
    cap = int(interstitial_frac * 2 * cells ** 3 * 1.5) + 16
    dom, box, st, itl = mmd.lattice_init(spec, itl_cap=cap)
    jit = 0.1 if n_types > 1 else 0.0
    if potential == "synthetic":
        tab = mmd.synth_tables(n_types=n_types, rc=RC, jitter=jit, seed=seed)
    else:
        tab = mmd.fs_tables(n_types=n_types, rc=RC, jitter=jit, seed=seed)
    st3 = mmd.stencil_build(A, RC + SKIN, 0)
    return dict(dom=dom, box=box, st=st, itl=itl, tab=tab, st3=st3, even=mmd.stencil_linear(dom, st3),
                odd=mmd.stencil_linear(dom, -st3), wide3=mmd.wide_stencil(A, RC, SKIN))


def n_atoms_of(w):
    li = w["st"].local_index()
    return int((w["st"].type[li] >= 0).sum()) + w["itl"].n


def pairs_in_cutoff(w, sample=20000, seed=0):
    """Mean number of in-cutoff lattice neighbours per atom (N_c) over a sample."""
    st, dom = w["st"], w["dom"]
    li = st.local_index()
    li = li[st.type[li] >= 0]
    rng = np.random.default_rng(seed)
    idx = rng.choice(li, size=min(sample, len(li)), replace=False)
    X = (idx % dom.ext[0]) - dom.ghost[0]
    par = X % 2
    tot = 0
    for p, offs in ((0, w["even"]), (1, w["odd"])):
        i = idx[par == p]
        j = i[:, None] + offs[None, :]
        d2 = (st.x[i, None] - st.x[j]) ** 2 + (st.y[i, None] - st.y[j]) ** 2 + (st.z[i, None] - st.z[j]) ** 2
        tot += ((d2 < RC * RC) & (st.type[j] >= 0)).sum()
    return tot / len(idx)


def open_ctx(w, prec, energy_every=ENERGY_EVERY, state=None, itl=None, cap=None):
    ctx = mmd.Context(0)
    ctx.set_domain(w["dom"])
    ctx.set_stencil(w["even"], w["odd"])
    ctx.set_tables(w["tab"], mmd.DEFAULT_MASSES)
    ctx.set_precision(prec)
    ctx.upload(state or w["st"], mmd.F_X | mmd.F_V | mmd.F_TYPE)
    ctx.md_setup(w["box"], A, DT, RC, energy_every=energy_every)
    it = itl if itl is not None else w["itl"]
    if it.n:
        ctx.itl_set(it, w["wide3"], cap=cap or max(it.n, 1))
    ctx.md_prime()
    return ctx


def forces(ctx, w):
    f = w["st"].copy()
    ctx.download(f, mmd.F_F)
    li = f.local_index()
    li = li[w["st"].type[li] >= 0]
    F = np.stack([f.fx[li], f.fy[li], f.fz[li]], 1)
    if w["itl"].n:
        gi = ctx.itl_get(w["itl"].n)
        F = np.concatenate([F, np.stack([gi.fx, gi.fy, gi.fz], 1)])
    return F


def run_gpu(w, prec, steps, warmup, keep_state=False, half=False):
    ctx = open_ctx(w, prec)
    if half:
        ctx.set_half_stencil(True)
    f0 = forces(ctx, w)
    e0 = ctx.energy()
    if warmup:
        ctx.md_run(warmup)
    ctx.sync()
    ctx.timing(True)
    ctx.timing_reset()
    en = np.zeros((steps, 2))
    t0 = time.perf_counter()
    ctx.md_run(steps, en)
    ctx.sync()
    t1 = time.perf_counter()
    ms, n = ctx.timing_get()
    ctx.timing(False)
    out = dict(seconds=t1 - t0, kernel_ms=ms, launches=n, energies=en, f0=f0, e0=e0)
    if keep_state:
        st = w["st"].copy()
        ctx.download(st, mmd.F_X | mmd.F_V | mmd.F_TYPE | mmd.F_F)
        out["state"] = st
        out["itl"] = ctx.itl_get(w["itl"].n) if w["itl"].n else w["itl"]
        out["f_end"] = forces(ctx, w)
    ctx.close()
    return out


def force_err(F, Fr):
    d = np.linalg.norm(F - Fr, axis=1)
    rms = np.sqrt((np.linalg.norm(Fr, axis=1) ** 2).mean())
    return float(np.linalg.norm(F - Fr) / np.linalg.norm(Fr)), float(d.max() / rms)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, steps, gpu_end=None):
    """The oracle on the host cores: the headline box for `steps` steps on all
    threads, the small single-element box on one thread, and (gpu_end) a
    per-atom check of the GPU run's final forces against the oracle on the
    same positions."""
    import oracle as O
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    T = O.Tables(w["tab"], mmd.DEFAULT_MASSES)
    out = {}
    if gpu_end is not None:
        st = gpu_end.copy()
        O.rho_df(w["dom"], w["even"], w["odd"], T, RC, st, nthreads=nthreads)
        O.ghost_df(w["dom"], st)
        O.force(w["dom"], w["even"], w["odd"], T, RC, st, nthreads=nthreads)
        li = st.local_index()
        occ = li[st.type[li] >= 0]
        Fo = np.stack([st.fx[occ], st.fy[occ], st.fz[occ]], 1)
        Fg = np.stack([gpu_end.fx[occ], gpu_end.fy[occ], gpu_end.fz[occ]], 1)
        out["gpu_check"] = {"what": "FP64 GPU forces after the timed steps vs the oracle on the same positions, "
                                    "every atom", "atoms": int(occ.size),
                            "max_abs_err_over_max_force": float(np.abs(Fg - Fo).max() / np.abs(Fo).max())}
    st = w["st"].copy()
    O.rho_df(w["dom"], w["even"], w["odd"], T, RC, st, nthreads=nthreads)
    O.ghost_df(w["dom"], st)
    O.force(w["dom"], w["even"], w["odd"], T, RC, st, nthreads=nthreads)
    n_atoms = n_atoms_of(w)
    t0 = time.perf_counter()
    O.md_run(w["dom"], w["even"], w["odd"], T, RC, w["box"], DT, steps, st, nthreads=nthreads, energies=False)
    dt = time.perf_counter() - t0
    # one thread on the small single-element box (20^3 cells, 16,000 atoms, 10 steps)
    w1 = make_workload(20, seed=7, n_types=1)
    s1 = w1["st"].copy()
    T1 = O.Tables(w1["tab"], mmd.DEFAULT_MASSES)
    O.rho_df(w1["dom"], w1["even"], w1["odd"], T1, RC, s1, nthreads=1)
    O.ghost_df(w1["dom"], s1)
    O.force(w1["dom"], w1["even"], w1["odd"], T1, RC, s1, nthreads=1)
    t1 = time.perf_counter()
    O.md_run(w1["dom"], w1["even"], w1["odd"], T1, RC, w1["box"], DT, 10, s1, nthreads=1, energies=False)
    d1 = time.perf_counter() - t1
    out.update(value=round(n_atoms * steps / dt, 1), unit="atom-steps/s", cores=nthreads, kind="port",
               sample=f"CPU oracle (FP64 C restatement of the reference EAM step, OpenMP, {nthreads} threads) on "
                      f"the same {n_atoms}-atom box for {steps} velocity-Verlet steps ({dt:.2f} s wall)",
               one_thread={"value": round(n_atoms_of(w1) * 10 / d1, 1), "unit": "atom-steps/s",
                           "sample": f"20^3 cells single element ({n_atoms_of(w1)} atoms), 10 steps, 1 thread"},
               nproc=os.cpu_count(), cpu_model=cpu_model())
    return out


def load_traffic():
    """HBM bytes per force launch from rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration (profiles/)."""
    p = os.path.join(ROOT, "profiles", "force_fp64_128cube_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def half_stencil(w, steps, warmup):
    """The Newton's-third-law half stencil (mmd_set_half_stencil) against the
    full stencil on the same box, FP64 and mixed: ms per step and the force
    pass (forward + backward launches for the half stencil)."""
    out = {"what": "each lattice pair evaluated once (forward entries), reactions pulled from per-entry pair-force "
                   "slots by a second pass; full stencil = every pair evaluated from both sides", "steps": steps}
    n_atoms = n_atoms_of(w)
    for prec, key in ((mmd.PREC_FP64, "fp64"), (mmd.PREC_MIXED, "mixed")):
        row = {}
        for half in (False, True):
            r = run_gpu(w, prec, steps, warmup, half=half)
            ms, n = r["kernel_ms"], r["launches"]
            row["half" if half else "full"] = {
                "value": round(n_atoms * steps / r["seconds"], 1), "unit": "atom-steps/s",
                "ms_per_step": round(r["seconds"] / steps * 1e3, 4),
                "force_ms": round(ms[mmd.K_FORCE] / max(n[mmd.K_FORCE], 1), 4)}
        out[key] = row
    return out


def force_err_along(w, steps, warmup, every=ENERGY_EVERY):
    """Mixed against FP64 forces on identical states every `every` steps of the
    mixed trajectory (untimed): the mixed run is stopped after each chunk, its
    state copied into an FP64 context and the forces of both compared."""
    ctx = open_ctx(w, mmd.PREC_MIXED)
    if warmup:
        ctx.md_run(warmup)
    out = []
    for s in range(every, steps + 1, every):
        ctx.md_run(every)
        st = w["st"].copy()
        ctx.download(st, mmd.F_X | mmd.F_V | mmd.F_TYPE | mmd.F_F)
        it = ctx.itl_get(w["itl"].n) if w["itl"].n else w["itl"]
        fm = forces(ctx, w)
        ref = open_ctx(w, mmd.PREC_FP64, state=st, itl=it)
        f64 = forces(ref, w)
        ref.close()
        nw, pa = force_err(fm, f64)
        out.append({"step": s, "normwise": nw, "max_over_rms": pa})
    ctx.close()
    return out


def mixed_vs_fp64(w, steps, warmup, migrate=False, desc=None, per_step_err=False):
    n_atoms = n_atoms_of(w)
    r64 = run_gpu(w, mmd.PREC_FP64, steps, warmup, keep_state=True)
    rmx = run_gpu(w, mmd.PREC_MIXED, steps, warmup, keep_state=True)
    v64 = n_atoms * steps / r64["seconds"]
    vmx = n_atoms * steps / rmx["seconds"]
    nw0, pa0 = force_err(rmx["f0"], r64["f0"])
    # FP64 forces on the mixed run's final state
    ctx = open_ctx(w, mmd.PREC_FP64, state=rmx["state"], itl=rmx["itl"])
    f_ref_end = forces(ctx, w)
    ctx.close()
    nw1, pa1 = force_err(rmx["f_end"], f_ref_end)
    mig = migration(w, rmx["state"], rmx["itl"]) if migrate else None
    e64, emx = r64["energies"].sum(1), rmx["energies"].sum(1)
    rec = np.arange(ENERGY_EVERY - 1, steps, ENERGY_EVERY)
    d64 = e64[rec] - (r64["e0"][0] + r64["e0"][1])
    dmx = emx[rec] - (rmx["e0"][0] + rmx["e0"][1])
    per_step = np.abs(dmx - d64) / n_atoms / (rec + 1 + warmup)
    mk, mn = rmx["kernel_ms"], rmx["launches"]
    # roofline of the mixed force kernel: the algorithmic flop count of the
    # pass; the mixed path stores FP32 spline corrections but evaluates in
    # FP64 (FP32 arithmetic only in the phi value of energy steps), so the
    # FP64 vector peak is the roof
    S, Nc = len(w["st3"]), pairs_in_cutoff(w)
    t_mf = mk[mmd.K_FORCE] / max(mn[mmd.K_FORCE], 1) * 1e-3
    ach = (8 * S + 40 * Nc) * n_atoms / t_mf / 1e12
    mixed_roof = {"kernel": "k_force<mixed>", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS,
                  "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4),
                  "flops_per_atom": {"fp64": round(8 * S + 40 * Nc, 1), "fp32": round(2 * Nc / ENERGY_EVERY, 1)},
                  "hbm_frac": round(57 * n_atoms / t_mf / 1e9 / PEAK_HBM_GBPS, 4)}
    desc = desc or (f"BCC {w['dom'].box[1]}^3 cells, 3-type alloy, 0.5% vacancies, "
                    f"{w['itl'].n} interstitials (0.2%), 1200 K, dt 1 fs")
    return {"workload": desc,
            "atoms": n_atoms, "steps": steps, "fp64_value": round(v64, 1), "value": round(vmx, 1),
            "speedup_vs_fp64": round(vmx / v64, 3),
            "force_rel_err_normwise": max(nw0, nw1), "force_err_max_over_rms": max(pa0, pa1),
            "energy_diff_per_atom_per_step_max": float(per_step.max()) if len(per_step) else None,
            "force_kernel_ms": round(mk[mmd.K_FORCE] / max(mn[mmd.K_FORCE], 1), 4),
            "roofline": mixed_roof,
            "rho_df_kernel_ms": round(mk[mmd.K_RHO] / max(mn[mmd.K_RHO], 1), 4),
            "interstitial_ms_per_step": round(mk[mmd.K_ITL] / steps, 4),
            **({"force_err_every_10_steps": force_err_along(w, steps, warmup)} if per_step_err else {}),
            **({"migration": mig} if mig else {})}


def migration(w, state, itl, reps=5):
    """Wigner-Seitz migration (mmd_itl_migrate) on the state a run ended in:
    wall time of the first call (the transitions since the start) and of calls
    with nothing left to move (two full-lattice passes plus a host sync)."""
    ctx = open_ctx(w, mmd.PREC_MIXED, state=state, itl=itl, cap=itl.n + 65536)
    ctx.sync()
    t0 = time.perf_counter()
    out, fill = ctx.itl_migrate()
    t1 = time.perf_counter()
    for _ in range(reps):
        ctx.itl_migrate()
    t2 = time.perf_counter()
    ctx.close()
    return {"to_interstitial": out, "to_lattice": fill, "first_call_ms": round((t1 - t0) * 1e3, 3),
            "steady_call_ms": round((t2 - t1) * 1e3 / reps, 3)}


def dropin_hooks(w, reps=3):
    """The reference hook sequence on host records: per step the rho pass (df fused) and the force pass, each
    copying the whole ghost-extended array of 104-byte _cuAtomElement records (pinned, as hip_create_atoms_mem
    allocates them) to the device, as the reference plugin does, by z-chunks while earlier chunks are computed;
    the outputs (rho and df, or f) are written into the occupied local records directly as their tiles finish
    (the reference copied the local layers back whole)."""
    import ctypes as C
    n = w["dom"].n_ext()
    nbytes = n * mmd.REF_AOS_DTYPE.itemsize
    p = C.c_void_p()
    mmd._check(mmd.lib().mmd_host_alloc(C.byref(p), nbytes))
    try:
        recs = np.frombuffer((C.c_char * nbytes).from_address(p.value), dtype=mmd.REF_AOS_DTYPE)
        st = w["st"]
        recs["id"] = np.arange(n)
        recs["type"] = st.type
        recs["x"] = np.stack([st.x, st.y, st.z], 1)
        recs["v"] = np.stack([st.vx, st.vy, st.vz], 1)
        with mmd.Context(0) as ctx:
            ctx.set_domain(w["dom"])
            ctx.set_stencil(w["even"], w["odd"])
            ctx.set_tables(w["tab"], mmd.DEFAULT_MASSES)
            ctx.rho_calc_aos(recs, RC)
            ctx.force_calc_aos(recs, RC)
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.rho_calc_aos(recs, RC)
                ctx.force_calc_aos(recs, RC)
            dt = (time.perf_counter() - t0) / reps
        del recs
    finally:
        mmd.lib().mmd_host_free(p)
    atoms = n_atoms_of(w)
    return {"what": "rho (fused df) + force hooks on pinned host _cuAtomElement records, PCIe transfers included "
                    "(records copied in by z-chunks overlapped with the pass and the output write-back)",
            "atom_passes_per_s": round(atoms / dt, 1), "ms_per_step": round(dt * 1e3, 3),
            "host_bytes_per_step": int(2 * nbytes + atoms * (16 + 24)),
            "reference_host_bytes_per_step": int(2 * (nbytes + w["dom"].box[2] * w["dom"].ext[0] * w["dom"].ext[1]
                                                      * 104))}


def replica_trace(n, seed, rate, t0=0.0, steps=200):
    """Replicas with box edge log-uniform in [10, 64] cells, 300-1500 K, Poisson arrivals at `rate`/s."""
    rng = np.random.default_rng(seed)
    cells = np.rint(np.exp(rng.uniform(np.log(10.0), np.log(64.0), n))).astype(int)
    temps = rng.uniform(300.0, 1500.0, n)
    arr = t0 + np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.full(n, t0)
    return [dict(id=i, cells=int(cells[i]), temperature=float(temps[i]), steps=steps, seed=seed * 100003 + i,
                 arrival_s=float(arr[i])) for i in range(n)]


def burst_capacity(srv, burst, seed, steps):
    """Atom-steps/s of a burst of the replica mix, all present at t = 0, measured while the admission queue is
    non-empty (from the first replica resident to the last admission), each replica's atom-steps spread evenly
    over its residency."""
    res, st = srv.run(replica_trace(burst, seed, 0.0, steps=steps))
    t0 = min(r["t_ready"] for r in res)
    t1 = max(r["t_admit"] for r in res)
    work = 0.0
    for r in res:
        span = r["t_done"] - r["t_ready"]
        overlap = max(0.0, min(r["t_done"], t1) - max(r["t_ready"], t0))
        work += r["n_atoms"] * steps * (overlap / span if span > 0 else 0.0)
    return work / (t1 - t0)


def serving(n_replicas, load, seed=77, burst=512, steps=200, use_graph=True, ghost_cells=2):
    """Replicas carry no interstitials, so their ghost shell only needs the
    stencil's reach (2 cells at rc + skin = 5.9 A).  The same burst is also
    served in FP64, so the mixed serving form is compared with the FP64 one
    on the same replicas (their temperatures and 200-step trajectories, which
    the synthetic forms disorder, slow both precisions alike)."""
    tab = mmd.synth_tables(n_types=3, rc=RC, jitter=0.1, seed=seed)
    opts = dict(a=A, cutoff=RC, skin=SKIN, dt=DT, sigma=0.05, max_replicas=64, hbm_fraction=0.5, chunk_steps=10,
                workers=16, use_graph=use_graph, ghost_cells=ghost_cells)
    with mmd.Server(tab, precision=mmd.PREC_FP64, **opts) as srv:
        capacity_fp64 = burst_capacity(srv, burst, seed + 1, steps)
    with mmd.Server(tab, precision=mmd.PREC_MIXED, **opts) as srv:
        capacity = burst_capacity(srv, burst, seed + 1, steps)
        reps = replica_trace(n_replicas, seed, 1.0, steps=steps)
        mean_work = np.mean([2 * r["cells"] ** 3 * r["steps"] for r in reps])
        rate = load * capacity / mean_work
        reps = replica_trace(n_replicas, seed, rate, steps=steps)
        log(f"bench: serving {n_replicas} replicas, capacity {capacity:.3e} atom-steps/s, "
            f"arrivals {rate:.2f}/s (load {load})")
        res, st = srv.run(reps)
    done = [r for r in res if r["status"] == 0]
    de = [abs(r["epot"] + r["ekin"] - r["epot0"] - r["ekin0"]) / r["n_atoms"] / steps for r in done]
    return {"workload": f"{n_replicas} replicas, box edge log-uniform 10-64 BCC cells (2k-524k atoms), 3-type alloy, "
                        f"300-1500 K, {steps} steps each, Poisson arrivals at {load:.0%} of measured capacity, "
                        "mixed precision, one batched launch per pass over (replica, block) work items, grids sized "
                        "to the work lists' capacities" + (", replayed from one captured HIP graph for every mix"
                                                            if use_graph else ", launched directly"),
            "value": round(st["atom_steps_per_s"], 1), "unit": "atom-steps/s",
            "capacity_atom_steps_per_s": round(capacity, 1),
            "capacity_fp64_serving_atom_steps_per_s": round(capacity_fp64, 1),
            "capacity_vs_fp64_serving": round(capacity / capacity_fp64, 3), "arrival_rate_per_s": round(rate, 3),
            "offered_load": load, "latency_p50_s": round(st["latency_p50_s"], 4),
            "latency_p99_s": round(st["latency_p99_s"], 4), "latency_mean_s": round(st["latency_mean_s"], 4),
            "wait_mean_s": round(st["wait_mean_s"], 4), "wall_s": round(st["wall_s"], 3),
            "atom_steps": st["atom_steps"], "completed": st["completed"], "failed": st["failed"],
            "max_active": st["max_active"], "batched_steps": st["batched_steps"],
            "graph_captures": st["graph_captures"],
            "total_energy_change_per_atom_per_step_median": float(np.median(de)) if de else None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-fs", action="store_true", help="skip the stable-potential (Finnis-Sinclair) FP64 line")
    ap.add_argument("--mixed-steps", type=int, default=200)
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3-cell large-box line")
    ap.add_argument("--no-half", action="store_true", help="skip the half-stencil comparison")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in hook measurement (its pipelined passes add small pair-kernel launches)")
    ap.add_argument("--half-steps", type=int, default=50)
    ap.add_argument("--large-steps", type=int, default=100)
    ap.add_argument("--serve-replicas", type=int, default=2000)
    ap.add_argument("--serve-load", type=float, default=0.75)
    ap.add_argument("--serve-no-graph", action="store_true", help="launch the batched step directly (profiling)")
    ap.add_argument("--serve-ghost", type=int, default=2, help="ghost shell of the serving replicas (cells)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only time a device-to-device streaming copy of BYTES (HBM bandwidth / counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs one sub-domain per GPU and has no multi-GPU "
              f"path; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        with mmd.Context(0) as ctx:
            ms = ctx.stream_copy(args.copy_probe, repeats=10)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms,
                          "GBps_read_plus_write": 2 * args.copy_probe / ms / 1e6}))
        return

    log(f"bench: building {args.cells}^3-cell 3-type BCC workload")
    w = make_workload(args.cells)
    n_atoms = n_atoms_of(w)
    S = len(w["st3"])
    Nc = pairs_in_cutoff(w)

    log(f"bench: FP64 {args.steps} steps (+{args.warmup} warmup), {n_atoms} atoms")
    r64 = run_gpu(w, mmd.PREC_FP64, args.steps, args.warmup, keep_state=not args.no_cpu)
    value = n_atoms * args.steps / r64["seconds"]
    ms, n = r64["kernel_ms"], r64["launches"]
    t_force = ms[mmd.K_FORCE] / max(n[mmd.K_FORCE], 1) * 1e-3
    t_rho = ms[mmd.K_RHO] / max(n[mmd.K_RHO], 1) * 1e-3
    flops_atom = 8 * S + 40 * Nc
    flops_force = flops_atom * n_atoms
    flops_rho = (8 * S + 10 * Nc + 10) * n_atoms
    bytes_force = 57 * n_atoms
    achieved = flops_force / t_force / 1e12
    roofline = {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "roof_note": "compute roof: the FP64 vector ALU (the kernel uses no matrix instructions; MI355X's FP64 "
                             "vector and matrix peaks are both 78.6 TFLOP/s)",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": load_traffic(),
                "kernel": "k_force<FP64>", "kernel_ms": round(t_force * 1e3, 4),
                "flops_per_atom": round(flops_atom, 1), "alg_bytes_per_atom": 57,
                "hbm_achieved_GBps": round(bytes_force / t_force / 1e9, 1),
                "hbm_frac": round(bytes_force / t_force / 1e9 / PEAK_HBM_GBPS, 4),
                "rho_df_kernel_ms": round(t_rho * 1e3, 4),
                "rho_df_tflops": round(flops_rho / t_rho / 1e12, 3)}
    out = {"metric": "atom-steps/s (FP64 EAM velocity-Verlet step)", "value": round(value, 1),
           "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(r64["seconds"] / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"BCC {args.cells}^3 cells, 3-type alloy 90/5/5, synthetic EAM tables (A, B, C "
                                  f"+-10%), 800 K, dt 1 fs, FP64",
                      "atoms": n_atoms, "stencil_entries": S, "pairs_in_cutoff": round(Nc, 2),
                      "energy_every": ENERGY_EVERY},
           "roofline": roofline,
           "kernel_ms_per_step": {k: round(ms[i] / max(args.steps, 1), 4) for i, k in
                                  enumerate(("rho_df", "force", "integrate", "ghost", "interstitial"))}}

    if not args.no_half:
        log(f"bench: Newton's-third-law half stencil vs full stencil, FP64 and mixed, {args.half_steps} steps")
        out["half_stencil"] = half_stencil(w, args.half_steps, args.warmup)

    if not args.no_dropin:
        log("bench: drop-in hook path on host records (PCIe included)")
        out["dropin_pcie"] = dropin_hooks(w)

    if not args.no_mixed:
        log(f"bench: mixed vs FP64 on the defect box (0.5% vacancies, 0.2% interstitials, 1200 K), "
            f"{args.mixed_steps} steps")
        wd = make_workload(args.cells, seed=2025, temperature=1200.0, vacancy_frac=0.005, interstitial_frac=0.002)
        out["mixed"] = mixed_vs_fp64(wd, args.mixed_steps, args.warmup, per_step_err=True)
        del wd

    if not args.no_large and not args.no_mixed:
        log(f"bench: large box, 256^3 cells, mixed vs FP64, {args.large_steps} steps")
        wl = make_workload(256, seed=2026, temperature=1800.0, vacancy_frac=0.01, interstitial_frac=0.005,
                           n_blobs=20, blob_sigma=8.0)
        out["large_box"] = mixed_vs_fp64(
            wl, args.large_steps, args.warmup,
            desc=f"BCC 256^3 cells ({n_atoms_of(wl)} atoms), 3-type alloy, 1% vacancies and {wl['itl'].n} "
                 "interstitials (0.5%) clustered in 20 seeded Gaussian blobs (sigma 8 A) at tetrahedral sites, "
                 "1800 K, dt 1 fs, fully device resident with in-place ghost refresh")
        del wl

    if not args.no_fs:
        log("bench: the same steps on a mechanically stable potential (Finnis-Sinclair Fe tables)")
        wf = make_workload(args.cells, potential="fs")
        rf = run_gpu(wf, mmd.PREC_FP64, args.steps, args.warmup)
        mf, nf = rf["kernel_ms"], rf["launches"]
        stable = {
            "workload": f"BCC {args.cells}^3 cells, 3-type alloy, Finnis-Sinclair Fe tables (+-10% per type / pair), "
                        "800 K, dt 1 fs, FP64; the synthetic forms leave the lattice far from equilibrium (it "
                        "disorders within ~100 steps), this potential keeps a thermal BCC crystal",
            "value": round(n_atoms_of(wf) * args.steps / rf["seconds"], 1), "unit": "atom-steps/s",
            "force_kernel_ms": round(mf[mmd.K_FORCE] / max(nf[mmd.K_FORCE], 1), 4),
            "rho_df_kernel_ms": round(mf[mmd.K_RHO] / max(nf[mmd.K_RHO], 1), 4)}
        del wf
        if not args.no_mixed:
            wfd = make_workload(args.cells, seed=2025, temperature=1200.0, vacancy_frac=0.005,
                                interstitial_frac=0.002, potential="fs")
            stable["mixed_defect_box"] = mixed_vs_fp64(wfd, args.mixed_steps, args.warmup, migrate=True)
            del wfd
        out["stable_potential"] = stable

    if not args.no_serve:
        out["serving"] = serving(args.serve_replicas, args.serve_load, use_graph=not args.serve_no_graph,
                                 ghost_cells=args.serve_ghost)
        out["serving"]["capacity_vs_fp64_path"] = round(out["serving"]["capacity_atom_steps_per_s"] / value, 3)

    if not args.no_cpu:
        log(f"bench: CPU oracle baseline, {args.cpu_steps} steps")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_steps, gpu_end=r64.get("state"))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
