"""Serving-loop scheduler on the CPU: the same admission / batching loop the
GPU server runs, driven by the library's simulated executor (virtual clock,
fake device-memory budget, linear cost model).  Checks the policy's
invariants: FIFO admission, slot and arena exclusivity, memory-bounded
admission, exact step accounting, continuous batching (late arrivals join
while earlier replicas run), rejection of replicas larger than the arena.
"""
import numpy as np
import pytest

import mmd

GHOST = 3


def trace(n, seed=1, rate=100.0, lo=10, hi=64, steps=200):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    cells = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(int)
    return [dict(id=1000 + i, cells=int(cells[i]), temperature=float(rng.uniform(300, 1500)), steps=steps,
                 seed=i, arrival_s=float(t[i])) for i in range(n)]


def residency_overlaps(res):
    """Pairs of admitted replicas whose [t_admit, t_done) intervals overlap."""
    adm = [r for r in res if r["slot"] >= 0]
    for i in range(len(adm)):
        for j in range(i + 1, len(adm)):
            a, b = adm[i], adm[j]
            if a["t_admit"] < b["t_done"] and b["t_admit"] < a["t_done"]:
                yield a, b


def test_replica_bytes_layout():
    b10, b11, b64 = (mmd.replica_bytes(c, GHOST) for c in (10, 11, 64))
    assert b10 < b11 < b64
    for c, b in ((10, b10), (64, b64)):
        n_ext = 2 * (c + 2 * GHOST) ** 3
        # x,y,z,type record 32 B + v, f (3 each) + rho + df 64 B + type 1 B per ext site, plus small tables
        assert 97 * n_ext <= b <= 97 * n_ext + 16 * 4096 + 3 * 8 * n_ext // 256 + 8192
        assert b % 256 == 0
    with pytest.raises(mmd.MmdError):
        mmd.replica_bytes(2, GHOST)  # box thinner than the ghost shell


def test_all_complete_fifo_and_exact_steps():
    reps = trace(120)
    res, st = mmd.simulate(reps, arena_bytes=3e9, max_replicas=8, chunk_steps=10)
    assert st["completed"] == len(reps) and st["failed"] == 0
    assert all(r["status"] == 0 for r in res)
    assert st["atom_steps"] == sum(2 * r["cells"] ** 3 * r["steps"] for r in reps)
    order = np.argsort([r["arrival_s"] for r in reps], kind="stable")
    admits = [res[i]["t_admit"] for i in order]
    assert all(a <= b for a, b in zip(admits, admits[1:])), "admission must follow arrival order"
    for r, q in zip(res, reps):
        assert r["id"] == q["id"]
        assert q["arrival_s"] <= r["t_admit"] <= r["t_ready"] < r["t_done"]
        assert r["n_atoms"] == 2 * q["cells"] ** 3
    assert 1 <= st["max_active"] <= 8
    lat = np.array([r["t_done"] - r["t_arrival"] for r in res])
    assert st["latency_p50_s"] == pytest.approx(np.sort(lat)[int(np.ceil(0.5 * len(lat))) - 1])
    assert st["latency_p99_s"] == pytest.approx(np.sort(lat)[int(np.ceil(0.99 * len(lat))) - 1])
    assert st["latency_p50_s"] <= st["latency_p99_s"]


def test_slots_and_arena_exclusive():
    reps = trace(150, seed=3, rate=400.0)
    arena = 1.5e9
    res, st = mmd.simulate(reps, arena_bytes=arena, max_replicas=6, chunk_steps=7)
    assert st["completed"] == len(reps)
    for a, b in residency_overlaps(res):
        assert a["slot"] != b["slot"]
        assert (a["arena_offset"] + a["arena_bytes"] <= b["arena_offset"]
                or b["arena_offset"] + b["arena_bytes"] <= a["arena_offset"])
    for r in res:
        assert r["arena_offset"] + r["arena_bytes"] <= arena
        assert 0 <= r["slot"] < 6
    # live bytes never exceed the arena and live replicas never exceed the slots
    events = sorted({r["t_admit"] for r in res})
    for t in events:
        live = [r for r in res if r["t_admit"] <= t < r["t_done"]]
        assert sum(r["arena_bytes"] for r in live) <= arena
        assert len(live) <= 6


def test_memory_bounds_admission():
    """With room for only two 64-cell replicas, no more than two are resident
    even though eight slots are free."""
    big = mmd.replica_bytes(64, GHOST)
    reps = [dict(cells=64, temperature=800.0, steps=50, seed=i, arrival_s=0.0) for i in range(6)]
    res, st = mmd.simulate(reps, arena_bytes=2 * big + 4096, max_replicas=8, chunk_steps=10)
    assert st["completed"] == 6
    assert st["max_active"] == 2
    for t in sorted({r["t_admit"] for r in res}):
        assert sum(1 for r in res if r["t_admit"] <= t < r["t_done"]) <= 2


def test_too_large_is_rejected_others_run():
    small = mmd.replica_bytes(12, GHOST)
    reps = [dict(cells=12, temperature=500.0, steps=20, seed=0, arrival_s=0.0),
            dict(cells=40, temperature=500.0, steps=20, seed=1, arrival_s=0.0),
            dict(cells=12, temperature=500.0, steps=20, seed=2, arrival_s=0.0)]
    res, st = mmd.simulate(reps, arena_bytes=3 * small, max_replicas=4)
    assert [r["status"] for r in res] == [0, mmd.E_NOMEM, 0]
    assert res[1]["slot"] == -1
    assert st["completed"] == 2 and st["failed"] == 1


def test_continuous_batching_admits_while_running():
    """A replica that arrives while a long one runs is admitted and finishes
    before it: the loop batches, it does not serve one request at a time."""
    reps = [dict(cells=40, temperature=800.0, steps=400, seed=0, arrival_s=0.0),
            dict(cells=10, temperature=800.0, steps=20, seed=1, arrival_s=0.05)]
    res, st = mmd.simulate(reps, arena_bytes=4e9, max_replicas=4, chunk_steps=10, step_atom_s=1e-9)
    long_, short = res
    assert short["t_admit"] < long_["t_done"]
    assert short["t_done"] < long_["t_done"]
    assert st["max_active"] == 2


def test_zero_steps_and_chunk_tail():
    reps = [dict(cells=10, temperature=300.0, steps=0, seed=0, arrival_s=0.0),
            dict(cells=10, temperature=300.0, steps=13, seed=1, arrival_s=0.0),
            dict(cells=11, temperature=300.0, steps=7, seed=2, arrival_s=0.01)]
    res, st = mmd.simulate(reps, arena_bytes=1e9, max_replicas=4, chunk_steps=5)
    assert st["completed"] == 3
    assert st["atom_steps"] == 2 * 1000 * 13 + 2 * 1331 * 7


def test_deterministic():
    reps = trace(60, seed=9, rate=300.0)
    a = mmd.simulate(reps, arena_bytes=1e9, max_replicas=5)
    b = mmd.simulate(reps, arena_bytes=1e9, max_replicas=5)
    assert a == b


def test_invalid_specs_rejected():
    for bad in (dict(cells=2), dict(steps=-1), dict(temperature=-1.0), dict(arrival_s=float("nan"))):
        r = dict(cells=10, temperature=300.0, steps=5, seed=0, arrival_s=0.0)
        r.update(bad)
        with pytest.raises(mmd.MmdError) as e:
            mmd.simulate([r], arena_bytes=1e9)
        assert e.value.code == mmd.E_INVALID
