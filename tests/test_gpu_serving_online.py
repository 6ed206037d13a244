"""The online serving surface: the loop runs on the server's own thread
while a second host thread submits jobs and the test thread polls them and
reads live statistics.  Every job must end with the result of its run
alone, and the FP64 NVT jobs with the oracle's accepted count and final
energy."""
import threading
import time

import numpy as np
import pytest

import mcgpu_amd as m
from gpu_checks import run_all
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def jobs():
    g = m.HostBox.generate
    return [
        (g(m.WATER, 300, density=0.997, cutoff=9.0, seed=51, mc_seed=3),
         dict(steps=4000, precision=m.FP64)),
        (g(m.WATER, 1000, density=1.0, cutoff=9.0, seed=52, mc_seed=4),
         dict(steps=6000, precision=m.MIXED, neighbor=m.NEIGHBOR_CELL)),
        (g(m.BICYCLIC, 250, density=1.10, cutoff=12.0, seed=53, mc_seed=5),
         dict(steps=2500, precision=m.FP64)),
        (g(m.WATER, 500, density=0.98, cutoff=9.0, seed=54, mc_seed=6),
         dict(steps=3000, precision=m.FP64, ensemble=m.NPT, volume_interval=500,
              pressure=1.4584e-5, max_volume_delta=0.002)),
        (g(m.MIXTURE, 400, n_second=150, density=0.9, cutoff=11.0, seed=55, mc_seed=7),
         dict(steps=3000, precision=m.FP64, neighbor=m.NEIGHBOR_CELL)),
        (g(m.BICYCLIC, 2000, density=1.10, cutoff=12.0, seed=56, mc_seed=8),
         dict(steps=3000, precision=m.MIXED, neighbor=m.NEIGHBOR_CELL)),
    ]


def alone(box, kw):
    b = m.Batch(2)
    s = b.add(box, **kw)
    (st,) = run_all(b, [s], kw["steps"], 500)
    b.close()
    return st


def test_online_submit_poll_and_live_stats():
    js = jobs()
    srv = m.Server(max_boxes=8, sweeps_per_tick=400, hbm_budget=1 << 30,
                   tick_budget_us=400.0)
    ids = [srv.submit(box, **kw) for box, kw in js[:2]]   # before the start
    srv.start()
    later = []

    def producer():
        for box, kw in js[2:] * 2:
            later.append((srv.submit(box, **kw), box, kw))
            time.sleep(0.01)

    th = threading.Thread(target=producer)
    th.start()
    seen = {}          # job id -> [states seen], progress seen while running
    clock = []
    while th.is_alive() or any(srv.poll(k).state != m.JOB_DONE
                               for k in ids + [k for k, _, _ in later]):
        for k in ids + [k for k, _, _ in later]:
            r = srv.poll(k)
            s = seen.setdefault(k, dict(states=set(), steps=[]))
            s["states"].add(r.state)
            if r.state == m.JOB_RUNNING:
                s["steps"].append(r.steps_done)
        st = srv.stats()
        clock.append((st.clock_s, st.total_moves, st.resident, st.queued))
        time.sleep(0.001)
    th.join()
    srv.wait()
    srv.stop()
    with pytest.raises(m.McgpuError, match="stopping"):
        srv.submit(js[0][0], steps=10)
    # live statistics moved forward while the loop ran
    assert all(b[0] >= a[0] and b[1] >= a[1] for a, b in zip(clock, clock[1:]))
    assert max(c[2] for c in clock) >= 2          # jobs overlapped on the device
    st = srv.stats()
    assert st.completed == st.submitted == len(ids) + len(later)
    # progress was observed part-way through some job, never going back
    partial = [s["steps"] for s in seen.values() if s["steps"]]
    assert partial and all(np.all(np.diff(p) >= 0) for p in partial)
    assert any(0 < v < max(kw["steps"] for _, kw in js) for p in partial for v in p)
    want = [alone(box, kw) for box, kw in js]
    everything = [(k, js[i][0], js[i][1]) for i, k in enumerate(ids)] + later
    for k, box, kw in everything:
        r = srv.poll(k)
        w = want[[b for b, _ in js].index(box)]
        assert r.state == m.JOB_DONE and r.status == 0, k
        assert r.steps == r.steps_done == kw["steps"], k
        assert r.accepted == w.accepted, k
        assert r.final_energy == pytest.approx(w.energy, rel=1e-12, abs=1e-9), k
        assert r.finish_s >= r.admit_s >= r.arrival_s, k
    for box, kw in js:
        if kw["precision"] != m.FP64 or kw.get("ensemble", m.NVT) != m.NVT:
            continue
        ob = O.OracleBox(box.arrays())
        fin, e0, nacc, _ = ob.run_nvt(box.mc_seed, kw["steps"], use_libc=False)
        w = want[[b for b, _ in js].index(box)]
        assert w.accepted == nacc
        assert w.energy == pytest.approx(fin, rel=1e-9, abs=1e-9 * abs(e0) + 1e-6)
