"""Serving loop: jobs of different kinds, sizes, precisions and ensembles
arrive over time and share the device (continuous batching).  Each job's
result must equal the same job run alone; a bad job fails alone; the HBM
budget bounds residency without losing jobs."""
import ctypes as C

import numpy as np
import pytest

import mcgpu_amd as m

pytestmark = pytest.mark.gpu


def job_boxes():
    return [
        (m.HostBox.generate(m.WATER, 128, density=0.997, cutoff=9.0, seed=31, mc_seed=3),
         dict(steps=3000, precision=m.FP64)),
        (m.HostBox.generate(m.WATER, 700, density=1.02, cutoff=9.0, seed=32, mc_seed=4),
         dict(steps=1200, precision=m.MIXED, neighbor=m.NEIGHBOR_CELL)),
        (m.HostBox.generate(m.BICYCLIC, 300, density=1.10, cutoff=12.0, seed=33, mc_seed=5),
         dict(steps=800, precision=m.MIXED)),
        (m.HostBox.generate(m.WATER, 300, density=0.97, cutoff=9.0, seed=34, mc_seed=6),
         dict(steps=2000, precision=m.FP64, ensemble=m.NPT, volume_interval=500,
              pressure=1.4584e-5, max_volume_delta=0.01)),
        (m.HostBox.generate(m.MIXTURE, 400, n_second=150, density=0.9, cutoff=11.0,
                            seed=35, mc_seed=7),
         dict(steps=1500, precision=m.FP64, neighbor=m.NEIGHBOR_CELL)),
        (m.HostBox.generate(m.WATER, 64, edge=12.4, cutoff=6.0, seed=36, mc_seed=8),
         dict(steps=5000, precision=m.MIXED)),
    ]


def alone(box, kw):
    b = m.Batch(2)
    s = b.add(box, **kw)
    for _ in range(10000):
        st = b.stats(s)
        if st.done:
            return st
        b.run(250)
    raise AssertionError("did not finish")


def test_each_job_same_alone_and_in_the_mix():
    jobs = job_boxes()
    want = [alone(box, kw) for box, kw in jobs]
    # a per-launch time limit splits each job's moves differently over
    # launches than the run alone; the results must not change
    srv = m.Server(max_boxes=8, sweeps_per_tick=250, hbm_budget=1 << 30,
                   tick_budget_us=300.0)
    for k, (box, kw) in enumerate(jobs * 2):   # every job twice, arrivals spread
        srv.submit(box, arrival_s=0.004 * k, **kw)
    srv.run()
    res = srv.results()
    assert len(res) == 2 * len(jobs)
    for r in res:
        w = want[r.job_id % len(jobs)]
        assert r.status == 0
        assert r.steps == w.steps_done
        assert r.accepted == w.accepted
        # bit-identical in practice (same kernel, same per-box work-group);
        # the stated tolerance is the FP64 bound
        assert r.final_energy == pytest.approx(w.energy, rel=1e-12, abs=1e-9)
        assert r.finish_s >= r.admit_s >= r.arrival_s
    st = srv.stats()
    assert st.total_moves == 2 * sum(kw["steps"] for _, kw in jobs)
    assert st.peak_resident >= 2   # jobs overlapped on the device


def test_bad_job_fails_alone():
    jobs = job_boxes()[:2]
    srv = m.Server(max_boxes=4, sweeps_per_tick=200, hbm_budget=1 << 28)
    srv.submit(jobs[0][0], arrival_s=0.0, **jobs[0][1])
    # NPT without a volume interval is rejected at admission
    srv.submit(jobs[1][0], arrival_s=0.001, steps=100, ensemble=m.NPT, volume_interval=0)
    srv.submit(jobs[1][0], arrival_s=0.002, **jobs[1][1])
    srv.run()
    res = srv.results()
    assert res[0].status == 0 and res[2].status == 0
    assert res[1].status == -1
    assert res[2].steps == jobs[1][1]["steps"]


def test_budget_limits_residency_without_losing_jobs():
    box = m.HostBox.generate(m.WATER, 200, density=0.997, cutoff=9.0, seed=40)
    d = box.desc()
    p = m.RunParams(1, 0, 0, 0, 500, 0.0, 0.0, 0.0, 0, 0)
    need = C.c_size_t()
    m.check(m.lib().mcgpu_box_footprint(C.byref(d), C.byref(p), C.byref(need)))
    srv = m.Server(max_boxes=16, sweeps_per_tick=100, hbm_budget=int(need.value * 3.5))
    for k in range(10):
        srv.submit(box, arrival_s=0.0, steps=500, seed=k + 1)
    srv.run()
    res = srv.results()
    assert all(r.status == 0 and r.steps == 500 for r in res)
    assert srv.stats().peak_resident == 3
