"""GPU parity on the benchmark workloads themselves.

512 water boxes of 64-4,096 molecules advanced together in one batch: the
FP64 traces of sampled boxes equal the oracle's runs of those boxes alone,
and their FP32-pair twins stay within 1e-4 kcal/mol of them.

A slice of the serving stream (Poisson arrivals, water and 16-site
bicyclic jobs of 128-16,384 molecules, NVT/NPT, FP64/FP32 pairs; steps
capped) through the continuous-batching server: every job equals its run
alone, and the FP64 jobs' accepted counts, volume moves and final energies
equal the oracle's."""
import numpy as np
import pytest

import bench
import mcgpu_amd as m
from gpu_checks import de_ok, run_all
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C2_STEPS = 300
STREAM_FIRST = 260
STREAM_JOBS = 40
STREAM_STEP_CAP = 1100  # two volume moves for the NPT jobs


def test_512_water_boxes_in_one_batch():
    boxes = bench.c2_boxes(512)
    sizes = np.array([hb.desc().n_molecules for hb in boxes])
    order = np.argsort(sizes)
    sample = [int(order[k]) for k in np.linspace(0, len(boxes) - 1, 8).astype(int)]
    b = m.Batch(len(boxes) + len(sample))
    slots, twins = {}, {}
    for i, hb in enumerate(boxes):
        slots[i] = b.add(hb, steps=C2_STEPS, seed=1 + i, precision=m.FP64,
                         neighbor=bench.neighbor_for(hb), record_trace=i in sample)
    for i in sample:
        twins[i] = b.add(boxes[i], steps=C2_STEPS, seed=1 + i, precision=m.MIXED,
                         neighbor=bench.neighbor_for(boxes[i]), record_trace=True)
    stats = run_all(b, list(slots.values()) + list(twins.values()), C2_STEPS, C2_STEPS)
    assert all(st.steps_done == C2_STEPS and st.fault == 0 for st in stats)
    for i in sample:
        hb = boxes[i]
        ob = O.OracleBox(hb.arrays())
        fin, e0, nacc, want = ob.run_nvt(1 + i, C2_STEPS, use_libc=False)
        st = b.stats(slots[i])
        tr = b.trace(slots[i])
        print(f"box {i}: {sizes[i]} molecules, {nacc} accepted")
        assert np.array_equal(tr.mol, want["mol"]), i
        assert np.array_equal(tr.accepted, want["accepted"]), i
        assert de_ok(tr.de, want), i
        assert st.accepted == nacc
        assert st.energy == pytest.approx(fin, rel=1e-10, abs=1e-7)
        tw = b.trace(twins[i])
        assert np.array_equal(tw.mol, want["mol"]), i
        assert np.array_equal(tw.accepted, want["accepted"]), i
        phys = np.abs(want["de"]) < 1e3
        err = np.abs(tw.de - want["de"])[phys]
        print(f"  FP32-pair twin: max |dE err| {err.max():.3e}")
        assert err.max() <= 1e-4
    b.close()


def stream_slice():
    """Jobs 260-299 of the stream (arrivals from 0), steps capped."""
    jobs = bench.make_stream(STREAM_FIRST + STREAM_JOBS)[STREAM_FIRST:]
    t0 = jobs[0]["arrival"]
    jobs = [dict(j, steps=min(j["steps"], STREAM_STEP_CAP), arrival=j["arrival"] - t0)
            for j in jobs]
    assert any(j["kind"] == m.BICYCLIC and j["n"] == 16384 for j in jobs)
    assert any(j["npt"] and not j["mixed"] for j in jobs)
    return jobs


def alone(box, kw):
    b = m.Batch(2)
    s = b.add(box, **kw)
    (st,) = run_all(b, [s], kw["steps"], 500)
    b.close()
    return st


def test_serving_stream_slice_alone_mix_and_oracle():
    jobs = stream_slice()
    pool = bench.BoxPool()
    srv = m.Server(max_boxes=64, sweeps_per_tick=2000, hbm_budget=16 << 30,
                   tick_budget_us=2000.0)
    for j in jobs:
        box = pool.get(j["kind"], j["n"])
        srv.submit(box, arrival_s=j["arrival"], **bench.run_kwargs(j, box))
    srv.run()
    res = {r.job_id: r for r in srv.results()}
    assert len(res) == len(jobs)
    assert srv.stats().peak_resident >= 4  # jobs overlapped on the device
    checked = 0
    for k, j in enumerate(jobs):
        box = pool.get(j["kind"], j["n"])
        kw = bench.run_kwargs(j, box)
        r = res[k]
        w = alone(box, kw)
        assert r.status == 0, k
        assert r.steps == w.steps_done == j["steps"], k
        assert r.accepted == w.accepted, k
        assert r.final_energy == pytest.approx(w.energy, rel=1e-12, abs=1e-9), k
        # FP64 jobs against the oracle (NPT ones up to 5,000 molecules)
        if j["mixed"] or (j["npt"] and j["n"] > 5000):
            continue
        a = box.arrays()
        ob = O.OracleBox(a)
        if j["npt"]:
            v0 = float(a["box"][0] * a["box"][1] * a["box"][2])
            fin, e0, nacc, _, v = ob.run_npt(j["mc_seed"], j["steps"], kw["volume_interval"],
                                             kw["pressure"],
                                             kw["max_volume_delta"] * v0, use_libc=False)
            assert w.volume_moves == len(v["step"]), k
            assert w.volume_accepted == int(v["accepted"].sum()), k
        else:
            fin, e0, nacc, _ = ob.run_nvt(j["mc_seed"], j["steps"], use_libc=False)
        print(f"job {k}: {'NPT' if j['npt'] else 'NVT'} {j['n']} molecules, "
              f"{nacc} accepted, E {fin:.6f}")
        assert r.accepted == nacc, k
        assert r.start_energy == pytest.approx(e0, rel=1e-11), k
        assert r.final_energy == pytest.approx(fin, rel=1e-9, abs=1e-9 * abs(e0) + 1e-6), k
        checked += 1
    assert checked >= 8
