"""GPU parity of the batched persistent path: device RNG + device Metropolis
per box must reproduce the reference's (and the oracle's) trial-move
sequences, alone and mixed with other boxes; mixed precision stays within
its stated bound; NPT runs reproduce the oracle's moves and volume moves."""
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from gpu_checks import P_ATM, DE_TOL, de_ok, npt_against_oracle, npt_matches_oracle, run_all
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return m.HostBox.from_state(str(GOLDEN / f"{name}.state"))


def test_batch_fp64_reproduces_reference_trace():
    ref = np.load(GOLDEN / "water256.ref.npz")
    steps = int(ref["steps"])
    b = m.Batch(4)
    s = b.add(load("water256"), steps=steps, seed=1, record_trace=True)
    (st,) = run_all(b, [s], steps, 2500)
    tr = b.trace(s)
    assert st.steps_done == steps and st.done
    assert np.array_equal(tr.mol, ref["mol"])
    assert np.array_equal(tr.accepted, ref["accepted"])
    assert np.max(np.abs(tr.de - ref["de"])) <= DE_TOL
    assert st.start_energy == pytest.approx(float(ref["e0"]), rel=1e-12)
    assert st.energy == pytest.approx(float(ref["final"]), rel=1e-12, abs=1e-7)
    assert st.accepted == int(ref["n_accepted"])


def test_batch_mix_each_box_equals_its_own_oracle_run():
    """Boxes of different kinds, sizes and seeds in one launch; each one's
    trace must equal the oracle run of that box alone."""
    steps = 1500
    boxes = [("water256", 1), ("water256", 7), ("bicyclic200", 3),
             ("mixture300", 11), ("water128_mp", 5)]
    b = m.Batch(8)
    slots, want = [], []
    for name, seed in boxes:
        hb = load(name)
        ob = O.OracleBox(hb.arrays())
        _, _, _, t = ob.run_nvt(seed, steps, use_libc=False)
        want.append(t)
        slots.append(b.add(hb, steps=steps, seed=seed, record_trace=True))
    run_all(b, slots, steps, 500)
    for s, w, (name, seed) in zip(slots, want, boxes):
        tr = b.trace(s)
        assert np.array_equal(tr.mol, w["mol"]), name
        assert np.array_equal(tr.accepted, w["accepted"]), name
        assert de_ok(tr.de, w), name


def test_batch_cells_match_brute_force_trace():
    hb = m.workload("bicyclic4000", n_molecules=1500)
    steps = 1500
    ob = O.OracleBox(hb.arrays())
    _, _, _, w = ob.run_nvt(hb.mc_seed, steps, use_libc=False)
    b = m.Batch(2)
    s = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
    run_all(b, [s], steps, 750)
    tr = b.trace(s)
    assert np.array_equal(tr.mol, w["mol"])
    assert np.array_equal(tr.accepted, w["accepted"])
    assert de_ok(tr.de, w)


def test_batch_mixed_precision_error():
    ref = np.load(GOLDEN / "water256.ref.npz")
    steps = int(ref["steps"])
    b = m.Batch(2)
    s = b.add(load("water256"), steps=steps, seed=1, precision=m.MIXED,
              record_trace=True)
    run_all(b, [s], steps, 5000)
    tr = b.trace(s)
    same = np.cumprod(tr.accepted == ref["accepted"]).astype(bool)
    n = int(same.sum())
    err = np.abs(tr.de[:n] - ref["de"][:n])
    print(f"batch mixed water256: {n}/{steps} aligned, max {err.max():.3e}, "
          f"p99 {np.quantile(err, 0.99):.3e}, median {np.median(err):.3e}")
    # accept/reject identical to the reference on every move of config 0
    assert n == steps
    assert np.array_equal(tr.mol, ref["mol"])
    assert err.max() <= 1e-4


def test_batch_npt_trace_matches_oracle():
    """Water/methanol mixture, all-pairs, FP64: 19 volume moves, some
    accepted and some rejected, each equal to the oracle's."""
    vg = npt_against_oracle(load("mixture300"), 2000, 100, 0.002, m.NEIGHBOR_BRUTE,
                            seed=1)
    assert 0 < vg["accepted"].sum() < len(vg["accepted"])


def test_batch_volume_moves_beside_running_boxes_under_a_tick_budget():
    """Budgeted ticks stop NPT boxes at their volume moves mid-tick; the next
    tick runs those volume moves beside the other boxes' moves and then moves
    the NPT boxes until the same deadline.  Every box's moves and volume
    moves must equal its own oracle run."""
    steps, vi, delta = 1500, 100, 0.002
    b = m.Batch(8)
    b.set_tick_budget(250.0)
    npt, nvt = [], []
    for name, seed in (("mixture300", 1), ("water256", 5)):
        hb = load(name)
        a = hb.arrays()
        v0 = float(a["box"][0] * a["box"][1] * a["box"][2])
        s = b.add(hb, steps=steps, seed=seed, ensemble=m.NPT, volume_interval=vi,
                  pressure=P_ATM, max_volume_delta=delta, record_trace=True)
        npt.append((s, hb, seed, delta * v0))
    for name, seed in (("water256", 7), ("bicyclic200", 3), ("water128_mp", 11)):
        hb = load(name)
        nvt.append((b.add(hb, steps=steps, seed=seed, record_trace=True), hb, seed))
    slots = [x[0] for x in npt + nvt]
    for _ in range(5000):  # ~40 moves per tick
        if all(b.stats(s).done for s in slots):
            break
        b.run(2000)
    else:
        raise AssertionError("boxes did not finish")
    for s, hb, seed, dv in npt:
        vg = npt_matches_oracle(b, s, hb, steps, vi, dv, seed)
        assert 0 < vg["accepted"].sum() < len(vg["accepted"])
    for s, hb, seed in nvt:
        _, _, _, t = O.OracleBox(hb.arrays()).run_nvt(seed, steps, use_libc=False)
        tr = b.trace(s)
        assert np.array_equal(tr.mol, t["mol"]) and np.array_equal(tr.accepted, t["accepted"])
        assert de_ok(tr.de, t)
    b.close()


def test_batch_remove_npt_boxes_at_their_volume_stop_and_reuse_the_slots():
    """An NPT box removed while it waits at its volume stop (between ticks)
    and one removed while the tick in flight runs its volume move and moves
    after it; new boxes then take the freed slots and memory.  The boxes
    that stay and the new ones must equal their own oracle runs."""
    hb = load("water256")
    npt = dict(ensemble=m.NPT, volume_interval=100, pressure=P_ATM, max_volume_delta=0.002)
    b = m.Batch(4)
    a1 = b.add(hb, steps=1000, seed=1, **npt)
    a2 = b.add(hb, steps=1000, seed=2, **npt)
    keep = b.add(hb, steps=600, seed=7, record_trace=True)
    b.run(100)  # both NPT boxes stop at their first volume move
    assert b.stats(a1).steps_done == 100 and b.stats(a1).volume_moves == 0
    b.remove(a1)  # nothing of its volume move is queued yet
    c = b.add(hb, steps=300, seed=3, record_trace=True)
    b.run(50, sync=False)  # a2 takes its volume move, then moves, in this tick
    b.remove(a2)  # waits for that tick
    d = b.add(hb, steps=300, seed=4, record_trace=True)
    run_all(b, [keep, c, d], 600, 100)
    for s, seed, steps in ((keep, 7, 600), (c, 3, 300), (d, 4, 300)):
        fin, _, nacc, t = O.OracleBox(hb.arrays()).run_nvt(seed, steps, use_libc=False)
        tr = b.trace(s)
        assert np.array_equal(tr.mol, t["mol"]) and np.array_equal(tr.accepted, t["accepted"])
        assert de_ok(tr.de, t)
        st = b.stats(s)
        assert st.steps_done == steps and st.accepted == nacc
        assert st.energy == pytest.approx(fin, rel=1e-11, abs=1e-7)
    b.close()


def test_batch_npt_volume_moves_that_change_the_cell_grid():
    """1,500 waters in a 35.6 A box hold 3 cells of 9 A per edge; the box,
    generated from a lattice, expands under volume steps of up to 10% of V0
    and passes 36 A, where 4 cells fit: volume moves build their lists on a
    grid of another size and undone ones swap the previous grid and lists
    back.  Moves and volume moves must equal the oracle's."""
    hb = m.HostBox.generate(m.WATER, 1500, edge=35.6, cutoff=9.0, temperature=298.0,
                            max_translate=0.15, max_rotate=15.0, seed=1500, mc_seed=5)
    assert int(35.6 // 9.0) == 3
    vg = npt_against_oracle(hb, 500, 25, 0.1, m.NEIGHBOR_CELL)
    edges = np.cbrt(vg["v_old"])  # the box's edge at each volume move
    print(f"volume moves: {len(vg['accepted'])}, accepted {int(vg['accepted'].sum())}, "
          f"edges {edges.min():.2f}-{edges.max():.2f} A")
    assert edges.min() < 36.0 < edges.max()


def test_batch_npt_large_molecules_with_cells():
    """NPT on 16-site molecules with 5 cells per edge: the volume moves'
    full energies take the parity pair split over all 27 cells."""
    vg = npt_against_oracle(m.workload("bicyclic4000", n_molecules=1500), 1000, 250,
                            0.002, m.NEIGHBOR_CELL)
    assert 0 < vg["accepted"].sum() < len(vg["accepted"])


def test_batch_npt_mixture20000_mixed_precision():
    """Config 3 at full size: 12,000 water + 8,000 methanol, cell lists, FP32
    pair terms, a volume move every 500 moves.  Moves and volume decisions
    equal the FP64 oracle's; the FP32-pair full energies stay within 1e-6
    relative (measured ~3e-8)."""
    npt_against_oracle(m.workload("mixture20000"), 2000, 500, 0.001, m.NEIGHBOR_CELL,
                       precision=m.MIXED, e_rel=1e-6)


def test_batch_admission_by_footprint():
    hb = load("water256")
    import ctypes as C
    d = hb.desc()
    p = m.RunParams(1, 0, 0, 0, 1000, 0.0, 0.0, 0.0, 0, 0)
    need = C.c_size_t()
    m.check(m.lib().mcgpu_box_footprint(C.byref(d), C.byref(p), C.byref(need)))
    b = m.Batch(8, hbm_budget=int(need.value * 2.5))
    s1 = b.add(hb, steps=1000)
    s2 = b.add(hb, steps=1000)
    with pytest.raises(m.McgpuError, match="budget"):
        b.add(hb, steps=1000)
    b.remove(s1)
    s3 = b.add(hb, steps=1000)
    b.run(100)
    assert b.stats(s3).steps_done == 100 and b.stats(s2).steps_done == 100


def test_batch_pipelined_ticks_peek_remove_and_admit():
    """The serving loop's order: launch a tick without waiting, read the
    consumed tick's statistics (peek), remove a box that the running tick
    does not run and admit another into the freed memory while the tick
    runs.  Every box must still equal its own oracle run."""
    import ctypes as C
    hb = load("water256")
    d = hb.desc()

    def need(steps):
        p = m.RunParams(1, 0, 0, 0, steps, 0.0, 0.0, 0.0, 0, 1)
        n = C.c_size_t()
        m.check(m.lib().mcgpu_box_footprint(C.byref(d), C.byref(p), C.byref(n)))
        return n.value

    def oracle(seed, steps):
        return O.OracleBox(hb.arrays()).run_nvt(seed, steps, use_libc=False)

    b = m.Batch(4, hbm_budget=need(900) + int(1.5 * need(300)))
    s1 = b.add(hb, steps=300, seed=1, record_trace=True)
    s2 = b.add(hb, steps=900, seed=7, record_trace=True)
    with pytest.raises(m.McgpuError, match="budget"):
        b.add(hb, steps=300, seed=3, record_trace=True)
    b.run(300, sync=False)  # tick 1: box 1 finishes
    fin1, e01, nacc1, _ = oracle(1, 300)
    p = b.peek(s1)  # admitted just before the tick in flight
    assert p.steps_done == 0 and not p.done
    assert p.start_energy == p.energy == pytest.approx(e01, rel=1e-11)
    b.run(300, sync=False)  # consumes tick 1, launches tick 2 (box 2 only)
    p1, p2 = b.peek(s1), b.peek(s2)
    assert p1.done and p1.steps_done == 300 and p2.steps_done == 300 and not p2.done
    assert p1.accepted == nacc1
    assert p1.energy == pytest.approx(fin1, rel=1e-11, abs=1e-7)
    b.remove(s1)
    s3 = b.add(hb, steps=300, seed=3, record_trace=True)
    p3 = b.peek(s3)
    assert p3.steps_done == 0 and p3.start_energy == pytest.approx(e01, rel=1e-11)
    st2, st3 = run_all(b, [s2, s3], 900, 300)
    assert st2.steps_done == 900 and st3.steps_done == 300
    for s, seed, steps in ((s2, 7, 900), (s3, 3, 300)):
        fin, e0, nacc, t = oracle(seed, steps)
        tr = b.trace(s)
        assert np.array_equal(tr.mol, t["mol"]) and np.array_equal(tr.accepted, t["accepted"])
        assert de_ok(tr.de, t)
        st = b.stats(s)
        assert st.accepted == nacc
        assert st.energy == pytest.approx(fin, rel=1e-11, abs=1e-7)
    b.close()


def test_batch_cells_equal_brute_over_boundary_crossings():
    """A box with 8 cells per dimension run long enough that many moves
    carry a molecule's primary across the periodic boundary (the cell range
    then spans the wrap): cell-list and all-molecule traces must agree and
    no box may stop on a device fault."""
    hb = m.HostBox.generate(m.BICYCLIC, 7000, density=1.10, cutoff=12.0, seed=91,
                            mc_seed=17)
    steps = 20000
    b = m.Batch(4)
    s_cell = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
    s_brute = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_BRUTE, record_trace=True)
    st = run_all(b, [s_cell, s_brute], steps, 5000)
    assert all(x.fault == 0 for x in st)
    tc, tb = b.trace(s_cell), b.trace(s_brute)
    assert np.array_equal(tc.mol, tb.mol)
    assert np.array_equal(tc.accepted, tb.accepted)
    assert de_ok(tc.de, tb.de)


def test_batch_cells_beyond_the_lds_occupancy_copy():
    """A grid of 11^3 = 1,331 cells (water, rc 5 A): more cells than a launch
    keeps occupancies for in LDS (1,024), so the moves read and update the
    occupancies in global memory; the trace must equal the all-molecule
    trace, with no fallback to the all-molecule scan.  A 4 x 4 x 4 grid of
    the same kind takes the LDS copy, checked the same way."""
    for n, want_cells in ((6000, 11 ** 3), (400, 4 ** 3)):
        hb = m.HostBox.generate(m.WATER, n, density=1.0, cutoff=5.0, seed=92 + n,
                                mc_seed=5)
        d = hb.desc()
        assert np.prod([int(np.floor(L / 5.0)) for L in d.box]) == want_cells
        steps = 6000
        b = m.Batch(4)
        s_cell = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
        s_brute = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_BRUTE, record_trace=True)
        st = run_all(b, [s_cell, s_brute], steps, 2000)
        assert all(x.fault == 0 for x in st)
        assert st[0].neighbor_fallbacks == 0
        tc, tb = b.trace(s_cell), b.trace(s_brute)
        assert np.array_equal(tc.mol, tb.mol)
        assert np.array_equal(tc.accepted, tb.accepted)
        assert de_ok(tc.de, tb.de)
        b.close()


def test_batch_autotuner_picks_a_size_and_keeps_results():
    """With a tick budget, the first box of a class runs one full tick at each
    work-group size and the class keeps the size with the most moves per tick.
    The box's moves and decisions do not depend on the sizes it ran at: its
    trace equals a run at the density-estimate size without tuning."""
    hb = m.HostBox.generate(m.WATER, 1000, density=0.997, cutoff=9.0, seed=5, mc_seed=3)
    steps = 3000
    b = m.Batch(4)
    b.set_tick_budget(300.0)
    s = b.add(hb, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
    run_all(b, [s], steps, 2000)
    tuned = b.tuning()
    assert len(tuned) == 1
    _, lanes, per_tick = tuned[0]
    print(f"tuned: {lanes} lanes, moves per tick at 256/512/1024: {per_tick}")
    assert all(x > 0 for x in per_tick)
    assert lanes == (256, 512, 1024)[int(np.argmax(per_tick))]
    assert b.group_threads(s) == lanes
    ref = m.Batch(2)
    ref.set_autotune(False)
    r = ref.add(hb, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
    run_all(ref, [r], steps, 2000)
    t1, t0 = b.trace(s), ref.trace(r)
    assert np.array_equal(t1.mol, t0.mol)
    assert np.array_equal(t1.accepted, t0.accepted)
    assert de_ok(t1.de, t0.de)
