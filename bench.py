#!/usr/bin/env python3
"""Benchmark of the MI355X energy path on the serving stream.

Workload (the stream the metric is quoted on): seeded Poisson arrivals at
200 jobs/s of Monte Carlo jobs drawn from two generators -- rigid TIP3P water
boxes (density U[0.95, 1.05] g/cm3, T U[250, 400] K, rc 9 A) and rigid
16-site planar bicyclic molecules (1.10 g/cm3, 298 K, rc 12 A) -- with
128-16,384 molecules and 1,000-100,000 steps (both log-uniform), half NVT
and half NPT (1 atm, a volume move every 500 steps), half FP64 and half
FP32-pair arithmetic.  Molecule counts are drawn from 24 log-spaced classes
so every class's box is generated once; each job has its own MC seed.

A step is 100 jobs of one continuous stream: the warm-up takes the first
W x 100 jobs (at most 2000 moves each, all arriving at once) and the timed
region the next K x 100 (K = 20 covers 2,000 jobs, ~44 M moves), served
through the continuous-batching server from the first arrival to the last
completion: value = trial moves / wall seconds, ms_per_step = wall / K.
Beside it:
  - the plain FP64 GPU path (host RNG, one fused trial-move launch per
    move) on a sample of the same stream's jobs, and the ratio;
  - p50/p99 job latency (completion - arrival);
  - the roofline of the dominant kernel (the persistent batch kernel):
    algorithmic bytes from the kernel's own work counters (24 B per
    candidate primary, 24 B per site loaded) / HIP-event launch time, on the
    timed stream and (capacity_achieved) on the saturated capacity run;
  - the reference's own CPU path (oracle/_ref) timed on host cores on a
    bounded sample of the same jobs (kind "reference"), or the oracle
    restatement when the reference build is absent (kind "port");
  - the same with one reference process per host core (up to 16), and the
    host CPU's model;
  - the full-energy kernel's site-pair rate against its VALU roof, and the
    per-move dE error of FP32 pair terms against FP64 on a box of each
    configuration;
  - configurations 0, 1 and 3 one box at a time: plain FP64 against the
    batched loop, and the move kernel's algorithmic GB/s (key "per_config");
  - 512 water boxes of 64-4,096 molecules advanced together for 20,000
    moves each (key "batch512").
Separate modes print one JSON object each and skip the stream:
--full-energy-only, --precision-only, --per-config-only, and --batch-512
(512 water boxes of 64-4,096 molecules advanced together for 20,000 moves
each, against the plain path).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("MC trial moves/s, batched FP32/FP64-mixed vs plain FP64 on 1 GPU; "
          "p50/p99 job latency")
HBM_PEAK_GBS = 8000.0
LAMBDA = 200.0
N_CLASSES = 24
JOBS_PER_STEP = 100


def log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def make_stream(n_jobs: int, seed: int = 2024):
    """The job stream: arrival times and per-job parameters (seeded).  Each
    job's values are drawn in turn, so a longer stream extends a shorter
    one: make_stream(n)[:k] == make_stream(k)."""
    import mcgpu_amd as m
    rng = np.random.default_rng(seed)
    classes = np.unique(np.round(np.exp(np.linspace(math.log(128), math.log(16384),
                                                     N_CLASSES))).astype(int))
    jobs = []
    t = 0.0
    for k in range(n_jobs):
        if k:
            t += float(rng.exponential(1.0 / LAMBDA))
        kind = m.WATER if rng.random() < 0.5 else m.BICYCLIC
        n = int(log_uniform(rng, 128, 16384))
        n = int(classes[np.argmin(np.abs(np.log(classes) - math.log(n)))])
        steps = int(log_uniform(rng, 1e3, 1e5))
        npt = rng.random() < 0.5
        mixed = rng.random() < 0.5
        jobs.append(dict(kind=kind, n=n, steps=steps, npt=npt, mixed=mixed,
                         mc_seed=int(rng.integers(1, 2**31 - 1)), arrival=t))
    return jobs


class BoxPool:
    """One generated box per (kind, size class), reused by every job of that
    class (each job has its own MC seed)."""

    def __init__(self, seed: int = 7):
        self.boxes = {}
        self.rng = np.random.default_rng(seed)

    def get(self, kind: int, n: int):
        import mcgpu_amd as m
        key = (kind, n)
        if key not in self.boxes:
            if kind == m.WATER:
                rho = float(self.rng.uniform(0.95, 1.05))
                temp = float(self.rng.uniform(250, 400))
                hb = m.HostBox.generate(m.WATER, n, density=rho, cutoff=9.0,
                                        temperature=temp, max_translate=0.15,
                                        max_rotate=15.0, seed=1000 + n)
            else:
                hb = m.HostBox.generate(m.BICYCLIC, n, density=1.10, cutoff=12.0,
                                        temperature=298.0, max_translate=0.15,
                                        max_rotate=15.0, seed=2000 + n)
            self.boxes[key] = hb
        return self.boxes[key]


def neighbor_for(box) -> int:
    """Cell lists when they shrink the scan, else the all-pairs scan (both
    give the same moves and energies): cells once every dimension holds at
    least 4 cells of width rc, so that the 27 cells around a molecule cover
    at most 27/64 of the box, and the box holds at least 1,000 molecules.
    With 3 cells per edge the 27 cells are the whole box and the cell
    bookkeeping only adds time.  Measured per move (FP64, one box alone, its
    tuned work-group size): water of 692-1,303 molecules (3 cells) 7.7-9.4 us
    all-pairs against 8.7-10.7 us with cells, bicyclic of 298-560 (3 cells)
    13.3-14.3 against 14.4-15.4 us; with 4 cells, water of 1,609 molecules
    9.4 against 8.8 us and of 2,454 11.1 against 9.6 us, bicyclic of 692 13.3
    against 14.2 us and of 1,055 15.6 against 15.3 us."""
    import mcgpu_amd as m
    d = box.desc()
    cells = min(d.box) >= 4 * d.cutoff and d.n_molecules >= 1000
    return m.NEIGHBOR_CELL if cells else m.NEIGHBOR_BRUTE


def run_kwargs(job, box):
    import mcgpu_amd as m
    kw = dict(steps=job["steps"], seed=job["mc_seed"],
              precision=m.MIXED if job["mixed"] else m.FP64,
              neighbor=neighbor_for(box))
    if job["npt"]:
        kw.update(ensemble=m.NPT, volume_interval=500, pressure=1.4584e-5,
                  max_volume_delta=0.01)
    return kw


def serve(jobs, pool, t_shift=True, sweeps=200, max_boxes=1024,
          budget=64 << 30, tick_us=0.0):
    import mcgpu_amd as m
    srv = m.Server(max_boxes=max_boxes, sweeps_per_tick=sweeps, hbm_budget=budget,
                   tick_budget_us=tick_us)
    t0 = jobs[0]["arrival"] if t_shift else 0.0
    for j in jobs:
        box = pool.get(j["kind"], j["n"])
        srv.submit(box, arrival_s=j["arrival"] - t0, **run_kwargs(j, box))
    srv.run()
    return srv.results(), srv.stats()


def plain_sample(jobs, pool, budget_s: float, max_moves: int = 2000):
    """Plain FP64 path on jobs of the stream in arrival order (each job run
    for min(steps, max_moves) moves) until budget_s of loop time."""
    import mcgpu_amd as m
    moves, secs, used = 0, 0.0, 0
    for j in jobs:
        box = pool.get(j["kind"], j["n"])
        with tempfile.TemporaryDirectory() as td:
            p = Path(td) / "b.state"
            box.write_state(str(p))
            hb = m.HostBox.from_state(str(p))
        steps = min(j["steps"], max_moves)
        res, _ = m.plain_run(hb, steps, precision=m.FP64,
                             neighbor=neighbor_for(hb), trace=False)
        moves += steps
        secs += res.loop_seconds
        used += 1
        if secs >= budget_s:
            break
    return moves / secs, used, moves


def _ref_loop_seconds(out: str) -> float:
    return float(re.search(r"Loop Seconds: ([0-9.eE+-]+)", out).group(1))


def cpu_baseline(jobs, pool, budget_s: float, max_cores: int = 16):
    """The reference's own CPU path (oracle/_ref/ref_trace, brute force as in
    its -s mode) on the first jobs of the stream, one per host core up to
    `max_cores`, each for about budget_s / jobs seconds of loop (moves
    estimated from its sites): run one after another on one thread (the
    baseline of record), then all at once, one process per job (jobs are
    independent boxes; aggregate = all moves / the longest loop).  Falls back
    to the oracle restatement, single thread only, when the reference build
    is absent."""
    import mcgpu_amd as m
    ref = ROOT / "oracle" / "_ref" / "ref_trace"
    kind = "reference" if ref.exists() else "port"
    cores = max(1, min(max_cores, os.cpu_count() or 1))
    sample = jobs[:cores]
    per_job = budget_s / len(sample)
    with tempfile.TemporaryDirectory() as td:
        runs = []
        for k, j in enumerate(sample):
            box = pool.get(j["kind"], j["n"])
            d = box.desc()
            # brute force costs ~ N x sites per move: ~4e6 / N moves/s for
            # 3-site water on one core
            rate = 1.2e7 / (d.n_molecules * (d.n_sites / d.n_molecules))
            steps = int(max(20, min(j["steps"], rate * per_job)))
            p = Path(td) / f"b{k}.state"
            box.write_state(str(p))
            runs.append((j, steps, p))
        moves = sum(st for _, st, _ in runs)
        secs = 0.0
        for j, steps, p in runs:
            if kind == "reference":
                out = subprocess.run([str(ref), str(p), str(steps), "-"],
                                     capture_output=True, text=True, check=True).stdout
                secs += _ref_loop_seconds(out)
            else:
                from oracle import oracle as O
                ob = O.OracleBox(m.HostBox.from_state(str(p)).arrays())
                t = time.perf_counter()
                ob.run_nvt(j["mc_seed"], steps, use_libc=True)
                secs += time.perf_counter() - t
        out = dict(value=moves / secs, unit="moves/s", cores=1, kind=kind,
                   sample=(f"first {len(runs)} jobs of the stream, {moves} moves in "
                           f"total (each capped to ~{per_job:.1f} s of loop), brute "
                           f"force, single thread, {secs:.1f} s of loop time"))
        if kind == "reference" and len(runs) > 1:
            procs = [subprocess.Popen([str(ref), str(p), str(steps), "-"],
                                      stdout=subprocess.PIPE, text=True)
                     for _, steps, p in runs]
            loops = []
            for pr in procs:
                o, _ = pr.communicate()
                if pr.returncode != 0:
                    raise RuntimeError("reference CPU run failed")
                loops.append(_ref_loop_seconds(o))
            out["all_cores"] = dict(
                value=moves / max(loops), unit="moves/s", cores=len(runs), kind=kind,
                sample=(f"the same {len(runs)} jobs and moves, one process per job, all "
                        f"at once; the longest loop took {max(loops):.1f} s"))
    return out


def cpu_model() -> str:
    """The host CPU's model name (what lscpu prints as 'Model name')."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# VALU issue slots per site pair in the full-energy kernel's inner loop
# (FP64 path without the per-pair wrap, read from this build's gfx950 code
# object: 27 full-rate VALU instructions -- displacement 3, r^2 3, the rsq
# refinement 5, LJ and Coulomb 10, the r^2 == 0 guard 3, the sum 1, two
# address steps -- plus v_rsq_f64 at quarter rate = 4 slots); with 256 CUs
# x 4 SIMDs x 16 lanes per cycle at 2.4 GHz this sets the pair-interaction
# roof.
VALU_PER_PAIR_FP64 = 31
VALU_LANE_OPS_PER_S = 256 * 4 * 16 * 2.4e9


def full_energy_measure():
    """The full-system energy kernel (one wave per molecule) on the
    20,000-molecule water/methanol mixture (rc 11 A, the NPT volume-move
    workload) and the 4,000-molecule bicyclic box (rc 12 A): site-pair
    interactions per second in FP64 and with FP32 pair terms, against the
    VALU roof, and the FP32-pair relative error of the total."""
    import mcgpu_amd as m
    out = {}
    for name in ("mixture20000", "bicyclic4000"):
        hb = m.workload(name)
        res = {}
        for prec, tag in ((m.FP64, "fp64"), (m.MIXED, "mixed")):
            db = m.DeviceBox(hb, precision=prec)
            db.neighbor(m.NEIGHBOR_CELL)
            db.enable_timing(True)
            e = db.system_energy()
            best = float("inf")
            for _ in range(5):
                e = db.system_energy()
                best = min(best, db.last_kernel_ms())
            res[tag] = (e, db.system_pairs(), best)
            db.close()
        e64, pairs, ms64 = res["fp64"]
        e32, _, ms32 = res["mixed"]
        roof = VALU_LANE_OPS_PER_S / VALU_PER_PAIR_FP64
        out[name] = {
            "site_pairs": int(pairs),
            "fp64_ms": ms64,
            "fp64_pairs_per_s": pairs / (ms64 / 1e3),
            "fp64_frac_of_valu_roof": pairs / (ms64 / 1e3) / roof,
            "mixed_ms": ms32,
            "mixed_pairs_per_s": pairs / (ms32 / 1e3),
            "mixed_rel_err": abs(e32 - e64) / abs(e64),
        }
    out["valu_roof_pairs_per_s"] = VALU_LANE_OPS_PER_S / VALU_PER_PAIR_FP64
    return out


def precision_error_measure(moves: int = 3000):
    """Per-move |dE_mixed - dE_FP64| (kcal/mol) on a box of each
    configuration: the same box and MC seed advanced with FP64 pair terms
    and with FP32 pair terms in one batch, compared over the prefix of moves
    on which both runs chose the same molecules and made the same
    decisions (after the first differing decision the states differ)."""
    import mcgpu_amd as m
    cases = [("c0_water256", m.workload("water256"), m.NEIGHBOR_BRUTE),
             ("c1_bicyclic4000", m.workload("bicyclic4000"), m.NEIGHBOR_CELL),
             ("c3_mixture20000", m.workload("mixture20000"), m.NEIGHBOR_CELL)]
    pool = BoxPool(seed=11)  # the serving stream's water generator
    for n in (64, 256, 1024, 4096):
        hb = pool.get(m.WATER, n)
        cases.append((f"c2_water{n}", hb, neighbor_for(hb)))
    out = {}
    for name, hb, nb in cases:
        b = m.Batch(2)
        s64 = b.add(hb, steps=moves, seed=1, precision=m.FP64, neighbor=nb,
                    record_trace=True)
        s32 = b.add(hb, steps=moves, seed=1, precision=m.MIXED, neighbor=nb,
                    record_trace=True)
        for _ in range(1000):
            if b.stats(s64).done and b.stats(s32).done:
                break
            b.run(2000)
        t64, t32 = b.trace(s64), b.trace(s32)
        b.close()
        same = np.cumprod((t64.mol == t32.mol) & (t64.accepted == t32.accepted)).astype(bool)
        k = int(same.sum())
        err = np.abs(t32.de[:k] - t64.de[:k])
        mag = np.abs(t64.de[:k])
        # trial positions that overlap another molecule reach |dE| ~ 1e19
        # (always rejected), where FP64 rounding alone is ~1e3 kcal/mol:
        # the absolute maximum is reported over |dE| < 1e3, the relative
        # one over all moves
        phys = mag < 1e3
        out[name] = dict(aligned_moves=k, moves=moves,
                         median=float(np.median(err)) if k else None,
                         p99=float(np.quantile(err, 0.99)) if k else None,
                         max_abs_dE_below_1e3=float(err[phys].max()) if phys.any() else None,
                         max_rel=float(np.max(err / np.maximum(mag, 1.0))) if k else None,
                         frac_within_1e_4=float(np.mean(err <= 1e-4)) if k else None)
    return out


def c2_boxes(n_boxes: int = 512, seed: int = 512):
    """The independent water boxes advanced together: 64-4096 molecules
    (log-uniform), density U[0.95, 1.05] g/cm3, T U[250, 400] K, rc 9 A,
    TIP3P (seeded)."""
    import mcgpu_amd as m
    rng = np.random.default_rng(seed)
    boxes = []
    for k, n in enumerate(log_uniform(rng, 64, 4096, n_boxes).astype(int)):
        boxes.append(m.HostBox.generate(
            m.WATER, int(n), density=float(rng.uniform(0.95, 1.05)), cutoff=9.0,
            temperature=float(rng.uniform(250, 400)), max_translate=0.15,
            max_rotate=15.0, seed=5000 + k))
    return boxes


def batch512_measure(n_boxes: int = 512, sweeps: int = 20000, seed: int = 512):
    """512 independent water boxes (c2_boxes) advanced together, every box
    `sweeps` moves: whole-batch moves/s with FP32 pair terms and with FP64,
    against the plain FP64 path on a sample of the same boxes."""
    import mcgpu_amd as m
    boxes = c2_boxes(n_boxes, seed)
    out = {"boxes": n_boxes, "sweeps": sweeps,
           "molecules_total": int(sum(b.desc().n_molecules for b in boxes))}
    # untimed: the first launch of each kernel loads its code object
    w = m.Batch(2)
    for i, hb in enumerate(boxes[:2]):
        w.add(hb, steps=10, seed=1 + i, neighbor=neighbor_for(hb))
    w.run(10)
    w.close()
    for name, prec in (("mixed", m.MIXED), ("fp64", m.FP64)):
        b = m.Batch(n_boxes)
        slots = [b.add(hb, steps=sweeps, seed=1 + i, precision=prec,
                       neighbor=neighbor_for(hb)) for i, hb in enumerate(boxes)]
        m.check(m.lib().mcgpu_device_synchronize())
        t0 = time.perf_counter()
        while not all(b.stats(s).done for s in slots):
            b.run(sweeps)
        dt = time.perf_counter() - t0
        st = [b.stats(s) for s in slots]
        moves = sum(x.steps_done for x in st)
        out[f"{name}_moves_per_s"] = moves / dt
        out[f"{name}_seconds"] = dt
        # move kernel: algorithmic bytes (24 B per candidate primary and per
        # site loaded) over the pass's wall time
        out[f"{name}_algorithmic_gb_per_s"] = 24.0 * sum(
            x.work_candidates + x.work_sites for x in st) / dt / 1e9
        b.close()
    sample = boxes[::32]
    moves = t = 0.0
    for i, hb in enumerate(sample):
        t0 = time.perf_counter()
        m.plain_run(hb, 2000, precision=m.FP64, neighbor=neighbor_for(hb), trace=False)
        t += time.perf_counter() - t0
        moves += 2000
    out["plain_fp64_moves_per_s"] = moves / t
    out["plain_sample"] = f"{len(sample)} of the boxes, 2000 moves each"
    out["ratio_mixed_vs_plain"] = out["mixed_moves_per_s"] / out["plain_fp64_moves_per_s"]
    return out


def per_config_measure(moves: int = 4000):
    """Configurations 0, 1 and 3 one box at a time (NVT; the NPT volume
    moves of configuration 3 are measured by full_energy): the plain FP64
    path (host RNG, one fused trial-move launch per move, PCIe included) and
    the batched device loop on the same box and seed, with FP64 and with
    FP32 pair terms.  For the batched loop: moves per second of kernel time
    (one box holds one CU, so this is the per-move latency) and the move
    kernel's algorithmic GB/s (24 B per candidate primary and per site
    loaded) over that kernel time."""
    import mcgpu_amd as m
    out = {}
    for tag, name in (("c0_water256", "water256"), ("c1_bicyclic4000", "bicyclic4000"),
                      ("c3_mixture20000", "mixture20000")):
        hb = m.workload(name)
        nb = neighbor_for(hb)
        res, _ = m.plain_run(hb, moves, precision=m.FP64, neighbor=nb, trace=False)
        row = {"moves": moves,
               "neighbor": "cells" if nb == m.NEIGHBOR_CELL else "all pairs",
               "plain_fp64_moves_per_s": moves / res.loop_seconds}
        for prec, p in ((m.FP64, "fp64"), (m.MIXED, "mixed")):
            b = m.Batch(2)
            b.set_tick_budget(2000.0)
            s = b.add(hb, steps=moves, precision=prec, neighbor=nb)
            while not b.stats(s).done:
                b.run(2000)
            avg_ms, launches = b.kernel_ms()  # since the batch was made
            ms = avg_ms * launches
            st = b.stats(s)
            row[f"batched_{p}_moves_per_s"] = st.steps_done / (ms / 1e3)
            row[f"batched_{p}_us_per_move"] = 1e3 * ms / st.steps_done
            row[f"batched_{p}_algorithmic_gb_per_s"] = (
                24.0 * (st.work_candidates + st.work_sites) / (ms / 1e3) / 1e9)
            b.close()
        row["ratio_batched_mixed_vs_plain"] = (row["batched_mixed_moves_per_s"]
                                               / row["plain_fp64_moves_per_s"])
        out[tag] = row
    return out


def pmc_traffic_ratio(steps: int, warmup: int):
    """HBM bytes per algorithmic byte of the batch kernel, from the committed
    rocprofv3 counter summary of this same workload (the same --steps and
    --warmup): FETCH_SIZE + WRITE_SIZE of every k_batch dispatch of the
    counted runs, scaled by a copy kernel's readings, over the algorithmic
    bytes those runs' serving loops reported.  None when the profile was
    taken on another workload.  (Counters slow the kernels, so ticks of a
    counted run carry more moves than ticks of this run; the ratio, not the
    bytes per launch, carries over.)"""
    p = ROOT / "profiles" / "k_batch_serving_pmc.txt"
    if not p.exists():
        return None
    text = p.read_text()
    args = re.search(r"bench_args:\s*--steps (\d+) --warmup (\d+)", text)
    val = re.search(r"traffic_per_algorithmic_byte:\s*([0-9.eE+]+)", text)
    if not args or not val or (int(args.group(1)), int(args.group(2))) != (steps, warmup):
        return None
    return float(val.group(1))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help=f"timed steps; a step is {JOBS_PER_STEP} jobs of the stream")
    ap.add_argument("--warmup", type=int, default=5,
                    help=f"untimed warm-up steps ({JOBS_PER_STEP} jobs each, "
                    "at most 2000 moves per job, all arriving at once)")
    ap.add_argument("--sweeps", type=int, default=2000, help="max moves per box per tick")
    ap.add_argument("--tick-us", type=float, default=2000.0,
                    help="per-launch time limit in microseconds")
    ap.add_argument("--plain-seconds", type=float, default=8.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch512", action="store_true",
                    help="skip the 512-water-box measurement")
    ap.add_argument("--full-energy-only", action="store_true",
                    help="measure only the full-energy kernel and print its JSON")
    ap.add_argument("--precision-only", action="store_true",
                    help="measure only the FP32-pair dE error and print its JSON")
    ap.add_argument("--batch-512", action="store_true",
                    help="measure only 512 water boxes advanced together and print its JSON")
    ap.add_argument("--per-config-only", action="store_true",
                    help="measure only configurations 0, 1 and 3 one box at a time "
                    "(plain vs batched, move-kernel GB/s) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (independent boxes would shard across GPUs with no "
              "exchange between them)", file=sys.stderr)
        return 2
    import mcgpu_amd as m
    if m.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.full_energy_only:
        print(json.dumps(full_energy_measure()), flush=True)
        return 0
    if args.precision_only:
        print(json.dumps(precision_error_measure()), flush=True)
        return 0
    if args.batch_512:
        print(json.dumps(batch512_measure()), flush=True)
        return 0
    if args.per_config_only:
        print(json.dumps(per_config_measure()), flush=True)
        return 0

    t_setup = time.perf_counter()
    # one continuous stream: the warm-up steps' jobs, then the timed steps'
    n_warm, n_timed = args.warmup * JOBS_PER_STEP, args.steps * JOBS_PER_STEP
    stream = make_stream(n_warm + n_timed)
    warm, timed = stream[:n_warm], stream[n_warm:]
    pool = BoxPool()
    for j in stream:
        pool.get(j["kind"], j["n"])
    setup_s = time.perf_counter() - t_setup
    print(f"# generated {len(pool.boxes)} boxes for {len(stream)} jobs in "
          f"{setup_s:.1f} s", file=sys.stderr, flush=True)

    os.environ.setdefault("MCGPU_SERVER_LOG", "10")
    # warm-up: same server path, saturated (all warm-up jobs at t=0)
    wticks = 0
    if warm:
        wj = [dict(j, arrival=0.0, steps=min(j["steps"], 2000)) for j in warm]
        _, wst = serve(wj, pool, t_shift=False, sweeps=args.sweeps, tick_us=args.tick_us)
        wticks = wst.ticks
    m.check(m.lib().mcgpu_device_synchronize(), "mcgpu_device_synchronize")
    print("# warm-up done", file=sys.stderr, flush=True)

    # timed region: exactly K steps of the stream (K x JOBS_PER_STEP jobs),
    # on the server's clock from its start to the last job's completion
    # (every job's result is read back, so the device is idle at both ends)
    res, st = serve(timed, pool, sweeps=args.sweeps, tick_us=args.tick_us)
    bad = [r for r in res if r.status != 0]
    if bad:
        print(f"bench.py: {len(bad)} jobs failed", file=sys.stderr)
        return 1
    moves = int(st.total_moves)
    value = moves / st.wall_s
    lat = np.array([r.finish_s - r.arrival_s for r in res])
    print(f"# serving: {len(res)} jobs, {moves} moves in {st.wall_s:.2f} s "
          f"({value:.4g} moves/s), {st.ticks} ticks, peak resident "
          f"{st.peak_resident}, p50 {np.median(lat):.3f} s p99 "
          f"{np.quantile(lat, 0.99):.3f} s", file=sys.stderr, flush=True)

    # saturated capacity: the same jobs, all arriving at t = 0
    cap_jobs = [dict(j, arrival=0.0) for j in timed[: min(len(timed), 600)]]
    _, cst = serve(cap_jobs, pool, t_shift=False, sweeps=args.sweeps, tick_us=args.tick_us)
    capacity = cst.total_moves / cst.wall_s
    print(f"# capacity run: {capacity:.4g} moves/s", file=sys.stderr, flush=True)

    # roofline of the batch kernel (timed serving run)
    kernel_s = st.kernel_ms_total / 1e3
    alg_bytes = 24.0 * (st.work_candidates + st.work_sites)
    achieved = alg_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    per_launch = alg_bytes / max(st.ticks, 1)
    ratio = pmc_traffic_ratio(args.steps, args.warmup)
    traffic = ratio * per_launch if ratio is not None else None
    # every serving loop of the process (warm-up, timed, capacity): what a
    # counter profile of this command covers
    alg_bytes_run = 24.0 * sum(float(x.work_candidates + x.work_sites)
                               for x in ([wst] if warm else []) + [st, cst])
    # the same kernel saturated: the capacity run's boxes over its kernel time
    cap_kernel_s = cst.kernel_ms_total / 1e3
    cap_achieved = (24.0 * (cst.work_candidates + cst.work_sites) / cap_kernel_s / 1e9
                    if cap_kernel_s > 0 else 0.0)

    plain_rate, plain_jobs, plain_moves = plain_sample(timed, pool, args.plain_seconds)
    print(f"# plain FP64 path: {plain_rate:.4g} moves/s over {plain_jobs} jobs",
          file=sys.stderr, flush=True)

    full = full_energy_measure()
    print(f"# full energy: {json.dumps(full)}", file=sys.stderr, flush=True)
    prec = precision_error_measure()
    print(f"# FP32-pair dE error: {json.dumps(prec)}", file=sys.stderr, flush=True)
    per_config = per_config_measure()
    print(f"# per configuration: {json.dumps(per_config)}", file=sys.stderr, flush=True)

    b512 = None
    if not args.no_batch512:
        b512 = batch512_measure()
        print(f"# 512 water boxes: {json.dumps(b512)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(timed, pool, args.cpu_seconds)
        cpu["cpu_model"] = cpu_model()
        cpu["cores_available"] = os.cpu_count()
        print(f"# CPU baseline: {json.dumps(cpu)}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * st.wall_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 coordinates/sums + f32-or-f64 pair terms",
        "data": "synthetic",
        "config": {
            "workload": ("serving stream: Poisson 200 jobs/s, water (c2 generator) "
                         "and 16-site bicyclic (c1 generator) boxes, 128-16384 "
                         "molecules, 1e3-1e5 steps, NVT/NPT, FP64/FP32-pair mixed"),
            "jobs": len(timed), "jobs_per_step": JOBS_PER_STEP,
            "first_job": n_warm, "moves": moves, "max_sweeps_per_tick": args.sweeps,
            "tick_budget_us": args.tick_us,
            "neighbor": ("cell lists at >= 4 cells of width rc per edge and >= 1000 "
                         "molecules, else all pairs (the plain path alike)"),
            "parallelism": "boxes in one persistent launch",
        },
        "p50_latency_s": float(np.median(lat)),
        "p99_latency_s": float(np.quantile(lat, 0.99)),
        "plain_fp64_moves_per_s": plain_rate,
        "plain_sample": f"{plain_jobs} jobs of the stream, {plain_moves} moves",
        "ratio_vs_plain": value / plain_rate,
        "capacity_moves_per_s": capacity,
        "capacity_ratio_vs_plain": capacity / plain_rate,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_per_algorithmic_byte": ratio,
            "kernel": ("k_batch (a tick: one graph launch of the 256-, 512- and 1024-lane "
                       "instances, and beside it, on the volume stream, the NPT volume "
                       "moves and then their boxes' late instances)"),
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_ms": st.kernel_ms_total / max(st.ticks, 1),
            # graph launches and algorithmic bytes of the whole process
            # (warm-up, timed and capacity serving), to normalise
            # per-process counter totals
            "graph_launches_in_run": int(wticks + st.ticks + cst.ticks),
            "algorithmic_bytes_in_run": alg_bytes_run,
            # the capacity run (every job at once): the kernel with the
            # device full rather than paced by arrivals
            "capacity_achieved": cap_achieved,
            "capacity_frac": cap_achieved / HBM_PEAK_GBS,
        },
        "full_energy": full,
        "fp32_pair_de_error_kcal_mol": prec,
        "per_config": per_config,
        "batch512": b512,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
