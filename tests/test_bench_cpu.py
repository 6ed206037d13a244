"""CPU checks of the benchmark's workload definitions: the job stream is
seeded and prefix-stable (warm-up and timed jobs are slices of one stream),
its parameters stay in the configured ranges, the neighbour policy picks the
scan per box class, and a request for more than one GPU is refused."""
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402
import mcgpu_amd as m  # noqa: E402


def test_stream_is_seeded_and_prefix_stable():
    long, short = bench.make_stream(300), bench.make_stream(120)
    assert long[:120] == short
    assert bench.make_stream(120) == short
    assert bench.make_stream(120, seed=7) != short


def test_stream_parameters_stay_in_their_ranges():
    jobs = bench.make_stream(2500)
    t = np.array([j["arrival"] for j in jobs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0.0)
    # Poisson arrivals at the configured rate: mean gap within 5% of 1/rate
    assert abs(np.diff(t).mean() * bench.LAMBDA - 1.0) < 0.05
    sizes = {j["n"] for j in jobs}
    assert min(sizes) == 128 and max(sizes) == 16384 and len(sizes) == bench.N_CLASSES
    steps = np.array([j["steps"] for j in jobs])
    assert steps.min() >= 1000 and steps.max() <= 100000
    for key in ("npt", "mixed"):
        share = np.mean([j[key] for j in jobs])
        assert 0.45 < share < 0.55, key
    kinds = {j["kind"] for j in jobs}
    assert kinds == {m.WATER, m.BICYCLIC}
    assert len({j["mc_seed"] for j in jobs}) == len(jobs)


@pytest.mark.parametrize("kind,n,cells", [
    (m.WATER, 128, False),      # under 3 cells per edge
    (m.WATER, 692, False),      # 3 cells: the 27 cells are the whole box
    (m.WATER, 1303, False),     # 3 cells
    (m.WATER, 1609, True),      # 4 cells, 1,609 molecules
    (m.WATER, 16384, True),     # 8 cells
    (m.BICYCLIC, 560, False),   # 3 cells
    (m.BICYCLIC, 692, False),   # 4 cells but under 1,000 molecules
    (m.BICYCLIC, 1055, True),   # 4 cells
])
def test_neighbor_policy_per_box_class(kind, n, cells):
    hb = bench.BoxPool().get(kind, n)
    d = hb.desc()
    per_edge = int(min(d.box) // d.cutoff)
    want = m.NEIGHBOR_CELL if cells else m.NEIGHBOR_BRUTE
    assert bench.neighbor_for(hb) == want, (n, per_edge)
    assert cells == (per_edge >= 4 and d.n_molecules >= 1000)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stdout + r.stderr


def test_log_uniform_bounds():
    rng = np.random.default_rng(0)
    x = bench.log_uniform(rng, 128, 16384, 10000)
    assert x.min() >= 128 and x.max() <= 16384
    # log-uniform: the median sits at the geometric mean
    assert abs(math.log(np.median(x)) - 0.5 * math.log(128 * 16384)) < 0.1
