"""Resume on the GPU: a box stopped after 3000 moves, saved to a .state file
with its resume record, read back and continued for 500 moves, makes the
same moves, decisions and energies as one run of 3500 moves -- NVT and NPT,
the NPT run split both on a volume stop and between two."""
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from gpu_checks import de_ok, run_all

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
P_ATM = 1.4584e-5


def run(hb, steps, kw, resume=None):
    b = m.Batch(2)
    s = b.add(hb, steps=steps, record_trace=True, resume=resume, **kw)
    (st,) = run_all(b, [s], steps, 700)
    out = dict(st=st, trace=b.trace(s), resume=b.get_resume(s), coords=b.download(s))
    if kw.get("ensemble") == m.NPT:
        out["vol"] = b.volume_trace(s)
    b.close()
    return out


def golden(name):
    return lambda: m.HostBox.from_state(str(GOLDEN / f"{name}.state"))


def split_run(make, kw, first, total, tmp_path, neighbor=m.NEIGHBOR_BRUTE):
    kw = dict(kw, neighbor=neighbor)
    hb = make()
    straight = run(hb, total, kw)
    a = run(hb, first, kw)
    # through a .state file: coordinates, edges and the resume record
    hb2 = make()
    hb2.set_coords(*a["coords"], edges=a["st"].box[:])
    hb2.resume = a["resume"]
    p = tmp_path / "cut.state"
    hb2.write_state(str(p), step=first)
    hb3 = m.HostBox.from_state(str(p))
    assert hb3.resume is not None and hb3.resume.steps_done == first
    b = run(hb3, total - first, kw, resume=hb3.resume)
    tm = np.concatenate([a["trace"].mol, b["trace"].mol])
    ta = np.concatenate([a["trace"].accepted, b["trace"].accepted])
    td = np.concatenate([a["trace"].de, b["trace"].de])
    assert np.array_equal(tm, straight["trace"].mol)
    assert np.array_equal(ta, straight["trace"].accepted)
    return straight, a, b, td


def test_resume_nvt_continues_bit_exactly(tmp_path):
    straight, a, b, td = split_run(golden("water256"), dict(seed=1), 3000, 3500, tmp_path)
    # same box contents, same kernel, same summation order: bit-exact
    assert np.array_equal(td, straight["trace"].de)
    assert b["st"].energy == straight["st"].energy
    assert a["st"].accepted + b["st"].accepted == straight["st"].accepted
    assert b["resume"].steps_done == 3500


@pytest.mark.parametrize("first", [3000, 2750], ids=["on-volume-stop", "between-stops"])
def test_resume_npt_keeps_the_volume_schedule(first, tmp_path):
    kw = dict(seed=1, ensemble=m.NPT, volume_interval=500, pressure=P_ATM,
              max_volume_delta=0.002)
    straight, a, b, td = split_run(golden("mixture300"), kw, first, 3500, tmp_path)
    assert a["resume"].volume_pending == (1 if first % 500 == 0 else 0)
    assert np.array_equal(td, straight["trace"].de)
    assert b["st"].energy == straight["st"].energy
    assert a["st"].volume_moves + b["st"].volume_moves == straight["st"].volume_moves == 6
    v = {k: np.concatenate([a["vol"][k], b["vol"][k]]) for k in ("accepted", "u", "e_new")}
    assert np.array_equal(v["accepted"], straight["vol"]["accepted"])
    assert np.array_equal(v["u"], straight["vol"]["u"])
    assert np.array_equal(v["e_new"], straight["vol"]["e_new"])
    assert tuple(b["st"].box) == tuple(straight["st"].box)


def test_resume_with_cells(tmp_path):
    """Cell lists are rebuilt in molecule order on admission, where the
    uninterrupted run's lists were updated move by move: the order of the
    pair sums may differ, so dE agrees to round-off, decisions exactly."""
    make = lambda: m.workload("bicyclic4000", n_molecules=1500)  # 5 cells per edge
    straight, a, b, td = split_run(make, dict(seed=3), 1500, 2000, tmp_path,
                                   neighbor=m.NEIGHBOR_CELL)
    assert de_ok(td, straight["trace"].de)
    assert b["st"].energy == pytest.approx(straight["st"].energy, rel=1e-11, abs=1e-8)
