"""Verlet lists on the GPU, with the reference's shell (1.2875 rc,
VerletStep.cpp:133): rebuilt whenever a committed primary has moved more
than skin/2 since the last build (VerletStep.cpp:186-207) and before a
proposal that would leave the skin, built from cells of width >= rc + skin
where the box holds three per edge, and giving the oracle's trace.  An
evaluation scans every molecule only for a proposal that moves a primary
more than skin/2 in one move (possible for large molecules rotated about a
distant site): the lists never go stale."""
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from gpu_checks import de_ok
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_verlet_20000_moves_rebuild_and_match_oracle():
    hb = m.HostBox.from_state(str(GOLDEN / "water256.state"))
    steps = 20000
    ob = O.OracleBox(hb.arrays())
    fin, e0, nacc, want = ob.run_nvt(hb.mc_seed, steps, use_libc=False)
    res, tr = m.plain_run(hb, steps, precision=m.FP64, neighbor=m.NEIGHBOR_VERLET)
    print(f"Verlet builds {res.neighbor_builds}, fallbacks {res.neighbor_fallbacks}")
    assert res.neighbor_builds > 1
    assert res.neighbor_fallbacks == 0
    assert np.array_equal(tr.mol, want["mol"])
    assert np.array_equal(tr.accepted, want["accepted"])
    assert de_ok(tr.de, want)
    assert res.final_energy == pytest.approx(fin, rel=1e-11, abs=1e-7)


def test_verlet_lists_built_from_cells():
    """1,500 16-site molecules (4 cells of >= rc + skin per edge): the lists
    come from the cell build, and every molecule energy equals the oracle's;
    a run of moves keeps the oracle's trace with rebuilds and no fallback."""
    hb = m.workload("bicyclic4000", n_molecules=1500)
    ob = O.OracleBox(hb.arrays())
    db = m.DeviceBox(hb)
    db.neighbor(m.NEIGHBOR_VERLET, skin=1.0)
    for i in (0, 1, 777, 1499):
        assert db.molecule_energy(i) == pytest.approx(ob.molecule_energy(i), rel=1e-12,
                                                      abs=1e-9)
    builds, fallbacks = db.neighbor_stats()
    assert builds == 1 and fallbacks == 0
    db.close()
    steps = 3000
    _, _, _, want = ob.run_nvt(hb.mc_seed, steps, use_libc=False)
    res, tr = m.plain_run(hb, steps, precision=m.FP64, neighbor=m.NEIGHBOR_VERLET)
    print(f"Verlet builds {res.neighbor_builds}, fallbacks {res.neighbor_fallbacks}")
    assert res.neighbor_builds > 1 and res.neighbor_fallbacks <= steps // 100
    assert np.array_equal(tr.mol, want["mol"])
    assert np.array_equal(tr.accepted, want["accepted"])
    assert de_ok(tr.de, want)


def test_verlet_strategy_calls_rebuild_after_set_molecule():
    """changeMolecule / rollback through mcgpu_set_molecule_coords: a jump
    far past skin/2 makes the next evaluation rebuild the lists first."""
    hb = m.HostBox.from_state(str(GOLDEN / "water256.state"))
    ob = O.OracleBox(hb.arrays())
    db = m.DeviceBox(hb)
    db.neighbor(m.NEIGHBOR_VERLET, skin=1.0)
    xyz = db.get_molecule(5)
    moved = xyz.copy()
    moved[:3] += 4.0  # 4 A along x
    db.set_molecule(5, moved)
    ob.a["x"][ob.a["mol_start"][5]:ob.a["mol_start"][5] + 3] += 4.0
    for i in (5, 6, 100):
        assert db.molecule_energy(i) == pytest.approx(ob.molecule_energy(i), rel=1e-12,
                                                      abs=1e-9)
    builds, fallbacks = db.neighbor_stats()
    assert builds == 2 and fallbacks == 0
