"""GPU parity at the edges of the input space the reference accepts:
molecules of the maximum site count (32, MCGPU_MAX_MOL_SITES), several
primary sites per molecule at a non-zero offset, dummy sites, ragged
boxes that mix 1-, 2-, 5- and 32-site molecules, and non-cubic boxes.  Boxes are built here from
plain arrays (the loader's descriptor layout), checked against the oracle's
energies and against its trial-move traces for the same seed."""
import numpy as np
import pytest

import mcgpu_amd as m
from edge_boxes import max_site_box, ragged_box, rotation, small_box
from gpu_checks import npt_against_oracle
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DE_TOL = 1e-9


# (builder, molecules, edge, cutoff): all-pairs boxes (fewer than 3 cells per
# edge) and boxes of exactly 3 cells per edge (cell width rc + twice the
# primary spread: 17.3 A for the 32-site kind, 19.6 A in the ragged box)
BOXES = {
    "max32_brute": (max_site_box, 48, 30.0, 11.0),
    "max32_cells": (max_site_box, 216, 56.0, 10.0),
    "ragged_brute": (ragged_box, 64, 26.0, 10.0),
    "ragged_cells": (ragged_box, 343, 62.0, 10.0),
    # non-cubic: 3 x 3 x 4 and 3 x 4 x 3 cells
    "noncubic32_cells": (max_site_box, 256, (56.0, 60.0, 75.0), 10.0),
    "noncubicragged_cells": (ragged_box, 400, (62.0, 80.0, 66.0), 10.0),
    # molecules of at most 5 sites (the half-shell full energy), 3 x 4 x 5 cells
    "noncubicsmall_cells": (small_box, 600, (40.0, 52.0, 64.0), 9.0),
}


def de_ok(got, want):
    scale = np.maximum(np.abs(want["old_e"]), np.abs(want["new_e"]))
    err = np.abs(got - want["de"])
    print(f"max |dE err| {err.max():.3e}")
    return np.all(err <= DE_TOL + 1e-12 * scale)


@pytest.mark.parametrize("name", sorted(BOXES))
def test_edge_boxes_energies_match_oracle(name):
    fn, n, edge, rc = BOXES[name]
    ab = fn(n, edge, rc, seed=1)
    ob = O.OracleBox(ab.arrays())
    e_ref = ob.system_energy()
    print(f"{name}: {n} molecules, E = {e_ref:.6f}")
    brute = m.DeviceBox(ab.desc())
    cell = m.DeviceBox(ab.desc())
    cell.neighbor(m.NEIGHBOR_CELL)
    for i in range(n):
        want = ob.molecule_energy(i)
        assert brute.molecule_energy(i) == pytest.approx(want, rel=1e-12, abs=DE_TOL), i
        assert cell.molecule_energy(i) == pytest.approx(want, rel=1e-12, abs=DE_TOL), i
    e_b, e_c = brute.system_energy(), cell.system_energy()
    assert e_b == pytest.approx(e_ref, rel=1e-12, abs=DE_TOL)
    assert e_c == pytest.approx(e_ref, rel=1e-12, abs=DE_TOL)
    # each real site pair of every included molecule pair, once
    assert brute.system_pairs() == cell.system_pairs() > 0


@pytest.mark.parametrize("name", sorted(BOXES))
def test_edge_boxes_batch_trace_matches_oracle(name):
    fn, n, edge, rc = BOXES[name]
    ab = fn(n, edge, rc, seed=2)
    steps = 1200
    ob = O.OracleBox(ab.arrays())
    _, e0, _, want = ob.run_nvt(ab.mc_seed, steps, use_libc=False)
    assert want["accepted"].sum() > 0
    b = m.Batch(2)
    nb = m.NEIGHBOR_CELL if name.endswith("cells") else m.NEIGHBOR_BRUTE
    s = b.add(ab, steps=steps, neighbor=nb, record_trace=True)
    b.run(steps)
    st = b.stats(s)
    assert st.done and st.steps_done == steps
    tr = b.trace(s)
    assert st.start_energy == pytest.approx(e0, rel=1e-12, abs=DE_TOL)
    assert np.array_equal(tr.mol, want["mol"])
    assert np.array_equal(tr.accepted, want["accepted"])
    assert de_ok(tr.de, want)


@pytest.mark.parametrize("name", ["max32_cells", "ragged_cells", "noncubicragged_cells"])
def test_edge_boxes_mixed_precision_error(name):
    fn, n, edge, rc = BOXES[name]
    ab = fn(n, edge, rc, seed=3)
    steps = 1200
    ob = O.OracleBox(ab.arrays())
    _, _, _, want = ob.run_nvt(ab.mc_seed, steps, use_libc=False)
    b = m.Batch(2)
    s = b.add(ab, steps=steps, neighbor=m.NEIGHBOR_CELL, precision=m.MIXED, record_trace=True)
    b.run(steps)
    tr = b.trace(s)
    same = np.cumprod(tr.accepted == want["accepted"]).astype(bool)
    k = int(same.sum())
    ok = np.abs(want["de"][:k]) < 1e3   # moves into overlaps: |dE| up to 1e12
    err = np.abs(tr.de[:k] - want["de"][:k])[ok]
    print(f"{name} mixed: {k} moves aligned, max |dE err| {err.max():.3e}")
    assert k >= 300
    assert err.max() <= 1e-4


@pytest.mark.parametrize("name,nb", [(k, k.split("_")[1]) for k in sorted(BOXES)] +
                         [(k, "verlet") for k in sorted(BOXES) if k.endswith("cells")])
def test_edge_boxes_single_box_trial_moves(name, nb):
    """The single-box fused trial move (mcgpu_trial_move) on the edge boxes:
    old and new molecule energies against the oracle's, the Metropolis
    decision of SimulationStep.cpp:794-803 from those energies and the
    caller's draw, and the commit on acceptance (the oracle's copy follows
    the accepted moves, so later moves see the committed coordinates).  The
    boxes of 3 cells per edge run with cell lists, and again with Verlet
    lists (1 A skin, so lists go stale and are rebuilt along the way)."""
    fn, n, edge, rc = BOXES[name]
    ab = fn(n, edge, rc, seed=4)
    ob = O.OracleBox(ab.arrays())
    a = ob.a
    db = m.DeviceBox(ab.desc())
    if nb == "verlet":
        db.neighbor(m.NEIGHBOR_VERLET, 1.0)
    elif nb == "cells":
        db.neighbor(m.NEIGHBOR_CELL)
    rng = np.random.default_rng(9)
    n_acc = 0
    for _ in range(150):
        i = int(rng.integers(n))
        s0, ln = int(a["mol_start"][i]), int(a["mol_len"][i])
        old = np.stack([a[k][s0:s0 + ln].copy() for k in "xyz"])
        pivot = old[:, :1]
        new = pivot + rotation(rng) @ (old - pivot) + rng.uniform(-0.5, 0.5, (3, 1))
        e_old = ob.molecule_energy(i)
        for k, c in enumerate("xyz"):
            a[c][s0:s0 + ln] = new[k]
        e_new = ob.molecule_energy(i)
        u = float(rng.random())
        de = e_new - e_old
        accept = de < 0 or np.exp(-de / a["kT"]) >= u
        res = db.trial_move(i, new.reshape(-1), u)
        assert res.old_energy == pytest.approx(e_old, rel=1e-12, abs=DE_TOL)
        assert res.new_energy == pytest.approx(e_new, rel=1e-12, abs=DE_TOL)
        assert bool(res.accepted) == accept
        if not accept:
            for k, c in enumerate("xyz"):
                a[c][s0:s0 + ln] = old[k]
        n_acc += accept
    print(f"{name}: {n_acc} of 150 accepted")
    assert 0 < n_acc < 150
    assert db.system_energy() == pytest.approx(ob.system_energy(), rel=1e-12, abs=DE_TOL)


def test_first_primaries_outside_the_box_with_cells():
    """Coordinates need not be wrapped into [0, L): molecules whose first
    primary lies a few A past a face, in a box of 5 cells per edge.
    Cell-list energies must equal the oracle's (all pairs, single wrap), and
    the batched trace with cells must equal the oracle's trace."""
    ab = max_site_box(512, 90.0, 10.0, seed=5)
    a = ab.a
    L = 90.0
    # the whole box shifted by half a lattice spacing: the top x layer and the
    # bottom y layer straddle the faces (pair displacements are unchanged)
    a["x"] += 5.625
    a["y"] -= 5.625
    p = a["primary_index"][a["mol_pidx_start"]]
    outside = (a["x"][p] >= L) | (a["x"][p] < 0) | (a["y"][p] >= L) | (a["y"][p] < 0)
    assert outside.sum() >= 20
    ob = O.OracleBox(ab.arrays())
    db = m.DeviceBox(ab.desc())
    db.neighbor(m.NEIGHBOR_CELL)
    brute = m.DeviceBox(ab.desc())
    for i in np.flatnonzero(outside)[:60]:
        want = ob.molecule_energy(int(i))
        assert db.molecule_energy(int(i)) == pytest.approx(want, rel=1e-12, abs=DE_TOL), i
        assert brute.molecule_energy(int(i)) == pytest.approx(want, rel=1e-12, abs=DE_TOL), i
    e_ref = ob.system_energy()
    assert db.system_energy() == pytest.approx(e_ref, rel=1e-12, abs=DE_TOL)
    assert brute.system_energy() == pytest.approx(e_ref, rel=1e-12, abs=DE_TOL)
    assert db.system_pairs() == brute.system_pairs() > 0
    steps = 1500
    _, e0, _, want = ob.run_nvt(ab.mc_seed, steps, use_libc=False)
    b = m.Batch(2)
    s = b.add(ab, steps=steps, neighbor=m.NEIGHBOR_CELL, record_trace=True)
    b.run(steps)
    tr = b.trace(s)
    assert b.stats(s).start_energy == pytest.approx(e0, rel=1e-12, abs=DE_TOL)
    assert np.array_equal(tr.mol, want["mol"])
    assert np.array_equal(tr.accepted, want["accepted"])
    assert de_ok(tr.de, want)


@pytest.mark.parametrize("name", ["max32_cells", "ragged_cells", "noncubic32_cells",
                                  "noncubicsmall_cells"])
def test_edge_boxes_npt_matches_oracle(name):
    """Batched NPT on the edge boxes with cell lists: the device-decided
    volume moves rescale by first primary, rebuild the cells and take a
    full energy over molecules of 1 to 32 sites; moves and volume moves
    equal the oracle's NPT loop, and the tracked energy a fresh oracle
    evaluation of the final box."""
    fn, n, edge, rc = BOXES[name]
    ab = fn(n, edge, rc, seed=6)
    vg = npt_against_oracle(ab, 1000, 200, 0.002, m.NEIGHBOR_CELL)
    print(f"{name}: volume moves accepted {vg['accepted'].sum()} of {len(vg['accepted'])}")


@pytest.mark.parametrize("nb", ["cells", "verlet"])
def test_set_molecule_jumps_across_cells(nb):
    """mcgpu_set_molecule_coords (the changeMolecule / rollback path) moving
    molecules far across the box, into other cells and past the faces: the
    neighbour structures must follow, so every molecule energy and the system
    energy still equal the oracle's."""
    fn, n, edge, rc = BOXES["noncubicragged_cells"]
    ab = fn(n, edge, rc, seed=7)
    ob = O.OracleBox(ab.arrays())
    a = ob.a
    L = np.asarray(a["box"])
    db = m.DeviceBox(ab.desc())
    if nb == "verlet":
        db.neighbor(m.NEIGHBOR_VERLET, 1.0)
    else:
        db.neighbor(m.NEIGHBOR_CELL)
    rng = np.random.default_rng(11)
    for i in rng.choice(n, 24, replace=False):
        s0, ln = int(a["mol_start"][i]), int(a["mol_len"][i])
        shift = rng.uniform(-0.5, 0.5, 3) * L
        shift[0] = 0.3 * L[0]     # along x: the top layers end up past the face
        xyz = np.stack([a[c][s0:s0 + ln] for c in "xyz"]) + shift[:, None]
        for k, c in enumerate("xyz"):
            a[c][s0:s0 + ln] = xyz[k]
        db.set_molecule(int(i), xyz.reshape(-1))
    for i in range(n):
        assert db.molecule_energy(i) == pytest.approx(ob.molecule_energy(i), rel=1e-12,
                                                      abs=DE_TOL), i
    assert db.system_energy() == pytest.approx(ob.system_energy(), rel=1e-12, abs=DE_TOL)
