"""GPU parity at the edges of the input space the reference accepts:
molecules of the maximum site count (32, MCGPU_MAX_MOL_SITES), several
primary sites per molecule at a non-zero offset, dummy sites, and ragged
boxes that mix 1-, 2-, 5- and 32-site molecules.  Boxes are built here from
plain arrays (the loader's descriptor layout), checked against the oracle's
energies and against its trial-move traces for the same seed."""
import ctypes as C

import numpy as np
import pytest

import mcgpu_amd as m
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DE_TOL = 1e-9


class ArrayBox:
    """A box described by numpy arrays; duck-types the HostBox methods that
    DeviceBox and Batch.add use (desc, mc_seed)."""

    def __init__(self, a: dict, mc_seed: int):
        self.a = a
        self.mc_seed = mc_seed

    def arrays(self):
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.a.items()}

    def desc(self) -> m.BoxDesc:
        a = self.a
        d = m.BoxDesc()
        d.n_molecules = len(a["mol_start"])
        d.n_sites = len(a["x"])
        d.n_primary = len(a["primary_index"])
        d.box[:] = [float(v) for v in a["box"]]
        d.cutoff, d.kT = a["cutoff"], a["kT"]
        d.max_translate, d.max_rotate = a["max_translate"], a["max_rotate"]
        dp = lambda k: a[k].ctypes.data_as(C.POINTER(C.c_double))
        ip = lambda k: a[k].ctypes.data_as(C.POINTER(C.c_int32))
        d.x, d.y, d.z = dp("x"), dp("y"), dp("z")
        d.sigma, d.epsilon, d.charge = dp("sigma"), dp("epsilon"), dp("charge")
        d.mol_start, d.mol_len = ip("mol_start"), ip("mol_len")
        d.mol_pidx_start, d.mol_pidx_count = ip("mol_pidx_start"), ip("mol_pidx_count")
        d.primary_index = ip("primary_index")
        return d


def _kind(n_sites, rng, primaries, dummies=()):
    """A rigid molecule: compact site geometry (1.1 A grid, first site at the
    origin), OPLS-like parameters, a neutral charge pattern."""
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(2), indexing="ij"),
                 -1).reshape(-1, 3)[:n_sites] * 1.1
    sigma = rng.uniform(2.6, 3.6, n_sites)
    eps = rng.uniform(0.03, 0.2, n_sites)
    q = np.where(np.arange(n_sites) % 2 == 0, 0.3, -0.3)
    if n_sites % 2:
        q[-1] = 0.0
    for k in dummies:
        sigma[k] = -1.0
    return dict(g=g, sigma=sigma, eps=eps, q=q, prim=list(primaries))


def _rotation(rng):
    qq = rng.normal(size=4)
    w, x, y, z = qq / np.linalg.norm(qq)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def build_box(kinds, n_mol, edge, cutoff, seed):
    """Molecules cycle through `kinds`, one per jittered lattice point."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n_mol ** (1 / 3)))
    h = edge / k
    pts = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[rng.permutation(len(pts))[:n_mol]] * h + 0.5 * h
    xs, sg, ep, qs, ms, ml, ps, pc, pi = [], [], [], [], [], [], [], [], []
    s0 = 0
    for i in range(n_mol):
        t = kinds[i % len(kinds)]
        n = len(t["sigma"])
        xyz = pts[i] + rng.uniform(-0.3, 0.3, 3) + (t["g"] - t["g"].mean(0)) @ _rotation(rng).T
        xs.append(xyz)
        sg.append(t["sigma"])
        ep.append(t["eps"])
        qs.append(t["q"])
        ms.append(s0)
        ml.append(n)
        ps.append(len(pi))
        pc.append(len(t["prim"]))
        pi.extend(s0 + p for p in t["prim"])
        s0 += n
    xyz = np.concatenate(xs)
    a = dict(x=np.ascontiguousarray(xyz[:, 0]), y=np.ascontiguousarray(xyz[:, 1]),
             z=np.ascontiguousarray(xyz[:, 2]),
             sigma=np.concatenate(sg), epsilon=np.concatenate(ep), charge=np.concatenate(qs),
             mol_start=np.array(ms, np.int32), mol_len=np.array(ml, np.int32),
             mol_pidx_start=np.array(ps, np.int32), mol_pidx_count=np.array(pc, np.int32),
             primary_index=np.array(pi, np.int32),
             box=np.array([edge] * 3), cutoff=cutoff, kT=0.5922, max_translate=0.4,
             max_rotate=12.0)
    return ArrayBox(a, mc_seed=seed + 1)


def max_site_box(n_mol, edge, cutoff, seed):
    rng = np.random.default_rng(100 + seed)
    return build_box([_kind(32, rng, (3, 28), dummies=(5, 20))], n_mol, edge, cutoff, seed)


def ragged_box(n_mol, edge, cutoff, seed):
    rng = np.random.default_rng(200 + seed)
    kinds = [_kind(1, rng, (0,)), _kind(32, rng, (0, 17, 31)), _kind(2, rng, (1,)),
             _kind(5, rng, (2,), dummies=(4,))]
    return build_box(kinds, n_mol, edge, cutoff, seed)


# (builder, molecules, edge, cutoff): all-pairs boxes (fewer than 3 cells per
# edge) and boxes of exactly 3 cells per edge (cell width rc + twice the
# primary spread: 17.3 A for the 32-site kind, 19.6 A in the ragged box)
BOXES = {
    "max32_brute": (max_site_box, 48, 30.0, 11.0),
    "max32_cells": (max_site_box, 216, 56.0, 10.0),
    "ragged_brute": (ragged_box, 64, 26.0, 10.0),
    "ragged_cells": (ragged_box, 343, 62.0, 10.0),
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


@pytest.mark.parametrize("name", ["max32_cells", "ragged_cells"])
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
