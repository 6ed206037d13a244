"""Boxes built from plain arrays in the loader's descriptor layout: rigid
molecules of up to 32 sites (MCGPU_MAX_MOL_SITES) on a jittered lattice, for
the edge-input parity tests and the descriptor-validation tests."""
import ctypes as C

import numpy as np

import mcgpu_amd as m


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


def kind(n_sites, rng, primaries, dummies=()):
    """A rigid molecule: compact site geometry (1.1 A grid, first site at the
    origin), OPLS-like parameters, a neutral charge pattern."""
    nz = 2 if n_sites <= 32 else 3
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(nz), indexing="ij"),
                 -1).reshape(-1, 3)[:n_sites] * 1.1
    sigma = rng.uniform(2.6, 3.6, n_sites)
    eps = rng.uniform(0.03, 0.2, n_sites)
    q = np.where(np.arange(n_sites) % 2 == 0, 0.3, -0.3)
    if n_sites % 2:
        q[-1] = 0.0
    for k in dummies:
        sigma[k] = -1.0
    return dict(g=g, sigma=sigma, eps=eps, q=q, prim=list(primaries))


def rotation(rng):
    qq = rng.normal(size=4)
    w, x, y, z = qq / np.linalg.norm(qq)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def build_box(kinds, n_mol, edge, cutoff, seed):
    """Molecules cycle through `kinds`, one per jittered lattice point.
    `edge`: one box edge, or three (a non-cubic box)."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n_mol ** (1 / 3)))
    edges = np.broadcast_to(np.asarray(edge, dtype=np.float64), (3,)).copy()
    h = edges / k
    pts = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[rng.permutation(len(pts))[:n_mol]] * h + 0.5 * h
    xs, sg, ep, qs, ms, ml, ps, pc, pi = [], [], [], [], [], [], [], [], []
    s0 = 0
    for i in range(n_mol):
        t = kinds[i % len(kinds)]
        n = len(t["sigma"])
        xyz = pts[i] + rng.uniform(-0.3, 0.3, 3) + (t["g"] - t["g"].mean(0)) @ rotation(rng).T
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
             box=edges, cutoff=cutoff, kT=0.5922, max_translate=0.4,
             max_rotate=12.0)
    return ArrayBox(a, mc_seed=seed + 1)


def max_site_box(n_mol, edge, cutoff, seed):
    rng = np.random.default_rng(100 + seed)
    return build_box([kind(32, rng, (3, 28), dummies=(5, 20))], n_mol, edge, cutoff, seed)


def ragged_box(n_mol, edge, cutoff, seed):
    rng = np.random.default_rng(200 + seed)
    kinds = [kind(1, rng, (0,)), kind(32, rng, (0, 17, 31)), kind(2, rng, (1,)),
             kind(5, rng, (2,), dummies=(4,))]
    return build_box(kinds, n_mol, edge, cutoff, seed)


def small_box(n_mol, edge, cutoff, seed):
    """Molecules of 1 to 5 sites (one kind with two primaries): the
    full-energy kernel takes its half-shell path for them."""
    rng = np.random.default_rng(300 + seed)
    kinds = [kind(1, rng, (0,)), kind(4, rng, (0, 3)), kind(5, rng, (2,), dummies=(4,)),
             kind(2, rng, (1,)), kind(3, rng, (0,))]
    return build_box(kinds, n_mol, edge, cutoff, seed)
