"""CPU tests of the oracle itself: it must be pinned before it checks the GPU.

Pins: glibc rand() and its known-answer values; closed forms of the pair
terms; the per-move traces of the reference's own CPU path (golden
fixtures written by tests/golden/make_golden.py from oracle/_ref); sum and
symmetry invariants; rigid-move invariants; the textbook tail correction.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

import mcgpu_amd as m

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["water256", "bicyclic200", "mixture300", "water128_mp"]


# ----------------------------------------------------------------- RNG

def test_rand_known_answers():
    # glibc srand(1) / srand(12345) first six outputs
    g = O.GlibcRng(1)
    assert [g.rand() for _ in range(6)] == [1804289383, 846930886, 1681692777,
                                            1714636915, 1957747793, 424238335]
    g = O.GlibcRng(12345)
    assert [g.rand() for _ in range(6)] == [383100999, 858300821, 357768173,
                                            455528251, 133005921, 116285904]


@pytest.mark.parametrize("seed", [0, 1, 7, 12345, 2**31 - 1, 2**31 + 5, 2**32 - 1])
def test_rand_restatement_matches_libc(seed):
    g = O.GlibcRng(seed)
    mine = [g.rand() for _ in range(2000)]
    assert mine == O.libc_rand_sequence(seed, 2000)


# ------------------------------------------------------- pair closed forms

def test_lj_closed_forms():
    lib = O.lib()
    for sig, eps in [(3.15061, 0.1521), (3.5, 0.066), (2.42, 0.03)]:
        # r = sigma -> 0 ; r = 2^(1/6) sigma -> -eps (minimum)
        assert abs(lib.orc_lj_params(sig, eps, sig, eps, sig * sig)) < 1e-12
        rmin = 2 ** (1 / 6) * sig
        assert lib.orc_lj_params(sig, eps, sig, eps, rmin * rmin) == pytest.approx(-eps, rel=1e-12)
    # geometric mixing: sigma_ab = sqrt(sa sb), eps_ab = sqrt(ea eb)
    sa, sb, ea, eb = 3.0, 3.5, 0.1, 0.2
    s, e = math.sqrt(sa * sb), math.sqrt(ea * eb)
    assert lib.orc_lj_params(sa, ea, sb, eb, s * s) == pytest.approx(0.0, abs=1e-12)
    rm = 2 ** (1 / 6) * s
    assert lib.orc_lj_params(sa, ea, sb, eb, rm * rm) == pytest.approx(-e, rel=1e-12)
    # r^2 == 0 contributes nothing
    assert lib.orc_lj_params(sa, ea, sb, eb, 0.0) == 0.0


def test_coulomb_closed_forms():
    lib = O.lib()
    assert lib.orc_coulomb_params(1.0, -1.0, 1.0) == pytest.approx(-332.06, rel=1e-15)
    assert lib.orc_coulomb_params(-0.834, 0.417, 2.0) == pytest.approx(
        -0.834 * 0.417 * 332.06 / 2.0, rel=1e-15)
    assert lib.orc_coulomb_params(1.0, 1.0, 0.0) == 0.0


def test_minimum_image_single_wrap():
    lib = O.lib()
    L = 10.0
    assert lib.orc_make_periodic(6.0, L) == pytest.approx(-4.0)
    assert lib.orc_make_periodic(-6.0, L) == pytest.approx(4.0)
    assert lib.orc_make_periodic(5.0, L) == 5.0       # exactly L/2 stays
    assert lib.orc_make_periodic(-5.0, L) == -5.0
    # one wrap only (SimulationStep.cpp:377-385): 16 -> 6, not -4
    assert lib.orc_make_periodic(16.0, L) == pytest.approx(6.0)


def _two_site_box(x1, x2, L=20.0, cutoff=9.0, q=(1.0, -1.0), sig=3.0, eps=0.2):
    a = dict(x=np.array([x1, x2]), y=np.array([1.0, 1.0]), z=np.array([1.0, 1.0]),
             sigma=np.array([sig, sig]), epsilon=np.array([eps, eps]),
             charge=np.array(q, dtype=float), mol_start=np.array([0, 1]),
             mol_len=np.array([1, 1]), mol_pidx_start=np.array([0, 1]),
             mol_pidx_count=np.array([1, 1]), primary_index=np.array([0, 1]),
             box=np.array([L, L, L]), cutoff=cutoff, kT=0.6, max_translate=0.1,
             max_rotate=10.0)
    return O.OracleBox(a)


@pytest.mark.parametrize("x1,x2,r", [(1.0, 4.0, 3.0), (1.0, 19.0, 2.0),
                                     (0.5, 8.5, 8.0)])
def test_two_molecule_energy_closed_form(x1, x2, r):
    b = _two_site_box(x1, x2)
    lj = 4 * 0.2 * ((3.0 / r) ** 12 - (3.0 / r) ** 6)
    coul = -332.06 / r
    assert b.molecule_energy(0, 0) == pytest.approx(lj + coul, rel=1e-12)
    assert b.system_energy() == pytest.approx(lj + coul, rel=1e-12)


def test_cutoff_is_inclusive_and_on_primaries():
    b = _two_site_box(1.0, 10.0, cutoff=9.0)       # exactly rc apart
    assert b.in_range(0, 1)
    b = _two_site_box(1.0, 10.0 + 1e-9, cutoff=9.0)
    assert not b.in_range(0, 1)
    assert b.molecule_energy(0, 0) == 0.0


def test_dummy_sites_never_interact():
    b = _two_site_box(1.0, 4.0)
    b.a["sigma"][1] = -1.0   # dummy (ZmatrixScanner convention)
    assert b.molecule_energy(0, 0) == 0.0


# ----------------------------------------------------------- invariants

def _load(name):
    return m.HostBox.from_state(str(GOLDEN / f"{name}.state"))


@pytest.mark.parametrize("name", ["water256", "mixture300", "water128_mp"])
def test_sum_invariant_and_symmetry(name):
    hb = _load(name)
    b = O.OracleBox(hb.arrays())
    n = len(b.a["mol_start"])
    full = sum(b.molecule_energy(i, 0) for i in range(n))
    half = b.system_energy()
    assert full == pytest.approx(2 * half, rel=1e-12, abs=1e-9)
    pairs = b.inrange_pairs()
    for i, j in pairs[:50]:
        assert b.pair_sum(i, j) == pytest.approx(b.pair_sum(j, i), rel=1e-12, abs=1e-12)
        assert b.in_range(j, i)


def test_whole_box_translation_invariance():
    hb = _load("water256")
    a = hb.arrays()
    b0 = O.OracleBox(a)
    e0 = b0.system_energy()
    L = a["box"][0]
    a2 = dict(a)
    # shift every site by the same vector and wrap molecules by primary
    shift = np.array([3.3, -7.1, 11.9])
    x, y, z = a["x"] + shift[0], a["y"] + shift[1], a["z"] + shift[2]
    for s, ms, ml, p in zip(range(len(a["mol_start"])), a["mol_start"], a["mol_len"],
                            a["primary_index"][a["mol_pidx_start"]]):
        for c in (x, y, z):
            c[ms:ms + ml] -= L * np.floor(c[p] / L)
    a2.update(x=x, y=y, z=z)
    assert O.OracleBox(a2).system_energy() == pytest.approx(e0, rel=1e-10)


def test_rigid_move_preserves_geometry_and_rollback_is_exact():
    hb = _load("bicyclic200")
    b = O.OracleBox(hb.arrays())
    before = b.molecule_xyz(5).reshape(3, -1).T.copy()
    rb = b.apply_move(5, 3, [0.1, -0.05, 0.12], [10.0, -7.0, 14.0])
    after = b.molecule_xyz(5).reshape(3, -1).T
    d0 = np.linalg.norm(before[:, None] - before[None], axis=-1)
    d1 = np.linalg.norm(after[:, None] - after[None], axis=-1)
    assert np.allclose(d0, d1, atol=1e-12)
    p = int(b.a["primary_index"][b.a["mol_pidx_start"][5]]) - int(b.a["mol_start"][5])
    assert np.all(after[p] >= 0) and np.all(after[p] <= b.L)
    lib = O.lib()
    lib.orc_rollback(b.ref, 5, rb.ctypes.data_as(O._d))
    assert np.array_equal(b.molecule_xyz(5).reshape(3, -1).T, before)


# ----------------------------------------- the reference's own CPU path

@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_trace(name, golden):
    ref = np.load(golden / f"{name}.ref.npz")
    hb = _load(name)
    b = O.OracleBox(hb.arrays())
    steps = int(ref["steps"])
    fin, e0, nacc, tr = b.run_nvt(1, steps, use_libc=True)
    assert e0 == float(ref["e0"])
    assert np.array_equal(tr["mol"], ref["mol"])
    assert np.array_equal(tr["accepted"], ref["accepted"])
    assert np.array_equal(tr["de"], ref["de"])          # bit for bit
    assert fin == float(ref["final"])
    assert nacc == int(ref["n_accepted"])


def test_oracle_restated_rng_gives_same_trace(golden):
    ref = np.load(golden / "water256.ref.npz")
    b = O.OracleBox(_load("water256").arrays())
    fin, e0, nacc, tr = b.run_nvt(1, 2000, use_libc=False)
    assert np.array_equal(tr["mol"], ref["mol"][:2000])
    assert np.array_equal(tr["de"], ref["de"][:2000])


def test_planted_mistake_is_caught(golden):
    """A transposed displacement convention or a dropped term must not pass
    the trace check: plant them in a copy of the oracle's inputs."""
    ref = np.load(golden / "water256.ref.npz")
    a = _load("water256").arrays()
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    a["charge"] = a["charge"] * 0.999   # a scaled Coulomb term
    b = O.OracleBox(a)
    _, _, _, tr = b.run_nvt(1, 200, use_libc=True)
    assert not np.allclose(tr["de"], ref["de"][:200], atol=1e-9)


# ------------------------------------------------------------------ LRC

def test_lrc_single_species_textbook():
    hb = _load("water256")
    a = hb.arrays()
    b = O.OracleBox(a)
    n = len(a["mol_start"])
    V = float(np.prod(a["box"]))
    rc = a["cutoff"]
    sig, eps = 3.15061, 0.1521   # only O has LJ
    tail = 2 * math.pi * n * n / V * 4 * eps * sig ** 6 * (
        sig ** 6 / (9 * rc ** 9) - 1 / (3 * rc ** 3))
    assert b.lrc_reference() == pytest.approx(tail, rel=1e-7)  # PI = 3.14159265
    assert b.lrc_species() == pytest.approx(tail, rel=1e-12)


def test_volume_scale_closed_forms():
    hb = _load("mixture300")
    b = O.OracleBox(hb.arrays())
    a = b.a
    ms, ml = a["mol_start"], a["mol_len"]
    p = a["primary_index"][a["mol_pidx_start"]]
    P0 = np.stack([a["x"][p], a["y"][p], a["z"][p]], 1).copy()
    inner0 = np.stack([a["x"] - np.repeat(a["x"][p], ml), a["y"] - np.repeat(a["y"][p], ml),
                       a["z"] - np.repeat(a["z"][p], ml)], 1).copy()
    lrc0 = b.lrc_species()
    V0 = float(np.prod(b.L))
    b.volume_scale(1.02)
    P1 = np.stack([a["x"][p], a["y"][p], a["z"][p]], 1)
    inner1 = np.stack([a["x"] - np.repeat(a["x"][p], ml), a["y"] - np.repeat(a["y"][p], ml),
                       a["z"] - np.repeat(a["z"][p], ml)], 1)
    assert np.allclose(P1, 1.02 * P0, rtol=1e-14, atol=1e-12)
    assert np.allclose(inner1, inner0, atol=1e-12)
    assert b.lrc_species() == pytest.approx(lrc0 * V0 / float(np.prod(b.L)), rel=1e-12)
    # the acceptance test itself is pinned against the reference's
    # acceptVolumeMove in test_npt_cpu.py
