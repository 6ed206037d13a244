"""NPT parts on the CPU: the volume-move acceptance test against the
reference's own SimCalcs::acceptVolumeMove (built from its sources into
oracle/_ref), the closed forms that fix that test, the K-species LRC against
the textbook single-species value and a numerical quadrature of the LJ tail,
the run-LRC policy of the library, and the oracle's NPT loop invariants."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate

import mcgpu_amd as m
from oracle import oracle as O
from edge_boxes import ArrayBox

GOLDEN = Path(__file__).resolve().parent / "golden"
RAND_MAX = 2147483647


def first_uniform(seed):
    """u = randomReal(0, 1) right after srand(seed) (MathLibrary.cpp:27-33)."""
    return O.GlibcRng(seed).rand() / RAND_MAX


def accept_prob(n, v0, v1, de, p, kT):
    return O.lib().orc_volume_accept_prob(int(n), float(v0), float(v1), float(de),
                                          float(p), float(kT))


# ------------------------------------------------- volume-move acceptance

def volume_grid(n_rows=600, seed=4):
    """Rows (seed, N, kT, V, Lx', Ly', Lz', dE, P): half random, half with dE
    placed 2% (in units of kT) either side of the decision boundary, so
    that every term of the test -- N ln(V'/V), dE and P dV, each against kT
    -- decides some rows."""
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(n_rows):
        s = int(rng.integers(1, 2**31 - 1))
        n = int(rng.choice([1, 37, 256, 300, 4000]))
        kT = float(rng.choice([0.4968, 0.5922, 0.7949]))
        L0 = rng.uniform(10.0, 60.0, 3)
        v0 = L0[0] * L0[1] * L0[2]
        f = float(rng.uniform(0.99, 1.01))
        L1 = L0 * f
        v1 = L1[0] * L1[1] * L1[2]
        p = float(rng.choice([0.0, 1.4584e-5, 1e-3, 2e-2]))
        if k % 2 == 0:
            de = float(rng.uniform(-5.0, 5.0))
        else:
            u = first_uniform(s)
            eps = 0.02 if k % 4 == 1 else -0.02
            de = kT * (n * (math.log(v1) - math.log(v0)) - math.log(u) + eps) - p * (v1 - v0)
        rows.append((s, n, kT, v0, L1[0], L1[1], L1[2], de, p))
    return rows


def test_volume_decisions_match_reference_acceptVolumeMove():
    rows = volume_grid()
    ref = O.reference_volume_decisions(GOLDEN / "water256.state", rows)
    if ref is None:
        pytest.skip("reference build (oracle/_ref/ref_volume) absent")
    assert len(ref) == len(rows)
    mine = np.array([int(accept_prob(n, v0, Lx * Ly * Lz, de, p, kT) >= first_uniform(s))
                     for (s, n, kT, v0, Lx, Ly, Lz, de, p) in rows], np.int8)
    assert np.array_equal(mine, ref)
    # the grid decides both ways, also among the rows next to the boundary
    assert 0.2 < ref.mean() < 0.8 and 0.2 < ref[1::2].mean() < 0.8


def test_volume_decisions_planted_sign_error_is_caught():
    """A P dV term of the wrong sign (planted in a copy of the test on the
    CPU) must disagree with the reference on the grid."""
    rows = volume_grid()
    ref = O.reference_volume_decisions(GOLDEN / "water256.state", rows)
    if ref is None:
        pytest.skip("reference build (oracle/_ref/ref_volume) absent")
    wrong = []
    for (s, n, kT, v0, Lx, Ly, Lz, de, p) in rows:
        v1 = Lx * Ly * Lz
        prob = math.exp(n * (math.log(v1) - math.log(v0)) - (de - p * (v1 - v0)) / kT)
        wrong.append(int(prob >= first_uniform(s)))
    assert not np.array_equal(np.array(wrong, np.int8), ref)


def test_volume_accept_prob_closed_forms():
    kT = 0.5922
    # V' = V: exp(-dE/kT) whatever N and P
    for n in (1, 300):
        for p in (0.0, 1e-3):
            for de in (-1.0, 0.0, 0.7):
                assert accept_prob(n, 1000.0, 1000.0, de, p, kT) == pytest.approx(
                    math.exp(-de / kT), rel=1e-15)
    # dE = 0, P = 0: (V'/V)^N
    assert accept_prob(300, 1000.0, 1010.0, 0.0, 0.0, kT) == pytest.approx(1.01 ** 300,
                                                                            rel=1e-12)
    # falls as P grows for V' > V, rises for V' < V
    ps = [0.0, 1e-5, 1e-4, 1e-3]
    up = [accept_prob(300, 1000.0, 1010.0, 0.3, p, kT) for p in ps]
    down = [accept_prob(300, 1000.0, 990.0, 0.3, p, kT) for p in ps]
    assert all(a > b for a, b in zip(up, up[1:]))
    assert all(a < b for a, b in zip(down, down[1:]))
    # P dV enters as dE does: P = 1e-3 over dV = 10 equals dE + 0.01
    assert accept_prob(300, 1000.0, 1010.0, 0.3, 1e-3, kT) == pytest.approx(
        accept_prob(300, 1000.0, 1010.0, 0.31, 0.0, kT), rel=1e-12)


# ------------------------------------------------------------------- LRC

def species_box(kinds, counts, edge=30.0, cutoff=9.0):
    """A box of molecule kinds [(sigma[], eps[], q[])] with counts; positions
    on a coarse lattice (the LRC reads only parameters, counts, V, rc)."""
    sig, eps, q, ms, ml = [], [], [], [], []
    for (s_, e_, q_), c in zip(kinds, counts):
        for _ in range(c):
            ms.append(len(sig))
            ml.append(len(s_))
            sig += list(s_)
            eps += list(e_)
            q += list(q_)
    n, ns = len(ms), len(sig)
    g = np.arange(ns, dtype=np.float64)
    a = dict(x=(g * 0.37) % edge, y=(g * 0.71) % edge, z=(g * 0.13) % edge,
             sigma=np.array(sig), epsilon=np.array(eps), charge=np.array(q),
             mol_start=np.array(ms, np.int32), mol_len=np.array(ml, np.int32),
             mol_pidx_start=np.arange(n, dtype=np.int32),
             mol_pidx_count=np.ones(n, np.int32),
             primary_index=np.array(ms, np.int32),
             box=np.array([edge, edge, edge]), cutoff=cutoff, kT=0.5922,
             max_translate=0.15, max_rotate=15.0)
    return a


def product_lrc(a, fn="mcgpu_lrc_species"):
    e = C.c_double()
    d = ArrayBox(a, 1).desc()
    m.check(getattr(m.lib(), fn)(C.byref(d), C.byref(e)), fn)
    return e.value


def textbook_one_species(n, v, rc, sig, eps):
    return 2 * math.pi * n * n / v * 4 * eps * sig ** 6 * (
        sig ** 6 / (9 * rc ** 9) - 1 / (3 * rc ** 3))


O_SITE = (3.15061, 0.1521)
WATER = ([3.15061, 0.0, 0.0], [0.1521, 0.0, 0.0], [-0.834, 0.417, 0.417])
# another species with the same LJ content: one O site, other sites without
# LJ (a 4-site model) or dummy (sigma = -1)
WATER4 = ([3.15061, 0.0, 0.0, 0.0], [0.1521, 0.0, 0.0, 0.0], [0.0, 0.52, 0.52, -1.04])
WATER_DUMMY = ([3.15061, -1.0, 0.0, 0.0], [0.1521, -1.0, 0.0, 0.0],
               [-0.834, 0.0, 0.417, 0.417])


@pytest.mark.parametrize("other", [WATER4, WATER_DUMMY], ids=["4-site", "dummy-site"])
def test_lrc_species_identical_lj_sites_give_one_species(other):
    """Two species whose LJ sites are identical are one species for the
    tail: the K-species form must give the textbook value for N_A + N_B."""
    a = species_box([WATER, other], [120, 80])
    v = float(np.prod(a["box"]))
    want = textbook_one_species(200, v, a["cutoff"], *O_SITE)
    assert O.OracleBox(a).lrc_species() == pytest.approx(want, rel=1e-12)
    assert product_lrc(a) == pytest.approx(want, rel=1e-12)


def lrc_quadrature(kinds, counts, v, rc):
    """sum_AB N_A N_B / (2V) sum_ab int_rc^inf 4 pi r^2 u_ab(r) dr, with u_ab
    the oracle's LJ pair energy (pinned by its closed forms)."""
    lj = O.lib().orc_lj_params
    total = 0.0
    for (sa, ea, _), na in zip(kinds, counts):
        for (sb, eb, _), nb in zip(kinds, counts):
            for s1, e1 in zip(sa, ea):
                for s2, e2 in zip(sb, eb):
                    if s1 < 0 or e1 < 0 or s2 < 0 or e2 < 0 or e1 * e2 == 0:
                        continue
                    f = lambda r: 4 * math.pi * r * r * lj(s1, e1, s2, e2, r * r)
                    val, _ = integrate.quad(f, rc, np.inf, epsabs=0, epsrel=1e-13,
                                            limit=200)
                    total += na * nb / (2 * v) * val
    return total


def test_lrc_species_matches_tail_quadrature():
    """A two-species box of different LJ sites (methanol-like C/O plus a
    2-site kind): the closed form, its mixing rules and its species counts
    against a numerical integral of the pair potential."""
    meoh = ([3.5, 3.12, 0.0, 2.5, 2.5, 2.5], [0.066, 0.17, 0.0, 0.03, 0.03, 0.03],
            [0.145, -0.683, 0.418, 0.04, 0.04, 0.04])
    other = ([2.9, -1.0], [0.09, -1.0], [0.0, 0.0])
    kinds, counts = [WATER, meoh, other], [150, 90, 40]
    a = species_box(kinds, counts, edge=33.0, cutoff=11.0)
    v = float(np.prod(a["box"]))
    want = lrc_quadrature(kinds, counts, v, a["cutoff"])
    assert O.OracleBox(a).lrc_species() == pytest.approx(want, rel=1e-9)
    assert product_lrc(a) == pytest.approx(want, rel=1e-9)
    # the reference's two-solvent formula counts N/2 of molecules 0's and 1's
    # kinds (both water here): far from the mixture's tail
    assert abs(product_lrc(a, "mcgpu_lrc_reference") - want) > 0.1 * abs(want)


def test_run_lrc_policy():
    """NVT runs start from the reference's LRC (start energies match the
    reference's); NPT runs from the K-species form."""
    hb = m.HostBox.from_state(str(GOLDEN / "mixture300.state"))
    d = hb.desc()
    ob = O.OracleBox(hb.arrays())
    assert m.lrc_for_run(d, m.NVT) == ob.lrc_reference()
    assert m.lrc_for_run(d, m.NPT) == pytest.approx(ob.lrc_species(), rel=1e-13)


# ------------------------------------------------------- oracle NPT loop

def test_oracle_npt_loop_invariants():
    """The oracle's NPT loop: the running energy equals a fresh evaluation
    at the final volume; every recorded decision is the acceptance test;
    volume moves are both accepted and rejected."""
    hb = m.HostBox.from_state(str(GOLDEN / "mixture300.state"))
    a = hb.arrays()
    v0 = float(a["box"][0] * a["box"][1] * a["box"][2])
    ob = O.OracleBox(a)
    p = 1.4584e-5
    fin, e0, nacc, t, v = ob.run_npt(1, 2000, 100, p, 0.002 * v0)
    assert len(v["step"]) == 19 and list(v["step"]) == list(range(100, 2000, 100))
    assert 0 < v["accepted"].sum() < len(v["accepted"])
    fresh = O.OracleBox(ob.a)
    fresh.box.L[:] = ob.box.L[:]
    assert fin == pytest.approx(fresh.system_energy() + fresh.lrc_species(), rel=1e-10)
    n = len(a["mol_start"])
    kT = float(a["kT"])
    for k in range(len(v["step"])):
        prob = accept_prob(n, v["v_old"][k], v["v_new"][k], v["e_new"][k] - v["e_old"][k],
                           p, kT)
        assert int(prob >= v["u"][k]) == v["accepted"][k]
    # without volume moves the loop is the NVT loop (LRC apart)
    ob2 = O.OracleBox(a)
    fin2, e02, _, t2, v2 = ob2.run_npt(1, 500, 1000, p, 0.002 * v0)
    ob3 = O.OracleBox(a)
    fin3, e03, _, t3 = ob3.run_nvt(1, 500, use_libc=False)
    assert len(v2["step"]) == 0
    assert np.array_equal(t2["mol"], t3["mol"]) and np.array_equal(t2["de"], t3["de"])
    assert fin2 - e02 == pytest.approx(fin3 - e03, rel=1e-12)
