"""CPU tests of the .config input route (z-matrix + OPLS .par loader).

The cases are the reference's integration tests (test/unittests/Integration,
MultipleSolvents): same z-matrices, box edges, molecule counts, cut-offs,
move sizes, seed and primary-atom strings.  Each test writes its .config the
way the reference's TestUtil.cpp:4-46 does and loads it through the C ABI.

Pins, per case (fixtures from tests/golden/make_config_golden.py, i.e. the
reference's own loader and CPU path):
* the loaded box is bit-identical to the reference's (sha256 of every site
  and molecule array, environment, seed, step count);
* the oracle, run on the loaded box, reproduces the reference's per-move
  trace bit for bit and its start energy and LRC;
and, from geometry alone: the z-matrix bond lengths, angles and dihedrals
are rebuilt inside every molecule, and the molecules sit on the FCC lattice.
Error cases follow the reference's tests that expect loadBoxData() to fail.
"""
import hashlib
import math
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from mcgpu_amd.api import HostBox
from oracle import oracle as O

from config_cases import ARRAYS, CASES, CFG, PAR, SMALL, load_case, write_config


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", list(CASES))
def test_loaded_box_is_bit_identical_to_reference(name, tmp_path):
    ref = np.load(CFG / f"{name}.load.npz")
    hb, settings = load_case(name, tmp_path)
    a = hb.arrays()
    h = ref["head_x"].size
    for k in ("x", "y", "z", "sigma", "epsilon", "charge"):
        assert np.array_equal(a[k][:h], ref[f"head_{k}"]), k
    for k in ARRAYS:
        assert sha(a[k]) == str(ref[f"sha_{k}"]), k
    assert int(ref["n_molecules"]) == len(a["mol_start"])
    assert int(ref["n_sites"]) == len(a["x"])
    assert int(ref["n_primary"]) == len(a["primary_index"])
    assert np.array_equal(a["box"], ref["box"])
    assert a["cutoff"] == float(ref["cutoff"])
    assert a["kT"] == float(ref["kT"])
    assert a["max_translate"] == float(ref["max_translate"])
    assert a["max_rotate"] == float(ref["max_rotate"])
    assert hb.mc_seed == int(ref["seed"])
    assert hb.temperature == float(ref["temperature"])
    assert settings["steps"] == int(ref["steps"])
    assert settings["random_seed"] == int(ref["seed"])


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_on_loaded_box_reproduces_reference_trace(name, tmp_path):
    ref = np.load(CFG / f"{name}.load.npz")
    hb, _ = load_case(name, tmp_path)
    a = hb.arrays()
    b = O.OracleBox(a)
    lrc = b.lrc_reference()
    if a["mol_len"][0] == a["mol_len"][1]:
        assert lrc == float(ref["lrc"])
    else:
        # molecules 0 and 1 differ in length: the reference's LRC reads past
        # its arrays (SerialCalcs.cpp:98-104); the species form stands in
        assert lrc == b.lrc_species()
    n = ref["trace_mol"].size if name in SMALL else 300
    fin, e0, nacc, tr = b.run_nvt(hb.mc_seed, n, use_libc=True)
    assert e0 == float(ref["e0_pairs"]) + lrc
    assert np.array_equal(tr["mol"], ref["trace_mol"][:n])
    assert np.array_equal(tr["accepted"], ref["trace_acc"][:n])
    assert np.array_equal(tr["de"], ref["trace_de"][:n])


def _zrows(z: str):
    """Atom rows of a z-matrix fixture, per molecule type (ids, bond, angle,
    dihedral partners and values), read independently of the loader."""
    types, cur = [], []
    for line in (CFG / z).read_text().splitlines()[1:]:
        t = line.split()
        if len(t) >= 11 and t[0].isdigit():
            cur.append(dict(id=int(t[0]), b=int(t[4]), bd=float(t[5]), a=int(t[6]),
                            av=float(t[7]), d=int(t[8]), dv=float(t[9])))
        elif "TERZ" in line or "Geometry Variations" in line:
            types.append(cur)
            cur = []
    return types


def _unwrapped(a, mol):
    s, n = int(a["mol_start"][mol]), int(a["mol_len"][mol])
    p = np.stack([a["x"][s:s + n], a["y"][s:s + n], a["z"][s:s + n]], axis=1)
    return p


def _angle(p, q, r):
    u, v = p - q, r - q
    c = np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def _dihedral(p0, p1, p2, p3):
    """Torsion p0-p1-p2-p3 in degrees (0 = cis), IUPAC sign."""
    b1, b2, b3 = p1 - p0, p2 - p1, p3 - p2
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    m1 = np.cross(n1, b2 / np.linalg.norm(b2))
    return math.degrees(math.atan2(np.dot(m1, n2), np.dot(n1, n2)))


@pytest.mark.parametrize("name", ["meoh256", "indole267", "water5120", "meoh_t3p256",
                                  "t3pdim_a"])
def test_internal_coordinates_follow_the_zmatrix(name, tmp_path):
    """Bond lengths, angles and torsions of every atom row are rebuilt (the
    reference converts degrees with pi = 3.14159265, so angles agree with
    the rows' values scaled by 3.14159265/pi)."""
    hb, _ = load_case(name, tmp_path)
    a = hb.arrays()
    types = _zrows(CASES[name][0])
    off = 0
    for t, rows in enumerate(types):
        p = _unwrapped(a, t)          # molecule t is the first of its type
        ids = {r["id"]: k for k, r in enumerate(rows)}
        for k, r in enumerate(rows):
            if k == 0:
                continue
            if r["b"] in ids:
                assert np.linalg.norm(p[k] - p[ids[r["b"]]]) == pytest.approx(
                    r["bd"], abs=1e-9)
            if k >= 2 and r["a"] in ids and r["b"] in ids:
                assert _angle(p[k], p[ids[r["b"]]], p[ids[r["a"]]]) == pytest.approx(
                    r["av"] * 3.14159265 / math.pi, abs=1e-6)
            if k >= 3 and all(x in ids for x in (r["b"], r["a"], r["d"])):
                got = _dihedral(p[k], p[ids[r["b"]]], p[ids[r["a"]]], p[ids[r["d"]]])
                want = r["dv"] * 3.14159265 / math.pi
                # the z-matrix torsion turns the opposite way to IUPAC's
                assert abs((got + want + 180) % 360 - 180) < 1e-6, (k, got, want)
        off += len(rows)


@pytest.mark.parametrize("name", ["meoh256", "meoh_t3p500", "indole267", "water5120"])
def test_molecules_sit_on_the_fcc_lattice(name, tmp_path):
    """Every type's first site is built at the origin, so molecule i's first
    site is its FCC point: unit-cell site i mod 4 of cell i div 4 (x fastest),
    shifted back by half a cell and wrapped into [0, L] by that same site."""
    hb, _ = load_case(name, tmp_path)
    a = hb.arrays()
    n = len(a["mol_start"])
    L = a["box"]
    cells = int((0.25 * n) ** (1 / 3) + 0.5)
    while 4 * cells ** 3 < n:
        cells += 1
    cl = L[0] / cells
    unit = np.array([[0, 0, 0], [.5, .5, 0], [0, .5, .5], [.5, 0, .5]]) * cl
    for i in range(n):
        s = int(a["mol_start"][i])
        c = i // 4
        want = (np.array([c % cells, (c // cells) % cells, c // (cells * cells)]) * cl
                + unit[i % 4] - 0.5 * cl)
        want = np.where(want < 0, want + L, np.where(want > L, want - L, want))
        got = np.array([a["x"][s], a["y"][s], a["z"][s]])
        assert np.allclose(got, want, atol=1e-9), i


@pytest.mark.parametrize("z,prim,edge,n", [
    ("meohtip3.z", "1", 23.7856, 256),          # meohtip3Test.cpp:51-55
    ("meohtip3.z", "[1,2]", 23.7856, 256),      # meohtip3Test.cpp:73-77
    ("t3pdim.z", "1", 26.15, 256),              # t3pdimTest.cpp:54-58
    ("t3pdim.z", "[1,2]", 26.15, 256),          # t3pdimTest.cpp:78-82
    ("indole4000.z", "1,2", 87.17, 4000),       # indole4000Test.cpp:164-169
    ("meoh500.z", "[1,2],[3,4]", 32.91, 500),   # Methanol500Test.cpp:185-189
])
def test_primary_definitions_must_match_molecule_types(z, prim, edge, n, tmp_path):
    cfg = write_config(tmp_path / "bad.config", z, edge, n, 100, 11.0, 0.12, 12.0, prim)
    with pytest.raises(m.McgpuError, match="primary-atom"):
        HostBox.from_config(str(cfg))


def test_config_errors_are_reported(tmp_path):
    base = ("meoh.z", 26.15, 256, 100, 11.0, 0.12, 12.0, "1")
    cfg = write_config(tmp_path / "k.config", *base, extra="colour=blue\n")
    with pytest.raises(m.McgpuError, match="unexpected key"):
        HostBox.from_config(str(cfg))
    with pytest.raises(m.McgpuError, match="unable to open configuration"):
        HostBox.from_config(str(tmp_path / "missing.config"))
    cfg = write_config(tmp_path / "p.config", *base, par=tmp_path / "none.par")
    with pytest.raises(m.McgpuError, match="OPLS file"):
        HostBox.from_config(str(cfg))
    # an OPLS type that the .par file lacks
    par = tmp_path / "tiny.par"
    par.write_text(PAR.read_text().splitlines(keepends=True)[0] +
                   "".join(l for l in PAR.read_text().splitlines(keepends=True)[1:]
                           if not l.split()[:1] == ["157"]))
    cfg = write_config(tmp_path / "t.config", *base, par=par)
    with pytest.raises(m.McgpuError, match="OPLS type 157"):
        HostBox.from_config(str(cfg))
    # a line that is neither comment nor key=value
    cfg = write_config(tmp_path / "w.config", *base, extra="[section]\n")
    with pytest.raises(m.McgpuError, match="not well formed"):
        HostBox.from_config(str(cfg))


def test_config_settings_are_reported(tmp_path):
    cfg = write_config(tmp_path / "s.config", "meoh.z", 26.15, 256, 777, 11.0, 0.12,
                       12.0, "1", seed="99",
                       extra="sim-name=demo\nstrategy=brute-force\nensemble=NPT\n"
                             "pressure=1.5\nvolume-interval=40\n")
    hb, s = HostBox.from_config(str(cfg))
    assert s["steps"] == 777 and s["random_seed"] == 99 and hb.mc_seed == 99
    assert s["sim_name"] == "demo" and s["strategy"] == "brute-force"
    assert s["ensemble"] == "NPT" and s["pressure"] == 1.5
    assert s["volume_interval"] == 40
    assert s["state_output"] == str(tmp_path)


def test_molecule_count_is_kept_when_not_a_multiple_of_the_types(tmp_path):
    """255 molecules of a two-type z-matrix: the reference builds 256 but
    simulates the 255 its environment asks for (FileUtilities.cpp:296-300,
    SimBoxBuilder.cpp:30)."""
    cfg = write_config(tmp_path / "odd.config", "meoh-t3p.z", 23.7856, 255, 10,
                       11.0, 0.12, 12.0, "[1],[1]")
    hb, s = HostBox.from_config(str(cfg))
    a = hb.arrays()
    assert len(a["mol_start"]) == 255 and s["n_types"] == 2
    full, _ = load_case("meoh_t3p256", tmp_path)
    b = full.arrays()
    ns = len(a["x"])
    assert np.array_equal(a["x"], b["x"][:ns]) and np.array_equal(a["z"], b["z"][:ns])
