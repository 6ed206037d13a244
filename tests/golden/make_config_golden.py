"""Writes the .config-route fixtures under tests/golden/config/ from the
reference's own CPU path (oracle/_ref/ref_load and ref_trace, built from the
reference sources by oracle/Makefile).  Run in a container that has the
reference tree:

    make -C oracle ref && python tests/golden/make_config_golden.py [--full]

The cases are the reference's integration tests (test/unittests/
Integration/*, MultipleSolvents/*): the same z-matrix files (copied here as
data), box edges, molecule counts, cut-offs, move sizes, seed 12345 and
primary-atom strings, each with the Final-Energy the test expects and its
margin.  For each case:

* ``<case>.load.npz``: the box the reference builds from the .config (FCC
  start, keepMoleculeInBox, SimBoxBuilder): sha256 digests of the site and
  molecule arrays (bit patterns), the first sites in full, the environment,
  the start pair energy and the LRC;
* the reference's per-move trace of the first moves (molecule, old and
  new molecule energy, dE, accepted) in the same file;
* with --full, ``reference_runs.json``: the reference's full run (the
  test's step count) -- final energy and accepted moves.

``oplsaa_subset.par`` holds the title and column rows of the reference's
resources/bossFiles/oplsaa.par and every row whose id is an OPLS type the
z-matrices use (atom rows and torsion rows alike, in file order).  The
reference itself runs on the full file (its loader also needs oplsaa.sb
beside it); the loader under test reads the subset, so equal boxes pin both
the row selection and the parse.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
CFG = HERE / "config"
ROOT = HERE.parents[1]
REF_LOAD = ROOT / "oracle" / "_ref" / "ref_load"
REF_TRACE = ROOT / "oracle" / "_ref" / "ref_trace"
REFTREE = Path("/root/reference")
FULL_PAR = REFTREE / "resources" / "bossFiles" / "oplsaa.par"

# name: (z-matrix fixture, edge, molecules, steps, cutoff, max-translation,
#        max-rotation, primary-atom, expected Final-Energy, margin,
#        trace moves, reference test)
CASES = {
    "water5120": ("watt4p.z", 55.0, 5120, 1000, 25.0, 0.15, 15.0, "1",
                  -14000.0, 250.0, 1000, "WaterTest.cpp:54"),
    "indole267": ("indole.z", 35.36, 267, 40000, 12.0, 0.12, 12.0, "1",
                  1832460.0, 100.0, 2000, "IndoleTest.cpp:42"),
    "meoh256": ("meoh.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1",
                91.0, 100.0, 2000, "MethanolIntegrationTest.cpp:43"),
    "indole4000_1": ("indole4000.z", 87.17, 4000, 100000, 12.0, 0.06, 6.0, "1",
                     47518400.0, 100.0, 2000, "indole4000Test.cpp:56"),
    "indole4000_2": ("indole4000.z", 87.17, 4000, 100000, 12.0, 0.06, 6.0,
                     "[1,2]", 47463200.0, 100.0, 2000, "indole4000Test.cpp:112"),
    "meoh500_1": ("meoh500.z", 32.91, 500, 100000, 11.0, 0.12, 12.0, "1",
                  1093.0, 100.0, 2000, "Methanol500Test.cpp:59"),
    "meoh500_2": ("meoh500.z", 32.91, 500, 100000, 11.0, 0.12, 12.0, "[1,2]",
                  1070.0, 100.0, 2000, "Methanol500Test.cpp:112"),
    "meoh_t3p256": ("meoh-t3p.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0,
                    "[1],[1]", -2007.0, 100.0, 2000, "meoh-t3pTest.cpp:30"),
    "meoh_t3p500": ("meoh-t3p.z", 32.91, 500, 100000, 11.0, 0.12, 12.0,
                    "[1],[1]", -3148.0, 100.0, 2000, "meoh-t3pTest.cpp:56"),
    "meoh_t3p8788": ("meoh-t3p.z", 85.57, 8788, 100000, 11.0, 0.12, 12.0,
                     "[1],[1]", -14118.0, 100.0, 2000, "meoh-t3pTest.cpp:82"),
    "meohtip3_a": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "1,2",
                   9803.0, 100.0, 2000, "meohtip3Test.cpp:98"),
    "meohtip3_b": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0,
                   "[1,2],[3,4]", 9770.0, 100.0, 2000, "meohtip3Test.cpp:155"),
    "meohtip3_c": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0,
                   "1,[1,2]", 9560.0, 100.0, 2000, "meohtip3Test.cpp:183"),
    "meohtip3_d": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0,
                   "[1,2],1", 9553.0, 100.0, 2000, "meohtip3Test.cpp:210"),
    "t3pdim_a": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1,2",
                 -1832.0, 100.0, 2000, "t3pdimTest.cpp:104"),
    "t3pdim_b": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0,
                 "[1,2],[3,4]", -1790.0, 100.0, 2000, "t3pdimTest.cpp:158"),
    "t3pdim_c": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1,[1,2]",
                 -1770.0, 100.0, 2000, "t3pdimTest.cpp:187"),
    "t3pdim_d": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "[1,2],1",
                 -1800.0, 100.0, 2000, "t3pdimTest.cpp:215"),
    "t3pdimone_1": ("t3pdimone.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1",
                    -1810.0, 100.0, 2000, "t3pdimoneTest.cpp:55"),
    "t3pdimone_2": ("t3pdimone.z", 26.15, 256, 100000, 11.0, 0.12, 12.0,
                    "[1,2]", -1770.0, 100.0, 2000, "t3pdimoneTest.cpp:106"),
}

# Sites kept in full in each fixture (the rest is pinned by digests).
HEAD_SITES = 96


def cxx_num(v: float) -> str:
    """A double as std::ostream writes it by default (6 significant digits),
    which is how the reference's tests write their .config files."""
    return "%g" % v


def write_config(path: Path, case: tuple, par: Path, zdir: Path) -> None:
    z, edge, n, steps, rc, mt, mr, prim = case[:8]
    e = cxx_num(edge)
    path.write_text(
        "# Size of periodic box (x, y, z in angstroms)\n"
        f"x={e}\ny={e}\nz={e}\n"
        f"# Temperature in Kelvin\ntemp={cxx_num(298.15)}\n"
        f"# Max translation\nmax-translation={cxx_num(mt)}\n"
        f"# Number of steps\nsteps={steps}\n"
        f"# Number of molecules\nmolecules={n}\n"
        f"# Path to opla.par file\nopla.par={par}\n"
        f"# Path to z matrix file\nz-matrix={zdir / z}\n"
        f"# Path to state input\nstate-input={path.parent}\n"
        f"# Path to state output directory\nstate-output={path.parent}\n"
        f"# Path to pdb output directory\npdb-output={path.parent}\n"
        f"# Cutoff distance in angstroms\ncutoff={cxx_num(rc)}\n"
        f"# Max rotation\nmax-rotation={cxx_num(mr)}\n"
        "# Random Seed Input\nrandom-seed=12345\n"
        f"# Primary Atom Index\nprimary-atom={prim}")


def write_par_subset() -> None:
    used = set()
    for z in sorted(CFG.glob("*.z")):
        for line in z.read_text().splitlines()[1:]:
            t = line.split()
            if len(t) >= 11 and t[0].isdigit() and t[2] != "-1":
                used.add(t[2])
    rows = FULL_PAR.read_text().splitlines(keepends=True)
    keep = rows[:2] + [r for r in rows[2:] if r.split()[:1] and r.split()[0] in used]
    (CFG / "oplsaa_subset.par").write_text("".join(keep))


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def parse_dump(path: Path) -> dict:
    sites, mols, pidx = [], [], []
    env = energy = None
    names = []
    for line in path.read_text().splitlines():
        t = line.split()
        if t[0] == "site":
            sites.append([float(v) for v in t[1:7]])
            names.append(t[7])
        elif t[0] == "mol":
            mols.append([int(v) for v in t[1:6]])
        elif t[0] == "pidx":
            pidx.append(int(t[1]))
        elif t[0] == "env":
            env = t[1:]
        elif t[0] == "energy":
            energy = [float(v) for v in t[1:3]]
    s = np.array(sites, dtype=np.float64)
    m = np.array(mols, dtype=np.int32)
    return dict(
        x=s[:, 0].copy(), y=s[:, 1].copy(), z=s[:, 2].copy(),
        sigma=s[:, 3].copy(), epsilon=s[:, 4].copy(), charge=s[:, 5].copy(),
        mol_start=m[:, 0].copy(), mol_len=m[:, 1].copy(), mol_type=m[:, 2].copy(),
        mol_pidx_start=m[:, 3].copy(), mol_pidx_count=m[:, 4].copy(),
        primary_index=np.array(pidx, dtype=np.int32), names=names,
        env=env, e0_pairs=energy[0], lrc=energy[1])


def parse_trace(path: Path) -> dict:
    rows = [l.split() for l in path.read_text().splitlines()
            if l and not l.startswith("#")]
    a = np.array(rows, dtype=object)
    return dict(mol=a[:, 1].astype(np.int32), old_e=a[:, 2].astype(np.float64),
                new_e=a[:, 3].astype(np.float64), de=a[:, 4].astype(np.float64),
                accepted=a[:, 5].astype(np.int8))


def make_case(name: str, case: tuple, work: Path) -> None:
    cfg = work / f"{name}.config"
    # the reference reads the very z-matrix bytes the loader under test reads
    write_config(cfg, case, FULL_PAR, CFG)
    dump = work / f"{name}.dump"
    subprocess.run([str(REF_LOAD), str(cfg), str(dump)], cwd=work, check=True,
                   capture_output=True)
    d = parse_dump(dump)
    trace = work / f"{name}.trace"
    subprocess.run([str(REF_TRACE), str(cfg), str(case[10]), str(trace)],
                   cwd=work, check=True, capture_output=True)
    t = parse_trace(trace)
    env = d["env"]
    arrays = ["x", "y", "z", "sigma", "epsilon", "charge", "mol_start",
              "mol_len", "mol_type", "mol_pidx_start", "mol_pidx_count",
              "primary_index"]
    out = {f"sha_{k}": np.array(digest(d[k])) for k in arrays}
    h = min(HEAD_SITES, d["x"].size)
    out.update({f"head_{k}": d[k][:h] for k in
                ("x", "y", "z", "sigma", "epsilon", "charge")})
    out.update(
        n_molecules=np.int32(env[0]), n_sites=np.int32(env[1]),
        n_primary=np.int32(env[2]),
        box=np.array([float(env[3]), float(env[4]), float(env[5])]),
        cutoff=np.float64(env[6]), temperature=np.float64(env[7]),
        kT=np.float64(env[8]), max_translate=np.float64(env[9]),
        max_rotate=np.float64(env[10]), seed=np.int32(env[11]),
        steps=np.int64(env[12]), e0_pairs=np.float64(d["e0_pairs"]),
        lrc=np.float64(d["lrc"]), head_names=np.array(d["names"][:h]),
        trace_mol=t["mol"], trace_de=t["de"], trace_acc=t["accepted"],
        trace_old=t["old_e"], trace_new=t["new_e"])
    np.savez_compressed(CFG / f"{name}.load.npz", **out)
    print(f"{name}: {env[0]} molecules, {env[1]} sites, E0 {d['e0_pairs']:.6f}",
          flush=True)


def full_run(name: str, case: tuple, work: Path) -> tuple:
    cfg = work / f"{name}.config"
    r = subprocess.run([str(REF_TRACE), str(cfg), str(case[3]), "-"], cwd=work,
                       check=True, capture_output=True, text=True)
    fin = acc = rej = e0p = lrc = None
    for line in r.stdout.splitlines():
        t = line.split()
        if line.startswith("# final"):
            fin, acc, rej = float(t[2]), int(t[4]), int(t[6])
        elif line.startswith("# E0_pairs"):
            e0p, lrc = float(t[2]), float(t[4])
    return name, dict(steps=case[3], final_energy=fin, accepted=acc,
                      rejected=rej, e0_pairs=e0p, lrc=lrc, expected=case[8],
                      margin=case[9], test=case[11],
                      reference_meets_expected=abs(fin - case[8]) <= case[9])


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true",
                    help="also run every case for the test's step count")
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    for p in (REF_LOAD, REF_TRACE, FULL_PAR):
        if not p.exists():
            raise SystemExit(f"{p} missing (make -C oracle ref)")
    write_par_subset()
    with tempfile.TemporaryDirectory() as tmp:
        work = Path(tmp)
        with cf.ThreadPoolExecutor(args.jobs) as ex:
            list(ex.map(lambda kv: make_case(kv[0], kv[1], work), CASES.items()))
        if args.full:
            with cf.ThreadPoolExecutor(args.jobs) as ex:
                runs = dict(ex.map(lambda kv: full_run(kv[0], kv[1], work),
                                   CASES.items()))
            (CFG / "reference_runs.json").write_text(
                json.dumps(runs, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
