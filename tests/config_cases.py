"""The reference's integration-test inputs as .config cases, shared by the
CPU and GPU tests of the .config route (test/unittests/Integration,
MultipleSolvents; fixtures from tests/golden/make_config_golden.py)."""
from pathlib import Path

from mcgpu_amd.api import HostBox

CFG = Path(__file__).resolve().parent / "golden" / "config"
PAR = CFG / "oplsaa_subset.par"

# name: (z-matrix, edge, molecules, steps, cutoff, max-translation,
#        max-rotation, primary-atom)  -- tests/golden/make_config_golden.py
CASES = {
    "water5120": ("watt4p.z", 55.0, 5120, 1000, 25.0, 0.15, 15.0, "1"),
    "indole267": ("indole.z", 35.36, 267, 40000, 12.0, 0.12, 12.0, "1"),
    "meoh256": ("meoh.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1"),
    "indole4000_1": ("indole4000.z", 87.17, 4000, 100000, 12.0, 0.06, 6.0, "1"),
    "indole4000_2": ("indole4000.z", 87.17, 4000, 100000, 12.0, 0.06, 6.0, "[1,2]"),
    "meoh500_1": ("meoh500.z", 32.91, 500, 100000, 11.0, 0.12, 12.0, "1"),
    "meoh500_2": ("meoh500.z", 32.91, 500, 100000, 11.0, 0.12, 12.0, "[1,2]"),
    "meoh_t3p256": ("meoh-t3p.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "[1],[1]"),
    "meoh_t3p500": ("meoh-t3p.z", 32.91, 500, 100000, 11.0, 0.12, 12.0, "[1],[1]"),
    "meoh_t3p8788": ("meoh-t3p.z", 85.57, 8788, 100000, 11.0, 0.12, 12.0, "[1],[1]"),
    "meohtip3_a": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "1,2"),
    "meohtip3_b": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "[1,2],[3,4]"),
    "meohtip3_c": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "1,[1,2]"),
    "meohtip3_d": ("meohtip3.z", 23.7856, 256, 100000, 11.0, 0.12, 12.0, "[1,2],1"),
    "t3pdim_a": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1,2"),
    "t3pdim_b": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "[1,2],[3,4]"),
    "t3pdim_c": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1,[1,2]"),
    "t3pdim_d": ("t3pdim.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "[1,2],1"),
    "t3pdimone_1": ("t3pdimone.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "1"),
    "t3pdimone_2": ("t3pdimone.z", 26.15, 256, 100000, 11.0, 0.12, 12.0, "[1,2]"),
}
SMALL = [k for k, v in CASES.items() if v[2] <= 500]
ARRAYS = ["x", "y", "z", "sigma", "epsilon", "charge", "mol_start", "mol_len",
          "mol_pidx_start", "mol_pidx_count", "primary_index"]


def write_config(path: Path, z: str, edge: float, n: int, steps: int, rc: float,
                 mt: float, mr: float, prim: str, seed: str = "12345",
                 extra: str = "", par: Path = PAR) -> Path:
    """A .config as TestUtil.cpp:4-46 writes it (doubles as ostream's %g)."""
    g = lambda v: "%g" % v
    path.write_text(
        "# Size of periodic box (x, y, z in angstroms)\n"
        f"x={g(edge)}\ny={g(edge)}\nz={g(edge)}\n"
        f"# Temperature in Kelvin\ntemp={g(298.15)}\n"
        f"# Max translation\nmax-translation={g(mt)}\n"
        f"# Number of steps\nsteps={steps}\n"
        f"# Number of molecules\nmolecules={n}\n"
        f"# Path to opla.par file\nopla.par={par}\n"
        f"# Path to z matrix file\nz-matrix={CFG / z}\n"
        f"# Path to state input\nstate-input={path.parent}\n"
        f"# Path to state output directory\nstate-output={path.parent}\n"
        f"# Path to pdb output directory\npdb-output={path.parent}\n"
        f"# Cutoff distance in angstroms\ncutoff={g(rc)}\n"
        f"# Max rotation\nmax-rotation={g(mr)}\n"
        f"# Random Seed Input\nrandom-seed={seed}\n"
        f"{extra}"
        f"# Primary Atom Index\nprimary-atom={prim}")
    return path


def load_case(name: str, tmp_path: Path):
    cfg = write_config(tmp_path / f"{name}.config", *CASES[name])
    return HostBox.from_config(str(cfg))
