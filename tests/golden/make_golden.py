"""Writes the golden fixtures under tests/golden/ from the reference's own
CPU path (oracle/_ref/ref_trace, built from the reference sources by
oracle/Makefile).  Run in a container that has the reference tree:

    python -m mcgpu_amd.build && python tests/golden/make_golden.py

For each case: the input box as a .state file (full precision) and the
reference's per-move trace (molecule, dE, accepted) plus its start/final
energies and accepted count, as a compressed .npz.

Note on seeds: the reference's .state reader loses the seed token (its
environment-line parser calls strtok re-entrantly), so every .state run of
the reference is seeded with srand(0), which glibc treats as srand(1).  The
fixtures therefore use MC seed 1.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import mcgpu_amd as m  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "ref_trace"

CASES = {
    # config-0 parity case: 256 TIP3P, L 19.73, rc 9, 298 K, 10,000 moves
    "water256": (lambda: m.workload("water256"), 10000, None),
    # 16-site rigid molecules, rc 12 (the config-1 generator at 200 molecules)
    "bicyclic200": (lambda: m.HostBox.generate(
        m.BICYCLIC, 200, density=1.10, cutoff=12.0, temperature=298.0,
        max_translate=0.15, max_rotate=15.0, seed=77, mc_seed=1), 3000, None),
    # two species (water + 6-site methanol), rc 11
    "mixture300": (lambda: m.HostBox.generate(
        m.MIXTURE, 300, n_second=120, density=0.90, cutoff=11.0,
        temperature=298.0, max_translate=0.15, max_rotate=15.0, seed=78,
        mc_seed=1), 3000, None),
    # two primary sites per molecule ("[1,2]": O and the first H)
    "water128_mp": (lambda: m.HostBox.generate(
        m.WATER, 128, density=0.997, cutoff=7.0, temperature=298.0,
        max_translate=0.15, max_rotate=15.0, seed=79, mc_seed=1), 3000, "[1,2]"),
}


def parse_trace(path: Path) -> dict:
    head = {}
    rows = []
    for line in path.read_text().splitlines():
        if line.startswith("#"):
            tok = line[1:].split()
            for k, v in zip(tok[0::2], tok[1::2]):
                head[k] = float(v)
        else:
            rows.append(line.split())
    a = np.array(rows, dtype=object)
    return dict(
        step=a[:, 0].astype(np.int64), mol=a[:, 1].astype(np.int32),
        old_e=a[:, 2].astype(np.float64), new_e=a[:, 3].astype(np.float64),
        de=a[:, 4].astype(np.float64), accepted=a[:, 5].astype(np.int8),
        e0=np.float64(head["E0"]), e0_pairs=np.float64(head["E0_pairs"]),
        lrc=np.float64(head["LRC"]), final=np.float64(head["final"]),
        n_accepted=np.int64(head["accepted"]))


def main() -> None:
    if not REF.exists():
        raise SystemExit(f"{REF} missing: run `make -C oracle ref` first")
    for name, (make, steps, primary) in CASES.items():
        state = HERE / f"{name}.state"
        box = make()
        box.write_state(str(state))
        if primary is not None:  # rewrite the primary definition token
            lines = state.read_text().splitlines(keepends=True)
            tok = lines[0].split()
            tok[9] = primary
            lines[0] = " ".join(tok) + " \n"
            state.write_text("".join(lines))
        out = HERE / f"_{name}.trace.txt"
        subprocess.run([str(REF), str(state), str(steps), str(out)], check=True,
                       stdout=subprocess.DEVNULL)
        t = parse_trace(out)
        out.unlink()
        np.savez_compressed(HERE / f"{name}.ref.npz", mol=t["mol"], de=t["de"],
                            old_e=t["old_e"], new_e=t["new_e"],
                            accepted=t["accepted"], e0=t["e0"],
                            e0_pairs=t["e0_pairs"], lrc=t["lrc"],
                            final=t["final"], n_accepted=t["n_accepted"],
                            steps=np.int64(steps))
        print(f"{name}: {steps} moves, E0 {float(t['e0']):.6f}, final "
              f"{float(t['final']):.6f}, accepted {int(t['n_accepted'])}")


if __name__ == "__main__":
    main()
