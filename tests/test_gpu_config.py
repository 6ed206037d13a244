"""GPU parity of the .config route: boxes loaded from the reference's own
integration-test inputs (z-matrix + OPLS .par, FCC start) must reproduce the
reference's CPU path on the same .config -- its per-move trace and, over the
test's full step count, its final energy and accepted-move count.

The reference's tests also publish an expected Final-Energy with a margin
(EXPECT_NEAR, +-100 or +-250).  The reference's current code, compiled here,
meets 9 of those 20 expectations; the rest are stale.  Where the reference
meets its expectation, this build must meet it too.

Tolerances: trial-move decisions and molecule choices must be identical; dE
within 1e-9 kcal/mol plus 1e-11 of max(|E_old|, |E_new|) of the move (the
sums run in another order than the reference's, and FCC starts overlap
molecules, so molecule energies reach 1e19 kcal/mol at trial positions).
Start and final pair energies within 1e-6 plus 1e-11 of the start energy's
magnitude, for the same reason.
"""
import json

import numpy as np
import pytest

import mcgpu_amd as m
from config_cases import CASES, CFG, load_case

pytestmark = pytest.mark.gpu
RUNS = json.loads((CFG / "reference_runs.json").read_text())


def de_ok(got, ref, n=None):
    """dE = E_new - E_old, each a sum of pair terms: the bound scales with
    the larger of the two energies, not with dE (FCC starts overlap
    molecules, so E reaches 1e19 kcal/mol at trial positions while dE may
    be small)."""
    n = ref["trace_de"].size if n is None else n
    want = ref["trace_de"][:n]
    scale = np.maximum(np.abs(ref["trace_old"][:n]), np.abs(ref["trace_new"][:n]))
    err = np.abs(got - want)
    print(f"max |dE err| {err.max():.3e}, max err/scale "
          f"{np.max(err / np.maximum(scale, 1e-300)):.3e}")
    return np.all(err <= 1e-9 + 1e-11 * scale)


def e_tol(e0):
    return 1e-6 + 1e-11 * abs(e0)


def lrc_of(hb):
    import ctypes as C
    d = hb.desc()
    e = C.c_double()
    m.check(m.lib().mcgpu_lrc_reference(C.byref(d), C.byref(e)))
    return e.value


def run_all(batch, slots, limit=100000):
    for _ in range(limit):
        stats = [batch.stats(s) for s in slots]
        if all(st.done for st in stats):
            return stats
        batch.run(5000)
    raise AssertionError("boxes did not finish")


def test_config_boxes_reproduce_reference_traces(tmp_path):
    """Every case in one batched launch (FP64, all molecule pairs tested)."""
    b = m.Batch(len(CASES) + 4)
    cases = {}
    for name in CASES:
        ref = np.load(CFG / f"{name}.load.npz")
        hb, _ = load_case(name, tmp_path)
        n = int(ref["trace_mol"].size)
        slot = b.add(hb, steps=n, record_trace=True)
        cases[name] = (slot, ref, lrc_of(hb))
    stats = run_all(b, [v[0] for v in cases.values()])
    for (name, (slot, ref, lrc)), st in zip(cases.items(), stats):
        print(f"[case] {name}", flush=True)
        tr = b.trace(slot)
        assert st.steps_done == ref["trace_mol"].size
        assert np.array_equal(tr.mol, ref["trace_mol"]), name
        assert np.array_equal(tr.accepted, ref["trace_acc"]), name
        assert de_ok(tr.de, ref), name
        e0 = float(ref["e0_pairs"])
        assert abs((st.start_energy - lrc) - e0) <= e_tol(e0), name


def test_config_full_runs_match_reference_final_energy(tmp_path):
    """The integration tests' full step counts (1,000 to 100,000 moves per
    box), every case at once."""
    b = m.Batch(len(CASES) + 4)
    cases = {}
    for name, case in CASES.items():
        hb, settings = load_case(name, tmp_path)
        assert settings["steps"] == RUNS[name]["steps"] == case[3]
        slot = b.add(hb, steps=settings["steps"])
        cases[name] = (slot, lrc_of(hb))
    stats = run_all(b, [v[0] for v in cases.values()])
    for (name, (slot, lrc)), st in zip(cases.items(), stats):
        r = RUNS[name]
        got = st.energy - lrc
        want = r["final_energy"] - r["lrc"]
        print(f"[case] {name}: final pairs {got:.6f} reference {want:.6f} "
              f"accepted {st.accepted} / {r['accepted']}", flush=True)
        assert st.accepted == r["accepted"], name
        assert abs(got - want) <= e_tol(r["e0_pairs"]), name
        if r["reference_meets_expected"]:
            assert abs(st.energy - r["expected"]) <= r["margin"], name


@pytest.mark.parametrize("name", ["meoh_t3p256", "t3pdim_c", "water5120"])
def test_plain_and_strategy_paths_on_config_boxes(name, tmp_path):
    ref = np.load(CFG / f"{name}.load.npz")
    n = min(int(ref["trace_mol"].size), 1000)
    hb, _ = load_case(name, tmp_path)
    lrc = lrc_of(hb)
    res, tr = m.plain_run(hb, n, precision=m.FP64)
    assert np.array_equal(tr.mol, ref["trace_mol"][:n])
    assert np.array_equal(tr.accepted, ref["trace_acc"][:n])
    assert de_ok(tr.de, ref, n)
    e0 = float(ref["e0_pairs"])
    assert abs((res.start_energy - lrc) - e0) <= e_tol(e0)
    hb2, _ = load_case(name, tmp_path)
    res2, tr2 = m.strategy_run(hb2, min(n, 300))
    k = tr2.mol.size
    assert np.array_equal(tr2.mol, ref["trace_mol"][:k])
    assert np.array_equal(tr2.accepted, ref["trace_acc"][:k])
    assert de_ok(tr2.de, ref, k)
