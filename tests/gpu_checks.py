"""Checks shared by the GPU parity tests: the dE tolerance, running a batch
to completion, and an NPT run compared move by move with the oracle."""
import numpy as np
import pytest

import mcgpu_amd as m
from oracle import oracle as O

DE_TOL = 1e-9


def de_ok(got, want, scale=None):
    """1e-9 kcal/mol plus 1e-12 of the move's energy scale: max(|E_old|,
    |E_new|) when the reference trace carries them (dE = E_new - E_old, each
    a sum of pair terms in another order than the reference's), else |dE|.
    The batched path also builds trial moves on the device, whose sin/cos
    may differ from glibc's in the last bit.  The relative term matters only
    at trial positions overlapping another molecule (energies up to 1e12
    kcal/mol in the generated 16-site boxes); on the 256-water box every
    move meets the absolute 1e-9 bound.  Moves of physical size (energy
    scale below 1e3 kcal/mol) must also meet 2e-10 absolutely, whatever the
    box (the largest error measured over every GPU test's moves of that
    size is 3.8e-11)."""
    if isinstance(want, (dict, np.lib.npyio.NpzFile)):
        scale = np.maximum(np.abs(want["old_e"]), np.abs(want["new_e"]))
        want = want["de"]
    elif scale is None:
        scale = np.abs(want)
    err = np.abs(got - want)
    phys = scale < 1e3
    phys_max = err[phys].max() if phys.any() else 0.0
    print(f"max |dE err| {err.max():.3e}, max err/scale "
          f"{np.max(err / np.maximum(scale, 1e-300)):.3e}, max on moves of "
          f"physical size {phys_max:.3e}")
    return bool(np.all(err <= DE_TOL + 1e-12 * scale) and phys_max <= 2e-10)


def run_all(batch, slots, total, chunk):
    """Runs until every box is done (volume moves pause a box between
    launches, so more launches than total/chunk may be needed)."""
    for _ in range(100 * (total // chunk + 1)):
        stats = [batch.stats(s) for s in slots]
        if all(st.done for st in stats):
            return stats
        batch.run(chunk)
    raise AssertionError("boxes did not finish")


P_ATM = 1.4584e-5  # 1 atm in kcal/mol/A^3


def npt_against_oracle(hb, steps, vi, delta, neighbor, precision=m.FP64, seed=None,
                       chunk=None, e_rel=1e-9):
    """Runs hb under NPT on the device (record_trace) and the oracle's NPT
    loop with the same seed: every move and every volume move must agree
    (npt_matches_oracle).  Returns the device's volume-move records."""
    a = hb.arrays()
    v0 = float(a["box"][0] * a["box"][1] * a["box"][2])
    seed = hb.mc_seed if seed is None else seed
    b = m.Batch(2)
    s = b.add(hb, steps=steps, seed=seed, ensemble=m.NPT, volume_interval=vi,
              pressure=P_ATM, max_volume_delta=delta, neighbor=neighbor,
              precision=precision, record_trace=True)
    run_all(b, [s], steps, chunk or vi)
    vg = npt_matches_oracle(b, s, hb, steps, vi, delta * v0, seed, precision, e_rel)
    b.close()
    return vg


def npt_matches_oracle(b, s, hb, steps, vi, dv, seed, precision=m.FP64, e_rel=1e-9):
    """Checks a finished NPT box of batch b (added with record_trace) against
    the oracle's NPT loop with the same seed and volume step dv: moves,
    volume moves, their energies, and the final running energy.  Each device
    volume decision is also replayed through the oracle's acceptance test
    (SimCalcs::acceptVolumeMove) on the device's own V, V', E, E' and u.
    Returns the device's volume-move records."""
    a = hb.arrays()
    ob = O.OracleBox(a)
    fin, e0, nacc, t, v = ob.run_npt(seed, steps, vi, P_ATM, dv, use_libc=False)
    st = b.stats(s)
    tr, vg = b.trace(s), b.volume_trace(s)
    assert st.steps_done == steps and st.done and st.fault == 0
    assert st.start_energy == pytest.approx(e0, rel=1e-11 if precision == m.FP64 else e_rel)
    assert np.array_equal(tr.mol, t["mol"])
    assert np.array_equal(tr.accepted, t["accepted"])
    if precision == m.FP64:
        assert de_ok(tr.de, t)
    else:
        phys = np.abs(t["de"]) < 1e3
        assert np.max(np.abs(tr.de - t["de"])[phys]) <= 1e-4
    n_vol = (steps - 1) // vi
    assert st.volume_moves == n_vol == len(v["step"]) == len(vg["step"])
    assert st.volume_accepted == int(v["accepted"].sum())
    assert np.array_equal(vg["step"], v["step"])
    assert np.array_equal(vg["accepted"], v["accepted"])
    assert np.array_equal(vg["u"], v["u"])
    assert np.allclose(vg["v_old"], v["v_old"], rtol=1e-12, atol=0)
    assert np.allclose(vg["v_new"], v["v_new"], rtol=1e-12, atol=0)
    for k in ("e_old", "e_new"):
        err = np.abs(vg[k] - v[k]) / np.maximum(np.abs(v[k]), 1.0)
        print(f"volume moves: max rel err {k} {err.max():.3e}")
        assert err.max() <= e_rel, k
    n, kT = len(a["mol_start"]), float(a["kT"])
    lib = O.lib()
    for k in range(n_vol):
        prob = lib.orc_volume_accept_prob(n, vg["v_old"][k], vg["v_new"][k],
                                          vg["e_new"][k] - vg["e_old"][k], P_ATM, kT)
        assert int(prob >= vg["u"][k]) == vg["accepted"][k], k
    # the running energy is a fresh evaluation of the final box (K-species LRC)
    x, y, z = b.download(s)
    fa = hb.arrays()
    fa.update(x=x, y=y, z=z, box=np.array(st.box[:]))
    fresh = O.OracleBox(fa)
    e = fresh.system_energy() + fresh.lrc_species()
    assert st.energy == pytest.approx(e, rel=1e-9 if precision == m.FP64 else 1e-7, abs=1e-6)
    assert st.energy == pytest.approx(fin, rel=1e-9 if precision == m.FP64 else 1e-7, abs=1e-6)
    return vg
