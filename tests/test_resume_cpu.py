"""The .state resume record on the CPU: written and read back bit-exactly,
absent from files without one, and skipped by the reference's own .state
reader (built from its sources into oracle/_ref/ref_load), which loads a
file with the record exactly as the same file without it."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_LOAD = Path(O.__file__).resolve().parent / "_ref" / "ref_load"


def sample_record():
    r = m.ResumeState()
    g = O.GlibcRng(12345)
    for _ in range(1000):
        g.rand()
    for k in range(31):
        r.rng_words[k] = g.s.state[k]
    r.rng_front, r.rng_rear = g.s.f, g.s.r
    r.volume_pending = 1
    r.steps_done, r.accepted, r.volume_moves, r.volume_accepted = 3000, 1234, 5, 3
    r.energy = -2345.678901234567891
    r.lrc = -12.3456789012345678
    r.volume_step = 130.77276476090854
    r.ensemble, r.volume_interval, r.pressure = m.NPT, 500, 1.4584e-5
    return r


def as_tuple(r):
    return (tuple(r.rng_words), r.rng_front, r.rng_rear, r.volume_pending, r.steps_done,
            r.accepted, r.volume_moves, r.volume_accepted, r.energy, r.lrc, r.volume_step,
            r.ensemble, r.volume_interval, r.pressure)


def test_resume_record_roundtrip(tmp_path):
    hb = m.HostBox.from_state(str(GOLDEN / "mixture300.state"))
    assert hb.resume is None
    r = sample_record()
    hb.resume = r
    p = tmp_path / "r.state"
    hb.write_state(str(p), step=3000)
    back = m.HostBox.from_state(str(p))
    assert back.resume is not None
    assert as_tuple(back.resume) == as_tuple(r)
    a, b = hb.arrays(), back.arrays()
    for k in ("x", "y", "z", "sigma", "epsilon", "charge"):
        assert np.array_equal(a[k], b[k]), k
    # cleared again: no record written
    back.resume = None
    back.write_state(str(tmp_path / "n.state"))
    assert m.HostBox.from_state(str(tmp_path / "n.state")).resume is None
    assert "#mcgpu-resume" not in (tmp_path / "n.state").read_text()


def test_reference_reader_skips_the_resume_record(tmp_path):
    if not REF_LOAD.exists():
        pytest.skip("reference build (oracle/_ref/ref_load) absent")
    hb = m.HostBox.from_state(str(GOLDEN / "mixture300.state"))
    plain, rec = tmp_path / "plain.state", tmp_path / "rec.state"
    hb.write_state(str(plain), step=3000)
    hb.resume = sample_record()
    hb.write_state(str(rec), step=3000)
    assert rec.read_text().startswith(plain.read_text())
    dumps = []
    for p in (plain, rec):
        out = tmp_path / (p.stem + ".dump")
        subprocess.run([str(REF_LOAD), str(p), str(out)], check=True, capture_output=True)
        dumps.append(out.read_text())
    assert dumps[0] == dumps[1]
    assert dumps[0].count("\nmol ") == 300
