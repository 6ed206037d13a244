"""CPU-side checks of the C ABI library: it loads, exports every declared
symbol, validates inputs, refuses to run without a GPU, and its host logic
(generators, .state I/O, LRC) is deterministic and matches the oracle."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import mcgpu_amd as m
from edge_boxes import build_box, kind
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
ERR_INVALID, ERR_NODEVICE = -1, -6   # MCGPU_ERR_INVALID, MCGPU_ERR_NODEVICE


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mcgpu_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = m.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("given,want", [(None, "16"), ("4", "16"), ("32", "32")])
def test_loading_the_library_raises_hardware_queues(given, want):
    """The package asks HIP for at least 16 hardware queues before the
    library starts HIP, and keeps a higher setting (fresh process each)."""
    import os
    import subprocess
    import sys
    env = dict(os.environ)
    env.pop("GPU_MAX_HW_QUEUES", None)
    if given is not None:
        env["GPU_MAX_HW_QUEUES"] = given
    code = ("import os, mcgpu_amd as m; m.lib(); "
            "print(os.environ['GPU_MAX_HW_QUEUES'])")
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd=ROOT,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split()[-1] == want


def test_no_device_here_fails_loudly():
    if m.device_count() > 0:
        pytest.skip("a GPU is visible")
    hb = m.HostBox.from_state(str(GOLDEN / "water256.state"))
    with pytest.raises(m.McgpuError, match="no HIP device"):
        m.DeviceBox(hb)


def test_invalid_descriptor_rejected_before_device():
    hb = m.HostBox.from_state(str(GOLDEN / "water256.state"))
    d = hb.desc()
    d.cutoff = -1.0
    b = C.c_void_p()
    assert m.lib().mcgpu_box_create(C.byref(d), 0, C.byref(b)) == -1
    assert b"cutoff" in m.lib().mcgpu_last_error()
    d = hb.desc()
    d.n_sites = d.n_sites - 1
    assert m.lib().mcgpu_box_create(C.byref(d), 0, C.byref(b)) == -1
    d = hb.desc()
    assert m.lib().mcgpu_box_create(C.byref(d), 7, C.byref(b)) == -1


def _edge_case(case):
    """A descriptor at or past one limit, built from plain arrays."""
    rng = np.random.default_rng(5)
    if case == "too_many_sites":        # 33 > MCGPU_MAX_MOL_SITES
        return build_box([kind(33, rng, (0,))], 8, 30.0, 10.0, 1), "molecule length"
    if case == "too_many_params":       # 5 x 32 distinct entries > 128
        kinds = [kind(32, rng, (0,)) for _ in range(5)]
        return build_box(kinds, 10, 40.0, 10.0, 1), "site-parameter"
    ab = build_box([kind(4, rng, (1,))], 27, 30.0, 10.0, 1)
    if case == "no_molecules":
        ab.a["mol_start"] = ab.a["mol_start"][:0]
        return ab, "n_molecules"
    if case == "primary_outside":
        ab.a["primary_index"][3] = 0    # molecule 3's primary in molecule 0
        return ab, "primary index outside"
    if case == "nan_coordinate":
        ab.a["y"][17] = np.nan
        return ab, "non-finite"
    if case == "ragged_ok":             # 1- to 32-site molecules: accepted
        return build_box([kind(1, rng, (0,)), kind(32, rng, (0, 31))], 16, 30.0, 10.0, 1), None
    raise ValueError(case)


@pytest.mark.parametrize("case", ["too_many_sites", "too_many_params", "no_molecules",
                                  "primary_outside", "nan_coordinate", "ragged_ok"])
def test_descriptor_limits(case):
    """Descriptors past a limit are refused before any device work, with a
    message naming the limit; one at the limits is accepted."""
    ab, why = _edge_case(case)
    d = ab.desc()
    b = C.c_void_p()
    rc = m.lib().mcgpu_box_create(C.byref(d), 0, C.byref(b))
    if why is None:
        nbytes = C.c_size_t()
        p = m.RunParams(1, 0, 0, 0, 100, 0.0, 0.0, 0.0, 0, 0)
        assert m.lib().mcgpu_box_footprint(C.byref(d), C.byref(p), C.byref(nbytes)) == 0
        assert nbytes.value > 0
        if m.device_count() == 0:
            assert rc == ERR_NODEVICE
        else:
            assert rc == 0
            m.lib().mcgpu_box_destroy(b)
        return
    assert rc == ERR_INVALID
    assert why in m.lib().mcgpu_last_error().decode()


def test_generators_are_deterministic_and_sized():
    a = m.workload("water256").arrays()
    b = m.workload("water256").arrays()
    for k in ("x", "y", "z", "charge"):
        assert np.array_equal(a[k], b[k])
    assert len(a["mol_start"]) == 256 and len(a["x"]) == 768
    assert a["box"][0] == 19.73 and a["cutoff"] == 9.0
    p = a["primary_index"][a["mol_pidx_start"]]
    assert np.all(a["x"][p] >= 0) and np.all(a["x"][p] < a["box"][0])
    c1 = m.workload("bicyclic4000").arrays()
    assert len(c1["x"]) == 64000
    assert c1["box"][0] == pytest.approx(89.10, abs=0.02)   # 1.10 g/cm3
    # neutral molecules, 9 heavy + 7 H
    q = c1["charge"].reshape(4000, 16)
    assert np.allclose(q.sum(1), 0, atol=1e-12)
    assert np.all(c1["sigma"].reshape(4000, 16)[:, 9:] == 2.42)


def test_state_roundtrip_is_bit_exact(tmp_path):
    hb = m.workload("water256", n_molecules=64, edge=12.44)
    p = tmp_path / "b.state"
    hb.write_state(str(p))
    hb2 = m.HostBox.from_state(str(p))
    a, b = hb.arrays(), hb2.arrays()
    for k in ("x", "y", "z", "sigma", "epsilon", "charge", "mol_start", "mol_len",
              "primary_index"):
        assert np.array_equal(a[k], b[k]), k
    assert hb2.mc_seed == hb.mc_seed and hb2.temperature == hb.temperature


def test_state_primary_definitions():
    hb = m.HostBox.from_state(str(GOLDEN / "water128_mp.state"))
    a = hb.arrays()
    assert np.all(a["mol_pidx_count"] == 2)
    assert np.array_equal(a["primary_index"][0::2], a["mol_start"])
    assert np.array_equal(a["primary_index"][1::2], a["mol_start"] + 1)


@pytest.mark.parametrize("name", ["water256", "mixture300", "bicyclic200"])
def test_product_lrc_matches_oracle(name):
    hb = m.HostBox.from_state(str(GOLDEN / f"{name}.state"))
    d = hb.desc()
    e = C.c_double()
    m.check(m.lib().mcgpu_lrc_reference(C.byref(d), C.byref(e)))
    ob = O.OracleBox(hb.arrays())
    assert e.value == ob.lrc_reference()
    m.check(m.lib().mcgpu_lrc_species(C.byref(d), C.byref(e)))
    assert e.value == pytest.approx(ob.lrc_species(), rel=1e-13)
