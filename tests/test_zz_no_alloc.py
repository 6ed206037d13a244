"""mcgpu_system_energy_noalloc through the C ABI: the same full energy as
mcgpu_system_energy, on two calls in a row (a 20,000-molecule box, then a
256-molecule box that leaves most of the shared array from the first call
untouched), in FP64 and with FP32 pair terms; boxes over the array's limit
are refused before any device work."""
import ctypes as C
from pathlib import Path

import pytest

import mcgpu_amd as m

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def noalloc(db):
    e = C.c_double()
    m.check(m.lib().mcgpu_system_energy_noalloc(db._b, C.byref(e)),
            "mcgpu_system_energy_noalloc")
    return e.value


@pytest.mark.parametrize("precision", [m.FP64, m.MIXED], ids=["fp64", "mixed"])
def test_noalloc_equals_system_energy_on_two_calls(precision):
    big = m.DeviceBox(m.workload("mixture20000"), precision=precision)
    big.neighbor(m.NEIGHBOR_CELL)
    small = m.DeviceBox(m.HostBox.from_state(str(GOLDEN / "water256.state")),
                        precision=precision)
    got_big = noalloc(big)
    got_small = noalloc(small)
    want_big, want_small = big.system_energy(), small.system_energy()
    # same kernels, same summation order
    assert got_big == want_big
    assert got_small == want_small
    assert got_small != got_big
    # and again, in the other order
    assert noalloc(small) == want_small
    assert noalloc(big) == want_big
    big.close()
    small.close()


def test_noalloc_refuses_boxes_beyond_its_limit():
    hb = m.HostBox.generate(m.WATER, 140000, density=0.997, cutoff=9.0, seed=3)
    db = m.DeviceBox(hb)
    with pytest.raises(m.McgpuError, match="scratch limit"):
        noalloc(db)
    db.close()
