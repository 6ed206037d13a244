"""MI355X-native MCGPU energy hot path.

The compute lives in ``libmcgpu.so`` (hand-written gfx950 HIP kernels behind
the C ABI declared in ``include/mcgpu_energy.h`` and ``include/mcgpu_host.h``).
This package only loads that library and describes its structures for
ctypes; there is no Python or CPU fallback.  Build with
``python -m mcgpu_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libmcgpu.so"
_lib = None

# Hardware queues for the process's HIP streams.  A batch keeps up to seven
# streams busy at once (a tick's three work-group-size groups; beside them
# the volume moves and the groups that move after them; admissions).  HIP
# maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and
# work on streams that share a queue runs in order, so a volume pass would
# wait behind a tick.  Raised to this many unless set higher; HIP reads it
# when its runtime starts, so it takes effect only before the process's
# first HIP call.
MIN_HW_QUEUES = 16

MCGPU_OK = 0
FP64 = 0
MIXED = 1
NEIGHBOR_BRUTE = 0
NEIGHBOR_CELL = 1
NEIGHBOR_VERLET = 2
NVT = 0
NPT = 1
JOB_QUEUED = 0
JOB_RUNNING = 1
JOB_DONE = 2
WL_WATER = 0
WL_BICYCLIC = 1
WL_MIXTURE = 2
MAX_MOL_SITES = 32

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)


class BoxDesc(C.Structure):
    _fields_ = [
        ("n_molecules", C.c_int32), ("n_sites", C.c_int32),
        ("n_primary", C.c_int32), ("reserved", C.c_int32),
        ("box", C.c_double * 3), ("cutoff", C.c_double), ("kT", C.c_double),
        ("max_translate", C.c_double), ("max_rotate", C.c_double),
        ("x", _d), ("y", _d), ("z", _d),
        ("sigma", _d), ("epsilon", _d), ("charge", _d),
        ("mol_start", _i32), ("mol_len", _i32), ("mol_pidx_start", _i32),
        ("mol_pidx_count", _i32), ("primary_index", _i32),
    ]


class MoveResult(C.Structure):
    _fields_ = [("old_energy", C.c_double), ("new_energy", C.c_double),
                ("delta", C.c_double), ("accepted", C.c_int32),
                ("reserved", C.c_int32)]


class RunParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint32), ("precision", C.c_int32), ("ensemble", C.c_int32),
        ("volume_interval", C.c_int32), ("steps", C.c_int64),
        ("pressure", C.c_double), ("max_volume_delta", C.c_double),
        ("energy_offset", C.c_double), ("neighbor", C.c_int32),
        ("record_trace", C.c_int32),
    ]


class BoxStats(C.Structure):
    _fields_ = [
        ("steps_done", C.c_int64), ("accepted", C.c_int64),
        ("volume_moves", C.c_int64), ("volume_accepted", C.c_int64),
        ("energy", C.c_double), ("start_energy", C.c_double),
        ("box", C.c_double * 3), ("work_candidates", C.c_int64),
        ("work_sites", C.c_int64), ("done", C.c_int32), ("fault", C.c_int32),
        ("neighbor_fallbacks", C.c_int64),
    ]


class VolumeRecord(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("v_old", C.c_double), ("v_new", C.c_double),
        ("e_old", C.c_double), ("e_new", C.c_double), ("u", C.c_double),
        ("accepted", C.c_int32), ("reserved", C.c_int32),
    ]


class ResumeState(C.Structure):
    _fields_ = [
        ("rng_words", C.c_int32 * 31), ("rng_front", C.c_int32),
        ("rng_rear", C.c_int32), ("volume_pending", C.c_int32),
        ("steps_done", C.c_int64), ("accepted", C.c_int64),
        ("volume_moves", C.c_int64), ("volume_accepted", C.c_int64),
        ("energy", C.c_double), ("lrc", C.c_double), ("volume_step", C.c_double),
        ("ensemble", C.c_int32), ("volume_interval", C.c_int32), ("pressure", C.c_double),
    ]


class WorkloadParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_molecules", C.c_int32), ("n_second", C.c_int32),
        ("lattice_jitter_permille", C.c_int32), ("density", C.c_double),
        ("edge", C.c_double), ("temperature", C.c_double), ("cutoff", C.c_double),
        ("max_translate", C.c_double), ("max_rotate", C.c_double),
        ("seed", C.c_uint64), ("mc_seed", C.c_uint32), ("reserved", C.c_int32),
    ]


class PlainResult(C.Structure):
    _fields_ = [
        ("start_energy", C.c_double), ("lrc", C.c_double),
        ("final_energy", C.c_double), ("accepted", C.c_int64),
        ("rejected", C.c_int64), ("loop_seconds", C.c_double),
        ("startup_seconds", C.c_double), ("neighbor_builds", C.c_int64),
        ("neighbor_fallbacks", C.c_int64),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("max_boxes", C.c_int32), ("sweeps_per_tick", C.c_int32),
                ("hbm_budget_bytes", C.c_size_t), ("tick_budget_us", C.c_double)]


class JobResult(C.Structure):
    _fields_ = [
        ("job_id", C.c_int64), ("arrival_s", C.c_double), ("admit_s", C.c_double),
        ("finish_s", C.c_double), ("final_energy", C.c_double),
        ("start_energy", C.c_double), ("steps", C.c_int64),
        ("accepted", C.c_int64), ("status", C.c_int32), ("state", C.c_int32),
        ("steps_done", C.c_int64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double), ("total_moves", C.c_int64), ("ticks", C.c_int64),
        ("kernel_ms_total", C.c_double), ("peak_resident", C.c_int64),
        ("peak_bytes", C.c_size_t), ("work_candidates", C.c_int64),
        ("work_sites", C.c_int64), ("resident", C.c_int32), ("reserved", C.c_int32),
        ("queued", C.c_int64), ("completed", C.c_int64), ("submitted", C.c_int64),
        ("clock_s", C.c_double),
    ]


class ConfigInfo(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("n_molecules", C.c_int32), ("n_types", C.c_int32),
        ("random_seed", C.c_int32), ("volume_interval", C.c_int32),
        ("ewald_radius", C.c_int32), ("reserved", C.c_int32),
        ("pressure", C.c_double), ("sim_name", C.c_char * 256),
        ("strategy", C.c_char * 64), ("ensemble", C.c_char * 64),
        ("state_output", C.c_char * 1024), ("pdb_output", C.c_char * 1024),
    ]


class McgpuError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Loads libmcgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise McgpuError(
                f"{_LIB_PATH} is missing: build it with `python -m mcgpu_amd.build`")
        try:
            queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
        except ValueError:
            queues = 0
        if queues < MIN_HW_QUEUES:
            os.environ["GPU_MAX_HW_QUEUES"] = str(MIN_HW_QUEUES)
        _lib = C.CDLL(str(_LIB_PATH))
        _lib.mcgpu_last_error.restype = C.c_char_p
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != MCGPU_OK:
        msg = lib().mcgpu_last_error().decode(errors="replace")
        raise McgpuError(f"{what or 'mcgpu call'} failed ({rc}): {msg}")


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mcgpu_device_count(C.byref(n)), "mcgpu_device_count")
    return n.value


from .api import *  # noqa: E402,F401,F403
