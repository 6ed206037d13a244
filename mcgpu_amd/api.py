"""Thin ctypes wrappers over the C ABI (no computation happens here)."""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np

from . import (BoxDesc, BoxStats, ConfigInfo, JobResult, MoveResult, PlainResult,
               ResumeState, RunParams, ServerConfig, ServerStats, VolumeRecord,
               WorkloadParams, check, lib)

__all__ = [
    "HostBox", "DeviceBox", "Batch", "Server", "plain_run", "strategy_run",
    "workload", "lrc_for_run", "WATER", "BICYCLIC", "MIXTURE",
]

WATER, BICYCLIC, MIXTURE = 0, 1, 2


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class HostBox:
    """A host-resident box owned by the library (mcgpu_hostbox)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    # -- construction
    @classmethod
    def generate(cls, kind: int, n_molecules: int, *, density: float = 0.0,
                 edge: float = 0.0, temperature: float = 298.0,
                 cutoff: float = 9.0, max_translate: float = 0.15,
                 max_rotate: float = 15.0, seed: int = 1, mc_seed: int = 1,
                 n_second: int = 0, jitter_permille: int = 100) -> "HostBox":
        p = WorkloadParams(kind, n_molecules, n_second, jitter_permille,
                           density, edge, temperature, cutoff, max_translate,
                           max_rotate, seed, mc_seed, 0)
        h = C.c_void_p()
        check(lib().mcgpu_workload_generate(C.byref(p), C.byref(h)),
              "mcgpu_workload_generate")
        return cls(h)

    @classmethod
    def from_state(cls, path: str) -> "HostBox":
        h = C.c_void_p()
        step = C.c_int64()
        check(lib().mcgpu_state_read(str(path).encode(), C.byref(h),
                                     C.byref(step)), "mcgpu_state_read")
        return cls(h)

    @classmethod
    def from_config(cls, path: str):
        """Loads a metrosim .config run (z-matrix + OPLS .par input); returns
        (box, settings) where settings holds the run's other keys."""
        h = C.c_void_p()
        info = ConfigInfo()
        check(lib().mcgpu_config_read(str(path).encode(), C.byref(h),
                                      C.byref(info)), "mcgpu_config_read")
        settings = dict(
            steps=info.steps, n_molecules=info.n_molecules, n_types=info.n_types,
            random_seed=info.random_seed, volume_interval=info.volume_interval,
            ewald_radius=info.ewald_radius, pressure=info.pressure,
            sim_name=info.sim_name.decode(), strategy=info.strategy.decode(),
            ensemble=info.ensemble.decode(),
            state_output=info.state_output.decode(),
            pdb_output=info.pdb_output.decode())
        return cls(h), settings

    def write_state(self, path: str, step: int = 0) -> None:
        check(lib().mcgpu_state_write(self._h, str(path).encode(),
                                      C.c_int64(step)), "mcgpu_state_write")

    @property
    def resume(self):
        """The box's resume record (ResumeState) or None."""
        r = ResumeState()
        present = C.c_int32()
        check(lib().mcgpu_hostbox_get_resume(self._h, C.byref(r), C.byref(present)),
              "mcgpu_hostbox_get_resume")
        return r if present.value else None

    @resume.setter
    def resume(self, r) -> None:
        check(lib().mcgpu_hostbox_set_resume(self._h, C.byref(r) if r is not None else None),
              "mcgpu_hostbox_set_resume")

    def set_coords(self, x, y, z, edges=None) -> None:
        """Overwrites the box's coordinates (and edges)."""
        px, py, pz = (C.POINTER(C.c_double)(), C.POINTER(C.c_double)(),
                      C.POINTER(C.c_double)())
        check(lib().mcgpu_hostbox_coords(self._h, C.byref(px), C.byref(py), C.byref(pz)))
        n = self.desc().n_sites
        for dst, src in ((px, x), (py, y), (pz, z)):
            np.ctypeslib.as_array(dst, shape=(n,))[:] = np.asarray(src, np.float64)
        if edges is not None:
            e = (C.c_double * 3)(*[float(v) for v in edges])
            check(lib().mcgpu_hostbox_set_edges(self._h, e))

    def free(self) -> None:
        if self._h:
            lib().mcgpu_hostbox_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- views
    @property
    def handle(self):
        return self._h

    def desc(self) -> BoxDesc:
        d = BoxDesc()
        check(lib().mcgpu_hostbox_desc(self._h, C.byref(d)), "mcgpu_hostbox_desc")
        return d

    @property
    def temperature(self) -> float:
        t = C.c_double()
        s = C.c_uint32()
        check(lib().mcgpu_hostbox_info(self._h, C.byref(t), C.byref(s)))
        return t.value

    @property
    def mc_seed(self) -> int:
        t = C.c_double()
        s = C.c_uint32()
        check(lib().mcgpu_hostbox_info(self._h, C.byref(t), C.byref(s)))
        return s.value

    def arrays(self, copy: bool = True) -> dict:
        """The descriptor arrays: copies by default; with copy=False, numpy
        views that are valid only while this HostBox is alive."""
        d = self.desc()
        ns, nm, npr = d.n_sites, d.n_molecules, d.n_primary
        if copy:
            v = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
        else:
            v = lambda p, n: np.ctypeslib.as_array(p, shape=(n,))
        return dict(
            x=v(d.x, ns), y=v(d.y, ns), z=v(d.z, ns), sigma=v(d.sigma, ns),
            epsilon=v(d.epsilon, ns), charge=v(d.charge, ns),
            mol_start=v(d.mol_start, nm), mol_len=v(d.mol_len, nm),
            mol_pidx_start=v(d.mol_pidx_start, nm),
            mol_pidx_count=v(d.mol_pidx_count, nm),
            primary_index=v(d.primary_index, npr),
            box=np.array(d.box[:]), cutoff=d.cutoff, kT=d.kT,
            max_translate=d.max_translate, max_rotate=d.max_rotate)


def workload(name: str, **kw) -> HostBox:
    """The benchmark/parity inputs by name (sizes from the workload table)."""
    if name == "water256":  # 256 TIP3P, L = 19.73 A, rc 9, 298 K
        args = dict(kind=WATER, n_molecules=256, edge=19.73, cutoff=9.0,
                    temperature=298.0, max_translate=0.15, max_rotate=15.0,
                    seed=20240601, mc_seed=1)
    elif name == "bicyclic4000":  # 4000 x 16 sites, 1.10 g/cm3, rc 12
        args = dict(kind=BICYCLIC, n_molecules=4000, density=1.10, cutoff=12.0,
                    temperature=298.0, max_translate=0.15, max_rotate=15.0,
                    seed=20240602, mc_seed=1)
    elif name == "mixture20000":  # 12000 water + 8000 methanol, rc 11
        args = dict(kind=MIXTURE, n_molecules=20000, n_second=8000, edge=96.45,
                    cutoff=11.0, temperature=298.0, max_translate=0.15,
                    max_rotate=15.0, seed=20240603, mc_seed=1)
    else:
        raise ValueError(f"unknown workload {name}")
    args.update(kw)
    return HostBox.generate(**args)


class DeviceBox:
    """A device-resident box (mcgpu_box)."""

    def __init__(self, host: HostBox | BoxDesc, precision: int = 0):
        d = host.desc() if isinstance(host, HostBox) else host
        self._keep = host
        self._b = C.c_void_p()
        check(lib().mcgpu_box_create(C.byref(d), int(precision), C.byref(self._b)),
              "mcgpu_box_create")
        self.n_mol = d.n_molecules
        self.n_site = d.n_sites
        self._len = np.ctypeslib.as_array(d.mol_len, shape=(d.n_molecules,)).copy()

    def close(self) -> None:
        if self._b:
            lib().mcgpu_box_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def neighbor(self, kind: int, skin: float = 1.0) -> None:
        check(lib().mcgpu_neighbor_rebuild(self._b, int(kind), C.c_double(skin)),
              "mcgpu_neighbor_rebuild")

    def set_precision(self, p: int) -> None:
        check(lib().mcgpu_box_set_precision(self._b, int(p)))

    def molecule_energy(self, mol: int, start: int = 0) -> float:
        e = C.c_double()
        check(lib().mcgpu_molecule_energy(self._b, int(mol), int(start), C.byref(e)),
              "mcgpu_molecule_energy")
        return e.value

    def system_energy(self) -> float:
        e = C.c_double()
        check(lib().mcgpu_system_energy(self._b, C.byref(e)), "mcgpu_system_energy")
        return e.value

    def system_pairs(self) -> int:
        n = C.c_int64()
        check(lib().mcgpu_system_pair_count(self._b, C.byref(n)))
        return n.value

    def trial_move(self, mol: int, new_xyz: np.ndarray, u: float) -> MoveResult:
        a = np.ascontiguousarray(new_xyz, dtype=np.float64).reshape(-1)
        r = MoveResult()
        check(lib().mcgpu_trial_move(self._b, int(mol), _dptr(a), C.c_double(u),
                                     C.byref(r)), "mcgpu_trial_move")
        return r

    def set_molecule(self, mol: int, xyz: np.ndarray) -> None:
        a = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        check(lib().mcgpu_set_molecule_coords(self._b, int(mol), _dptr(a)),
              "mcgpu_set_molecule_coords")

    def get_molecule(self, mol: int) -> np.ndarray:
        out = np.zeros(3 * int(self._len[mol]), dtype=np.float64)
        check(lib().mcgpu_get_molecule_coords(self._b, int(mol), _dptr(out)))
        return out

    def download(self):
        x = np.zeros(self.n_site)
        y = np.zeros(self.n_site)
        z = np.zeros(self.n_site)
        check(lib().mcgpu_download_coords(self._b, _dptr(x), _dptr(y), _dptr(z)))
        return x, y, z

    def upload(self, x, y, z) -> None:
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        check(lib().mcgpu_upload_coords(self._b, _dptr(x), _dptr(y), _dptr(z)))

    def volume_scale(self, f: float) -> None:
        check(lib().mcgpu_volume_scale(self._b, C.c_double(f)), "mcgpu_volume_scale")

    def size(self) -> np.ndarray:
        e = np.zeros(3)
        check(lib().mcgpu_box_size(self._b, _dptr(e)))
        return e

    def enable_timing(self, on: bool = True) -> None:
        check(lib().mcgpu_box_enable_timing(self._b, int(on)))

    def neighbor_stats(self):
        """(Verlet list builds, evaluations that fell back to scanning every
        molecule) since the box was created."""
        b, f = C.c_int64(), C.c_int64()
        check(lib().mcgpu_box_neighbor_stats(self._b, C.byref(b), C.byref(f)),
              "mcgpu_box_neighbor_stats")
        return b.value, f.value

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        check(lib().mcgpu_last_kernel_ms(self._b, C.byref(v)))
        return v.value


@dataclass
class RunTrace:
    mol: np.ndarray
    de: np.ndarray
    accepted: np.ndarray


def _run(fn, args, steps: int, trace: bool):
    res = PlainResult()
    if trace:
        tm = np.zeros(steps, np.int32)
        td = np.zeros(steps, np.float64)
        ta = np.zeros(steps, np.int8)
        ptrs = (_iptr(tm), _dptr(td), ta.ctypes.data_as(C.POINTER(C.c_int8)))
    else:
        ptrs = (None, None, None)
    check(fn(*args, C.c_int64(steps), *ptrs, C.byref(res)), fn.__name__)
    return res, (RunTrace(tm, td, ta) if trace else None)


def plain_run(box: HostBox, steps: int, precision: int = 0, neighbor: int = 0,
              trace: bool = True):
    """Plain path: host RNG, one fused trial-move launch per move."""
    return _run(lib().mcgpu_plain_run, (box.handle, int(precision), int(neighbor)),
                steps, trace)


def strategy_run(box: HostBox, steps: int, trace: bool = True):
    """Reference-shaped strategy calls (2 energy calls + move/rollback)."""
    return _run(lib().mcgpu_strategy_run, (box.handle,), steps, trace)


def lrc_for_run(d: BoxDesc, ensemble: int = 0) -> float:
    """The LRC a run starts from: the reference's formula under NVT, the
    K-species form under NPT (mcgpu_lrc_for_run)."""
    e = C.c_double()
    check(lib().mcgpu_lrc_for_run(C.byref(d), C.c_int32(int(ensemble)), C.byref(e)),
          "mcgpu_lrc_for_run")
    return e.value


class Batch:
    """Many independent boxes advanced by one persistent launch."""

    def __init__(self, max_boxes: int, hbm_budget: int = 0):
        self._b = C.c_void_p()
        check(lib().mcgpu_batch_create(int(max_boxes), C.c_size_t(hbm_budget),
                                       C.byref(self._b)), "mcgpu_batch_create")
        self._keep = {}

    def close(self):
        if self._b:
            lib().mcgpu_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, box: HostBox, *, steps: int, seed: int | None = None,
            precision: int = 0, ensemble: int = 0, volume_interval: int = 0,
            pressure: float = 0.0, max_volume_delta: float = 0.0,
            energy_offset: float | None = None, neighbor: int = 0,
            record_trace: bool = False, resume=None) -> int:
        """Adds a box; with resume (a ResumeState) it continues that record's
        move sequence (mcgpu_batch_add_resumed)."""
        d = box.desc()
        if energy_offset is None:
            energy_offset = lrc_for_run(d, ensemble)
        p = RunParams(box.mc_seed if seed is None else seed, precision, ensemble,
                      volume_interval, steps, pressure, max_volume_delta,
                      energy_offset, neighbor, int(record_trace))
        slot = C.c_int32()
        if resume is not None:
            check(lib().mcgpu_batch_add_resumed(self._b, C.byref(d), C.byref(p),
                                                C.byref(resume), C.byref(slot)),
                  "mcgpu_batch_add_resumed")
        else:
            check(lib().mcgpu_batch_add(self._b, C.byref(d), C.byref(p), C.byref(slot)),
                  "mcgpu_batch_add")
        self._keep[slot.value] = (box, d.n_sites, steps)
        return slot.value

    def remove(self, slot: int) -> None:
        check(lib().mcgpu_batch_remove(self._b, int(slot)), "mcgpu_batch_remove")
        self._keep.pop(slot, None)

    def run(self, sweeps: int, sync: bool = True) -> None:
        check(lib().mcgpu_batch_run(self._b, int(sweeps), int(sync)), "mcgpu_batch_run")

    def stats(self, slot: int) -> BoxStats:
        s = BoxStats()
        check(lib().mcgpu_batch_stats(self._b, int(slot), C.byref(s)), "mcgpu_batch_stats")
        return s

    def peek(self, slot: int) -> BoxStats:
        """Statistics as of the last tick consumed, without waiting for a
        tick in flight (run(sync=False) then peek overlaps host work with
        the tick)."""
        s = BoxStats()
        check(lib().mcgpu_batch_peek(self._b, int(slot), C.byref(s)), "mcgpu_batch_peek")
        return s

    def download(self, slot: int):
        n = self._keep[slot][1]
        x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().mcgpu_batch_download(self._b, int(slot), _dptr(x), _dptr(y),
                                         _dptr(z)), "mcgpu_batch_download")
        return x, y, z

    def trace(self, slot: int) -> RunTrace:
        cap = self._keep[slot][2]
        tm = np.zeros(cap, np.int32)
        td = np.zeros(cap, np.float64)
        ta = np.zeros(cap, np.int8)
        n = C.c_int64()
        check(lib().mcgpu_batch_trace(self._b, int(slot), _iptr(tm), _dptr(td),
                                      ta.ctypes.data_as(C.POINTER(C.c_int8)),
                                      C.c_int64(cap), C.byref(n)), "mcgpu_batch_trace")
        k = n.value
        return RunTrace(tm[:k], td[:k], ta[:k])

    def get_resume(self, slot: int) -> ResumeState:
        r = ResumeState()
        check(lib().mcgpu_batch_get_resume(self._b, int(slot), C.byref(r)),
              "mcgpu_batch_get_resume")
        return r

    def volume_trace(self, slot: int) -> dict:
        """The NPT volume moves of a record_trace box, as arrays: step,
        v_old, v_new, e_old, e_new, u, accepted."""
        steps = self._keep[slot][2]
        cap = steps + 1
        n = C.c_int64()
        check(lib().mcgpu_batch_volume_trace(self._b, int(slot), None, C.c_int64(0),
                                             C.byref(n)), "mcgpu_batch_volume_trace")
        out = (VolumeRecord * max(cap, 1))()
        check(lib().mcgpu_batch_volume_trace(self._b, int(slot), out, C.c_int64(cap),
                                             C.byref(n)), "mcgpu_batch_volume_trace")
        recs = out[:n.value]
        cols = {k: np.array([getattr(r, k) for r in recs], dt) for k, dt in
                (("step", np.int64), ("v_old", np.float64), ("v_new", np.float64),
                 ("e_old", np.float64), ("e_new", np.float64), ("u", np.float64),
                 ("accepted", np.int8))}
        return cols

    def set_tick_budget(self, microseconds: float) -> None:
        check(lib().mcgpu_batch_set_tick_budget(self._b, C.c_double(microseconds)))

    def set_autotune(self, on: bool) -> None:
        check(lib().mcgpu_batch_set_autotune(self._b, int(bool(on))))

    def tuning(self):
        """Tuned box classes: [(key, lanes chosen, (moves/tick at 256, 512, 1024))]."""
        n = C.c_int32()
        check(lib().mcgpu_batch_tuning(self._b, None, None, None, 0, C.byref(n)))
        k = n.value
        keys = (C.c_int64 * max(k, 1))()
        thr = (C.c_int32 * max(k, 1))()
        mv = (C.c_double * max(3 * k, 1))()
        check(lib().mcgpu_batch_tuning(self._b, keys, thr, mv, k, C.byref(n)))
        return [(keys[i], thr[i], tuple(mv[3 * i:3 * i + 3])) for i in range(k)]

    def group_threads(self, slot: int) -> int:
        t = C.c_int32()
        check(lib().mcgpu_batch_group(self._b, int(slot), C.byref(t)))
        return t.value

    def kernel_ms(self):
        ms = C.c_double()
        n = C.c_int64()
        check(lib().mcgpu_batch_kernel_ms(self._b, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Server:
    """Continuous-batching request loop over device-resident boxes."""

    def __init__(self, max_boxes: int, sweeps_per_tick: int, hbm_budget: int,
                 tick_budget_us: float = 0.0):
        cfg = ServerConfig(max_boxes, sweeps_per_tick, hbm_budget, tick_budget_us)
        self._s = C.c_void_p()
        check(lib().mcgpu_server_create(C.byref(cfg), C.byref(self._s)),
              "mcgpu_server_create")
        self._boxes = []
        self._lock = threading.Lock()

    def close(self):
        if self._s:
            lib().mcgpu_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, box: HostBox, arrival_s: float = -1.0, *, steps: int,
               precision: int = 0, ensemble: int = 0, volume_interval: int = 0,
               pressure: float = 0.0, max_volume_delta: float = 0.0,
               neighbor: int = 0, seed: int | None = None) -> int:
        """Queues a job (any thread, before or while the loop runs);
        arrival_s < 0 means now on the server clock."""
        d = box.desc()
        p = RunParams(box.mc_seed if seed is None else seed, precision, ensemble,
                      volume_interval, steps, pressure, max_volume_delta,
                      lrc_for_run(d, ensemble), neighbor, 0)
        jid = C.c_int64()
        with self._lock:
            self._boxes.append(box)   # the server reads it at admission
        check(lib().mcgpu_server_submit(self._s, box.handle, C.byref(p),
                                        C.c_double(arrival_s), C.byref(jid)),
              "mcgpu_server_submit")
        return jid.value

    @property
    def n(self) -> int:
        return int(self.stats().submitted)

    def start(self) -> None:
        """Starts the loop on the server's own thread; returns at once."""
        check(lib().mcgpu_server_start(self._s), "mcgpu_server_start")

    def wait(self, job_id: int = -1, timeout_s: float = -1.0) -> None:
        """Blocks until the job (every submitted job for -1) is done."""
        check(lib().mcgpu_server_wait(self._s, C.c_int64(job_id), C.c_double(timeout_s)),
              "mcgpu_server_wait")

    def stop(self) -> None:
        """Refuses new jobs, finishes the submitted ones, joins the loop."""
        check(lib().mcgpu_server_stop(self._s), "mcgpu_server_stop")

    def poll(self, job_id: int) -> JobResult:
        r = JobResult()
        check(lib().mcgpu_server_poll(self._s, C.c_int64(job_id), C.byref(r)),
              "mcgpu_server_poll")
        return r

    def run(self) -> None:
        """Start, wait for every submitted job, stop."""
        check(lib().mcgpu_server_run(self._s), "mcgpu_server_run")

    def results(self) -> list:
        n_sub = self.n
        out = (JobResult * max(n_sub, 1))()
        n = C.c_int64()
        check(lib().mcgpu_server_results(self._s, out, C.c_int64(n_sub), C.byref(n)))
        return [out[i] for i in range(n.value)]

    def stats(self) -> ServerStats:
        s = ServerStats()
        check(lib().mcgpu_server_get_stats(self._s, C.byref(s)))
        return s
