"""ctypes wrapper of the CPU oracle (oracle/bin/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_SO = _DIR / "bin" / "liboracle.so"
REF_TRACE = _DIR / "_ref" / "ref_trace"
REF_VOLUME = _DIR / "_ref" / "ref_volume"
_lib = None

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)


class _Box(C.Structure):
    _fields_ = [
        ("n_mol", C.c_int32), ("n_site", C.c_int32), ("n_pidx", C.c_int32),
        ("L", C.c_double * 3), ("cutoff", C.c_double), ("kT", C.c_double),
        ("max_translate", C.c_double), ("max_rotate", C.c_double),
        ("x", _d), ("y", _d), ("z", _d),
        ("sigma", _d), ("epsilon", _d), ("charge", _d),
        ("mol_start", _i32), ("mol_len", _i32), ("mol_pidx_start", _i32),
        ("mol_pidx_count", _i32), ("pidx", _i32),
    ]


class _Rng(C.Structure):
    _fields_ = [("state", C.c_int32 * 31), ("f", C.c_int32), ("r", C.c_int32)]


class _Trace(C.Structure):
    _fields_ = [("mol", _i32), ("vertex", _i32), ("old_e", _d), ("new_e", _d),
                ("de", _d), ("accepted", C.POINTER(C.c_int8))]


class _Npt(C.Structure):
    _fields_ = [("volume_interval", C.c_int32), ("reserved", C.c_int32),
                ("pressure", C.c_double), ("dv", C.c_double)]


class _VTrace(C.Structure):
    _fields_ = [("cap", C.c_int64), ("n", C.c_int64), ("step", C.POINTER(C.c_int64)),
                ("v_old", _d), ("v_new", _d), ("e_old", _d), ("e_new", _d), ("u", _d),
                ("accepted", C.POINTER(C.c_int8))]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _SO.exists():
            subprocess.run(["make", "-s", "-C", str(_DIR), "oracle"], check=True)
        _lib = C.CDLL(str(_SO))
        for name in ["orc_make_periodic", "orc_dist2", "orc_lj", "orc_coulomb",
                     "orc_lj_params", "orc_coulomb_params", "orc_pair_sum",
                     "orc_molecule_energy", "orc_system_energy",
                     "orc_lrc_reference", "orc_lrc_species", "orc_run_nvt",
                     "orc_random_real", "orc_volume_accept_prob", "orc_run_npt"]:
            getattr(_lib, name).restype = C.c_double
        _lib.orc_lj_params.argtypes = [C.c_double] * 5
        _lib.orc_coulomb_params.argtypes = [C.c_double] * 3
        _lib.orc_make_periodic.argtypes = [C.c_double, C.c_double]
        _lib.orc_volume_accept_prob.argtypes = [C.c_int32] + [C.c_double] * 5
        _lib.orc_random_real.argtypes = [C.POINTER(_Rng), C.c_double, C.c_double]
        _lib.orc_glibc_rand.restype = C.c_int32
        _lib.orc_libc_rand.restype = C.c_int32
        _lib.orc_inrange_pairs.restype = C.c_int64
        _lib.orc_volume_scale.argtypes = [C.c_void_p, C.c_double]
    return _lib


class OracleBox:
    """A private copy of a box's arrays in the oracle's layout."""

    def __init__(self, arrays: dict):
        self.a = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                  for k, v in arrays.items()}
        a = self.a
        for k in ("x", "y", "z", "sigma", "epsilon", "charge"):
            a[k] = np.ascontiguousarray(a[k], dtype=np.float64)
        for k in ("mol_start", "mol_len", "mol_pidx_start", "mol_pidx_count",
                  "primary_index"):
            a[k] = np.ascontiguousarray(a[k], dtype=np.int32)
        b = _Box()
        b.n_mol = len(a["mol_start"])
        b.n_site = len(a["x"])
        b.n_pidx = len(a["primary_index"])
        b.L[:] = [float(v) for v in a["box"]]
        b.cutoff, b.kT = a["cutoff"], a["kT"]
        b.max_translate, b.max_rotate = a["max_translate"], a["max_rotate"]
        dp = lambda k: a[k].ctypes.data_as(_d)
        ip = lambda k: a[k].ctypes.data_as(_i32)
        b.x, b.y, b.z = dp("x"), dp("y"), dp("z")
        b.sigma, b.epsilon, b.charge = dp("sigma"), dp("epsilon"), dp("charge")
        b.mol_start, b.mol_len = ip("mol_start"), ip("mol_len")
        b.mol_pidx_start, b.mol_pidx_count = ip("mol_pidx_start"), ip("mol_pidx_count")
        b.pidx = ip("primary_index")
        self.box = b

    @property
    def ref(self):
        return C.byref(self.box)

    @property
    def L(self):
        return np.array(self.box.L[:])

    def molecule_energy(self, mol: int, start: int = 0) -> float:
        return lib().orc_molecule_energy(self.ref, int(mol), int(start))

    def system_energy(self) -> float:
        return lib().orc_system_energy(self.ref)

    def pair_sum(self, m1: int, m2: int) -> float:
        return lib().orc_pair_sum(self.ref, int(m1), int(m2))

    def in_range(self, m1: int, m2: int) -> bool:
        return bool(lib().orc_in_range(self.ref, int(m1), int(m2)))

    def lrc_reference(self) -> float:
        return lib().orc_lrc_reference(self.ref)

    def lrc_species(self) -> float:
        return lib().orc_lrc_species(self.ref)

    def inrange_pairs(self) -> np.ndarray:
        n = lib().orc_inrange_pairs(self.ref, None, C.c_int64(0))
        out = np.zeros(2 * max(n, 1), np.int32)
        lib().orc_inrange_pairs(self.ref, out.ctypes.data_as(_i32), C.c_int64(n))
        return out[: 2 * n].reshape(-1, 2)

    def volume_scale(self, f: float) -> None:
        lib().orc_volume_scale(self.ref, C.c_double(f))

    def apply_move(self, mol: int, vertex: int, d, r) -> np.ndarray:
        d = np.ascontiguousarray(d, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        rb = np.zeros(3 * int(self.a["mol_len"][mol]))
        lib().orc_apply_move(self.ref, int(mol), int(vertex), d.ctypes.data_as(_d),
                             r.ctypes.data_as(_d), rb.ctypes.data_as(_d))
        return rb

    def molecule_xyz(self, mol: int) -> np.ndarray:
        s, n = int(self.a["mol_start"][mol]), int(self.a["mol_len"][mol])
        return np.concatenate([self.a["x"][s:s + n], self.a["y"][s:s + n],
                               self.a["z"][s:s + n]])

    def run_nvt(self, seed: int, steps: int, use_libc: bool = True):
        """The reference loop; returns (final_E, E0, n_acc, trace dict)."""
        t, tr = self._trace(steps)
        e0 = C.c_double()
        nacc = C.c_int64()
        fin = lib().orc_run_nvt(self.ref, C.c_uint32(seed), C.c_int64(steps),
                                int(use_libc), C.byref(tr), C.byref(e0), C.byref(nacc))
        return fin, e0.value, nacc.value, t

    def _trace(self, steps):
        t = {k: np.zeros(steps, dt) for k, dt in
             (("mol", np.int32), ("vertex", np.int32), ("old_e", np.float64),
              ("new_e", np.float64), ("de", np.float64), ("accepted", np.int8))}
        tr = _Trace(t["mol"].ctypes.data_as(_i32), t["vertex"].ctypes.data_as(_i32),
                    t["old_e"].ctypes.data_as(_d), t["new_e"].ctypes.data_as(_d),
                    t["de"].ctypes.data_as(_d),
                    t["accepted"].ctypes.data_as(C.POINTER(C.c_int8)))
        return t, tr

    def run_npt(self, seed: int, steps: int, volume_interval: int, pressure: float,
                dv: float, use_libc: bool = False):
        """The build's NPT loop (volume move after every volume_interval
        moves, fixed-width step dv in A^3).  Returns (final_E, E0, n_acc,
        move trace dict, volume-move trace dict)."""
        t, tr = self._trace(steps)
        cap = steps // max(volume_interval, 1) + 1
        v = {k: np.zeros(cap, dt) for k, dt in
             (("step", np.int64), ("v_old", np.float64), ("v_new", np.float64),
              ("e_old", np.float64), ("e_new", np.float64), ("u", np.float64),
              ("accepted", np.int8))}
        vt = _VTrace(cap, 0, v["step"].ctypes.data_as(C.POINTER(C.c_int64)),
                     *(v[k].ctypes.data_as(_d) for k in ("v_old", "v_new", "e_old",
                                                          "e_new", "u")),
                     v["accepted"].ctypes.data_as(C.POINTER(C.c_int8)))
        p = _Npt(int(volume_interval), 0, float(pressure), float(dv))
        e0 = C.c_double()
        nacc = C.c_int64()
        fin = lib().orc_run_npt(self.ref, C.c_uint32(seed), C.c_int64(steps), C.byref(p),
                                int(use_libc), C.byref(tr), C.byref(vt), C.byref(e0),
                                C.byref(nacc))
        n = int(vt.n)
        return fin, e0.value, nacc.value, t, {k: a[:n] for k, a in v.items()}


def reference_volume_decisions(state_path, rows):
    """The reference's own SimCalcs::acceptVolumeMove (oracle/_ref/ref_volume)
    on rows of (seed, n_molecules, kT, v_old, Lx, Ly, Lz, dE, pressure);
    returns its decisions (0/1), or None when the reference build is absent."""
    if not REF_VOLUME.exists():
        return None
    text = "".join("%d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % tuple(r)
                   for r in rows)
    out = subprocess.run([str(REF_VOLUME), str(state_path)], input=text,
                         capture_output=True, text=True, check=True).stdout
    return np.array([int(x) for x in out.split()], np.int8)


class GlibcRng:
    """The restated glibc TYPE_3 generator."""

    def __init__(self, seed: int):
        self.s = _Rng()
        lib().orc_glibc_srand(C.byref(self.s), C.c_uint32(seed))

    def rand(self) -> int:
        return lib().orc_glibc_rand(C.byref(self.s))


def libc_rand_sequence(seed: int, n: int) -> list:
    lib().orc_libc_srand(C.c_uint32(seed))
    return [lib().orc_libc_rand() for _ in range(n)]
