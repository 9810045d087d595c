"""ctypes binding of libnbx.so (include/nbx.h).

The library is loaded from ``nbodysimulator.jl_amd/lib/libnbx.so`` (built in-tree by
``nbodysimulator.jl_amd/build.py``).  There is no fallback: if the library is
missing, every entry point raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libnbx.so")

NBX_OK = 0
NBX_ERR_INVALID = -1
NBX_ERR_BOUNDS = -2
NBX_ERR_UNSUPPORTED = -3
NBX_ERR_DEVICE = -4
NBX_ERR_CAPACITY = -5
NBX_ERR_NONFINITE = -6
RDF_NOALLOC_MAXBIN = 16384   # NBX_RDF_NOALLOC_MAXBIN

BC_INFINITE, BC_CUBIC, BC_PERIODIC = 0, 1, 2
SYS_POTENTIAL, SYS_WATER = 0, 1
PREC_FP64_REF, PREC_MIXED = 0, 1
THERM_NONE, THERM_ANDERSEN, THERM_BERENDSEN, THERM_LANGEVIN, THERM_NOSEHOOVER = 0, 1, 2, 3, 4

_P = C.c_void_p
_D = C.c_double
_I32 = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol here is declared in include/nbx.h
SIGNATURES = {
    "nbx_create": (C.c_int, [C.POINTER(_P), _I64, _I32, _DP, _DP]),
    "nbx_destroy": (C.c_int, [_P]),
    "nbx_last_error": (C.c_char_p, [_P]),
    "nbx_version": (C.c_char_p, []),
    "nbx_device_count": (C.c_int, []),
    "nbx_set_boundary": (C.c_int, [_P, _I32, _DP]),
    "nbx_set_kb": (C.c_int, [_P, _D]),
    "nbx_set_precision": (C.c_int, [_P, _I32]),
    "nbx_add_lennard_jones": (C.c_int, [_P, _D, _D, _D]),
    "nbx_add_electrostatic": (C.c_int, [_P, _D, _D]),
    "nbx_add_reaction_field": (C.c_int, [_P, _D, _D, _D]),
    "nbx_add_gravitational": (C.c_int, [_P, _D]),
    "nbx_set_moments": (C.c_int, [_P, _DP]),
    "nbx_add_magnetostatic": (C.c_int, [_P, _D]),
    "nbx_set_spcfw": (C.c_int, [_P, _D, _D, _D, _D]),
    "nbx_set_thermostat": (C.c_int, [_P, _I32, _D, _D]),
    "nbx_set_skin": (C.c_int, [_P, _D]),
    "nbx_set_state": (C.c_int, [_P, _DP, _DP]),
    "nbx_get_state": (C.c_int, [_P, _DP, _DP]),
    "nbx_rhs": (C.c_int, [_P, _DP, _DP, _DP]),
    "nbx_accelerations": (C.c_int, [_P, _I32, _DP, _I64, _I64, _DP]),
    "nbx_vv_run": (C.c_int, [_P, _D, _I64, _U64, _I64, _DP, _DP]),
    "nbx_em_run": (C.c_int, [_P, _D, _I64, _U64, _I64, _DP, _DP]),
    "nbx_steps_taken": (_I64, [_P]),
    "nbx_get_thermostat_state": (C.c_int, [_P, _DP, _DP]),
    "nbx_energies": (C.c_int, [_P, _DP, _DP, _DP]),
    "nbx_temperature": (C.c_int, [_P, _DP]),
    "nbx_rdf": (C.c_int, [_P, _DP, _I64, _I64, _DP, _DP]),
    "nbx_rdf_noalloc": (C.c_int, [_P, _DP, _I64, _I64, _DP, _DP]),
    "nbx_time_force_kernel": (C.c_int, [_P, _I32, _DP, C.POINTER(C.c_char_p)]),
    "nbx_count_pairs": (C.c_int, [_P, C.POINTER(_I64)]),
    "nbx_set_kernel_timing": (C.c_int, [_P, _I32]),
    "nbx_get_kernel_timing": (C.c_int, [_P, _DP, C.POINTER(_I64), C.POINTER(C.c_char_p)]),
    "nbx_copy_probe": (C.c_int, [_P, _I64, _I32, _DP]),
    "nbx_valu_probe": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "nbx_get_stats": (C.c_int, [_P, C.POINTER(_I64), _I32]),
    "nbx_autotune": (C.c_int, [_P, _D, _I32, _DP, _I32]),
    "nbx_batch_create": (C.c_int, [C.POINTER(_P), _I32, C.POINTER(_I64), _DP, _DP, _D, _D,
                                   _D, _D]),
    "nbx_batch_destroy": (C.c_int, [_P]),
    "nbx_batch_last_error": (C.c_char_p, [_P]),
    "nbx_batch_set_thermostat": (C.c_int, [_P, _I32, _I32, _D, _D]),
    "nbx_batch_set_skin": (C.c_int, [_P, _D]),
    "nbx_batch_autotune": (C.c_int, [_P, _D, _I64, _DP]),
    "nbx_batch_set_state": (C.c_int, [_P, _DP, _DP]),
    "nbx_batch_get_state": (C.c_int, [_P, _DP, _DP]),
    "nbx_batch_run": (C.c_int, [_P, _D, _I64, _U64, _I64]),
    "nbx_batch_energies": (C.c_int, [_P, _DP, _DP, C.POINTER(_I64)]),
    "nbx_batch_stats": (C.c_int, [_P, C.POINTER(_I64), _I32]),
    "nbx_model_create": (C.c_int, [C.POINTER(_P), _U64]),
    "nbx_model_destroy": (C.c_int, [_P]),
    "nbx_model_last_error": (C.c_char_p, [_P]),
    "nbx_model_set": (C.c_int, [_P, _DP, _DP]),
    "nbx_model_get": (C.c_int, [_P, _DP, _DP]),
    "nbx_model_forward": (C.c_int, [_P, _DP, _I64, _DP]),
    "nbx_model_gradient": (C.c_int, [_P, _DP, _DP, _I64, _DP, _DP]),
    "nbx_model_add_samples": (C.c_int, [_P, _DP, _DP, _I64]),
    "nbx_model_get_samples": (C.c_int, [_P, _I64, _I64, _DP, _DP, C.POINTER(_I64)]),
    "nbx_model_clear_samples": (C.c_int, [_P]),
    "nbx_model_fit_normalization": (C.c_int, [_P]),
    "nbx_model_train": (C.c_int, [_P, _I64, _I32, _D, _U64, _DP]),
    "nbx_model_evaluate": (C.c_int, [_P, _I64, _I64, _DP, _DP]),
    "nbx_features": (C.c_int, [_P, _DP, _DP]),
    "nbx_learn_energy": (C.c_int, [_P, _DP, _DP]),
    "nbx_learn_collect": (C.c_int, [_P, _P, _D, _I32, _I64, _I32, _U64]),
    "nbx_learned_run": (C.c_int, [_P, _P, _D, _I32, _I64, _D, _I32, _DP]),
}

MLP_INPUTS = 35
MLP_OUTPUTS = 6
MLP_PARAMS = 38406
MLP_NORM = 76

_lib = None


def load() -> C.CDLL:
    """Load libnbx.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libnbx.so not found at {LIB_PATH}; build it with "
            "`python nbodysimulator.jl_amd/build.py` (the HIP engine has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NBXError(RuntimeError):
    """A libnbx call failed; ``code`` is the NBX_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def raise_for(code: int, msg: str):
    if code == NBX_ERR_BOUNDS:
        raise IndexError(msg)          # Julia: BoundsError
    if code == NBX_ERR_INVALID:
        raise ValueError(msg)          # Julia: ArgumentError / MethodError
    if code == NBX_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if code == NBX_ERR_NONFINITE:
        raise FloatingPointError(msg)
    raise NBXError(code, msg)


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def as_c3n(a, n: int) -> np.ndarray:
    """Coordinates as a C-contiguous float64 (n, 3) array: the memory layout of a Julia
    column-major 3 x n matrix."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.shape == (3, n) and n != 3:
        arr = np.ascontiguousarray(arr.T)
    if arr.shape != (n, 3):
        raise ValueError(f"expected coordinates of shape ({n}, 3), got {arr.shape}")
    return arr


class Context:
    """Owns one nbx_ctx (one HIP stream on the current device)."""

    def __init__(self, n: int, system_kind: int, masses, charges=None):
        lib = load()
        self._lib = lib
        self.n = int(n)
        self.masses = np.ascontiguousarray(np.asarray(masses, dtype=np.float64))
        self.charges = (None if charges is None else
                        np.ascontiguousarray(np.asarray(charges, dtype=np.float64)))
        h = _P()
        rc = lib.nbx_create(C.byref(h), self.n, system_kind, dptr(self.masses),
                            dptr(self.charges))
        if rc != NBX_OK:
            raise_for(rc, lib.nbx_last_error(None).decode())
        self._h = h

    # -- plumbing -------------------------------------------------------------
    def _check(self, rc: int):
        if rc != NBX_OK:
            raise_for(rc, self._lib.nbx_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nbx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration ---------------------------------------------------------
    def set_boundary(self, kind: int, box=None):
        b = np.zeros(6) if box is None else np.ascontiguousarray(np.asarray(box, np.float64))
        self._check(self._lib.nbx_set_boundary(self._h, kind, dptr(b)))

    def set_kb(self, kb: float):
        self._check(self._lib.nbx_set_kb(self._h, float(kb)))

    def set_precision(self, mode: int):
        self._check(self._lib.nbx_set_precision(self._h, mode))

    def add_lennard_jones(self, eps, sigma, R):
        self._check(self._lib.nbx_add_lennard_jones(self._h, float(eps), float(sigma), float(R)))

    def add_electrostatic(self, k, R):
        self._check(self._lib.nbx_add_electrostatic(self._h, float(k), float(R)))

    def add_reaction_field(self, k, R, eps_rf):
        self._check(self._lib.nbx_add_reaction_field(self._h, float(k), float(R), float(eps_rf)))

    def add_gravitational(self, G):
        self._check(self._lib.nbx_add_gravitational(self._h, float(G)))

    def set_moments(self, mm):
        """Magnetic moments, (n, 3) (the memory layout of a Julia 3 x n matrix)."""
        self._mm = as_c3n(mm, self.n)
        self._check(self._lib.nbx_set_moments(self._h, dptr(self._mm)))

    def add_magnetostatic(self, mu_4pi):
        self._check(self._lib.nbx_add_magnetostatic(self._h, float(mu_4pi)))

    def set_spcfw(self, rOH, aHOH, kbond, kangle):
        self._check(self._lib.nbx_set_spcfw(self._h, float(rOH), float(aHOH), float(kbond),
                                            float(kangle)))

    def set_thermostat(self, kind: int, T: float = 0.0, param: float = 1.0):
        self._check(self._lib.nbx_set_thermostat(self._h, kind, float(T), float(param)))

    def set_skin(self, skin: float):
        self._check(self._lib.nbx_set_skin(self._h, float(skin)))

    # -- state and evaluation ------------------------------------------------------
    def set_state(self, u, v=None):
        u = as_c3n(u, self.n)
        v = None if v is None else as_c3n(v, self.n)
        self._check(self._lib.nbx_set_state(self._h, dptr(u), dptr(v)))

    def get_state(self):
        u = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        self._check(self._lib.nbx_get_state(self._h, dptr(u), dptr(v)))
        return u, v

    def rhs(self, u, v=None) -> np.ndarray:
        u = as_c3n(u, self.n)
        v = None if v is None else as_c3n(v, self.n)
        dv = np.empty((self.n, 3))
        self._check(self._lib.nbx_rhs(self._h, dptr(u), dptr(v), dptr(dv)))
        return dv

    def accelerations(self, term: int, u, i_begin: int, i_end: int, dv: np.ndarray):
        """Add term `term`'s contribution for atoms [i_begin, i_end) into dv (n, 3)."""
        u = as_c3n(u, self.n)
        if dv.dtype != np.float64 or not dv.flags.c_contiguous or dv.shape != (self.n, 3):
            raise ValueError("dv must be a C-contiguous float64 (n, 3) array")
        self._check(self._lib.nbx_accelerations(self._h, int(term), dptr(u), int(i_begin),
                                                int(i_end), dptr(dv)))
        return dv

    def vv_run(self, dt: float, nsteps: int, seed: int = 0, save_every: int = 0):
        nframes = 1 + nsteps // save_every if save_every > 0 else 0
        uf = vf = None
        if nframes:
            uf = np.empty((nframes, self.n, 3))
            vf = np.empty((nframes, self.n, 3))
        self._check(self._lib.nbx_vv_run(self._h, float(dt), int(nsteps), int(seed),
                                          int(save_every), dptr(uf), dptr(vf)))
        return uf, vf

    def em_run(self, dt: float, nsteps: int, seed: int = 0, save_every: int = 0):
        """Euler-Maruyama steps of the Langevin SDE (nbx_em_run); frames as vv_run."""
        nframes = 1 + nsteps // save_every if save_every > 0 else 0
        uf = vf = None
        if nframes:
            uf = np.empty((nframes, self.n, 3))
            vf = np.empty((nframes, self.n, 3))
        self._check(self._lib.nbx_em_run(self._h, float(dt), int(nsteps), int(seed),
                                          int(save_every), dptr(uf), dptr(vf)))
        return uf, vf

    def steps_taken(self) -> int:
        return int(self._lib.nbx_steps_taken(self._h))

    def thermostat_state(self):
        """(zeta, xi) of a Nose-Hoover thermostat: the reference's u[3n+1], v[3n+1]."""
        z, x = C.c_double(), C.c_double()
        self._check(self._lib.nbx_get_thermostat_state(self._h, C.byref(z), C.byref(x)))
        return z.value, x.value

    def energies(self):
        ke, pe, pg = C.c_double(), C.c_double(), C.c_double()
        self._check(self._lib.nbx_energies(self._h, C.byref(ke), C.byref(pe), C.byref(pg)))
        return ke.value, pe.value, pg.value

    def temperature(self) -> float:
        t = C.c_double()
        self._check(self._lib.nbx_temperature(self._h, C.byref(t)))
        return t.value

    def rdf(self, frames: np.ndarray, maxbin: int = 1000):
        f = np.ascontiguousarray(np.asarray(frames, dtype=np.float64))
        if f.ndim != 3 or f.shape[1:] != (self.n, 3):
            raise ValueError("frames must have shape (nframes, n, 3)")
        rs = np.empty(maxbin)
        gr = np.empty(maxbin)
        self._check(self._lib.nbx_rdf(self._h, dptr(f), f.shape[0], maxbin, dptr(rs), dptr(gr)))
        return rs, gr

    def rdf_noalloc(self, frames: np.ndarray, maxbin: int = 1000):
        """rdf through nbx_rdf_noalloc: no device allocation (library-owned histogram of
        at most RDF_NOALLOC_MAXBIN bins)."""
        f = np.ascontiguousarray(np.asarray(frames, dtype=np.float64))
        if f.ndim != 3 or f.shape[1:] != (self.n, 3):
            raise ValueError("frames must have shape (nframes, n, 3)")
        rs = np.empty(maxbin)
        gr = np.empty(maxbin)
        self._check(self._lib.nbx_rdf_noalloc(self._h, dptr(f), f.shape[0], maxbin, dptr(rs),
                                              dptr(gr)))
        return rs, gr

    def set_kernel_timing(self, enable: bool) -> None:
        """Time every force-kernel launch of later mixed-path runs (HIP event pairs on the
        context's stream; no step graphs while on).  Resets the counters."""
        self._check(self._lib.nbx_set_kernel_timing(self._h, 1 if enable else 0))

    def kernel_timing(self):
        """(average ms, launches, kernel name) of the force launches timed so far."""
        ms = C.c_double()
        cnt = _I64()
        name = C.c_char_p()
        self._check(self._lib.nbx_get_kernel_timing(self._h, C.byref(ms), C.byref(cnt),
                                                    C.byref(name)))
        return ms.value, cnt.value, name.value.decode()

    def time_force_kernel(self, iters: int):
        ms = C.c_double()
        name = C.c_char_p()
        self._check(self._lib.nbx_time_force_kernel(self._h, int(iters), C.byref(ms),
                                                    C.byref(name)))
        return ms.value, name.value.decode()

    def count_pairs(self) -> int:
        p = C.c_int64()
        self._check(self._lib.nbx_count_pairs(self._h, C.byref(p)))
        return p.value

    def valu_probe(self, kind: int, reps: int = 3) -> float:
        """Wave64 VALU instructions per SIMD per ns of an independent stream of one
        instruction kind (0 v_fmac_f32, 1 v_mul_f32, 2 v_rsq_f32)."""
        r = C.c_double()
        self._check(self._lib.nbx_valu_probe(self._h, int(kind), int(reps), C.byref(r)))
        return r.value

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """Average duration (ms) of a plain copy kernel moving `nbytes` (read + write)."""
        ms = C.c_double()
        self._check(self._lib.nbx_copy_probe(self._h, int(nbytes), int(iters), C.byref(ms)))
        return ms.value

    def autotune(self, dt: float, trial_steps: int = 200) -> dict:
        """Time the candidate launch shapes of the mixed path and keep the fastest."""
        out = np.zeros(6)
        self._check(self._lib.nbx_autotune(self._h, float(dt), int(trial_steps), dptr(out), 6))
        return dict(shape0=out[0], shape1=out[1], best_ms_per_step=out[2],
                    entry_ms_per_step=out[3], candidates=int(out[4]), build_every=int(out[5]))

    # -- learned stride -------------------------------------------------------
    def features(self):
        """Descriptor of the current state: (feats (n, 35), frames (n, 6))."""
        f = np.empty((self.n, MLP_INPUTS))
        fr = np.empty((self.n, 6))
        self._check(self._lib.nbx_features(self._h, dptr(f), dptr(fr)))
        return f, fr

    def learn_energy(self):
        """(kinetic, shifted Lennard-Jones potential) energy from the Verlet list."""
        ke, pe = C.c_double(), C.c_double()
        self._check(self._lib.nbx_learn_energy(self._h, C.byref(ke), C.byref(pe)))
        return ke.value, pe.value

    def learn_collect(self, model, dt, stride, nframes, atoms_per_frame, seed=0):
        self._check(self._lib.nbx_learn_collect(self._h, model._h, float(dt), int(stride),
                                                int(nframes), int(atoms_per_frame), int(seed)))

    def learned_run(self, model, dt, stride, nsteps, drift_tol=1e-4, check_every=25) -> dict:
        r = np.zeros(9)
        self._check(self._lib.nbx_learned_run(self._h, model._h, float(dt), int(stride),
                                              int(nsteps), float(drift_tol), int(check_every),
                                              dptr(r)))
        return dict(strided_steps=int(r[0]), plain_steps=int(r[1]), tripped=bool(r[2]),
                    strided_before_trip=int(r[3]), E0=r[4], E=r[5], max_rel_dev=r[6],
                    list_builds=int(r[7]), trip_rel_dev=r[8])

    def stats(self) -> dict:
        out = (C.c_int64 * 8)()
        self._check(self._lib.nbx_get_stats(self._h, out, 8))
        return dict(list_builds=out[0], list_slots=out[1], cells_per_side=out[2],
                    lanes=out[3], graph_steps=out[4], build_every=out[5],
                    stale_replays=out[6], resorts=out[7])


class Batch:
    """Owns one nbx_batch: independent LJ replicas integrated by one persistent launch."""

    def __init__(self, natoms, box_L, masses, eps=1.0, sigma=1.0, R=2.5, kb=1.0):
        lib = load()
        self._lib = lib
        self.natoms = np.ascontiguousarray(np.asarray(natoms, dtype=np.int64))
        self.nrep = len(self.natoms)
        self.ntot = int(self.natoms.sum())
        box = np.ascontiguousarray(np.asarray(box_L, dtype=np.float64))
        ms = np.ascontiguousarray(np.asarray(masses, dtype=np.float64))
        h = _P()
        rc = lib.nbx_batch_create(C.byref(h), self.nrep,
                                  self.natoms.ctypes.data_as(C.POINTER(_I64)), dptr(box),
                                  dptr(ms), float(eps), float(sigma), float(R), float(kb))
        if rc != NBX_OK:
            raise_for(rc, lib.nbx_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != NBX_OK:
            raise_for(rc, self._lib.nbx_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nbx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_thermostat(self, replica, kind, T=0.0, param=1.0):
        self._check(self._lib.nbx_batch_set_thermostat(self._h, int(replica), int(kind),
                                                       float(T), float(param)))

    def set_skin(self, skin):
        self._check(self._lib.nbx_batch_set_skin(self._h, float(skin)))

    def autotune(self, dt, trial_steps=200) -> dict:
        """Pick the Verlet skin by timing (nbx_batch_autotune); the state is unchanged."""
        out = np.zeros(3)
        self._check(self._lib.nbx_batch_autotune(self._h, float(dt), int(trial_steps),
                                                 dptr(out)))
        return dict(skin=out[0], ms_per_step=out[1], entry_ms_per_step=out[2])

    def set_state(self, u, v=None):
        u = as_c3n(u, self.ntot)
        v = None if v is None else as_c3n(v, self.ntot)
        self._check(self._lib.nbx_batch_set_state(self._h, dptr(u), dptr(v)))

    def get_state(self):
        u = np.empty((self.ntot, 3))
        v = np.empty((self.ntot, 3))
        self._check(self._lib.nbx_batch_get_state(self._h, dptr(u), dptr(v)))
        return u, v

    def run(self, dt, nsteps, seed=0, steps_per_launch=0):
        self._check(self._lib.nbx_batch_run(self._h, float(dt), int(nsteps), int(seed),
                                            int(steps_per_launch)))

    def energies(self):
        ke = np.empty(self.nrep)
        pe = np.empty(self.nrep)
        pairs = np.empty(self.nrep, dtype=np.int64)
        self._check(self._lib.nbx_batch_energies(self._h, dptr(ke), dptr(pe),
                                                 pairs.ctypes.data_as(C.POINTER(_I64))))
        return ke, pe, pairs

    def stats(self):
        out = (C.c_int64 * 3)()
        self._check(self._lib.nbx_batch_stats(self._h, out, 3))
        return dict(list_builds=out[0], steps_taken=out[1], list_slots=out[2])


def device_count() -> int:
    return int(load().nbx_device_count())


class Model:
    """Owns one nbx_model: the learned-stride MLP, its normalization and training set."""

    def __init__(self, seed: int = 0):
        lib = load()
        self._lib = lib
        h = _P()
        rc = lib.nbx_model_create(C.byref(h), int(seed))
        if rc != NBX_OK:
            raise_for(rc, lib.nbx_model_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != NBX_OK:
            raise_for(rc, self._lib.nbx_model_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nbx_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self):
        p = np.empty(MLP_PARAMS)
        nm = np.empty(MLP_NORM)
        self._check(self._lib.nbx_model_get(self._h, dptr(p), dptr(nm)))
        return p, nm

    def set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        nm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float64)
        if p is not None and p.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        if nm is not None and nm.shape != (MLP_NORM,):
            raise ValueError(f"expected {MLP_NORM} normalization values")
        self._check(self._lib.nbx_model_set(self._h, dptr(p), dptr(nm)))

    def forward(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.float64).reshape(-1, MLP_INPUTS)
        out = np.empty((f.shape[0], MLP_OUTPUTS))
        self._check(self._lib.nbx_model_forward(self._h, dptr(f), f.shape[0], dptr(out)))
        return out

    def gradient(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float64).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, MLP_OUTPUTS)
        g = np.empty(MLP_PARAMS)
        loss = C.c_double()
        self._check(self._lib.nbx_model_gradient(self._h, dptr(f), dptr(t), f.shape[0], dptr(g),
                                                 C.byref(loss)))
        return g, loss.value

    def add_samples(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float64).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, MLP_OUTPUTS)
        if f.shape[0] != t.shape[0]:
            raise ValueError("feature and target row counts differ")
        self._check(self._lib.nbx_model_add_samples(self._h, dptr(f), dptr(t), f.shape[0]))

    @property
    def num_samples(self) -> int:
        c = C.c_int64()
        self._check(self._lib.nbx_model_get_samples(self._h, 0, 0, None, None, C.byref(c)))
        return c.value

    def get_samples(self, first=0, n=None):
        if n is None:
            n = self.num_samples - first
        f = np.empty((n, MLP_INPUTS))
        t = np.empty((n, MLP_OUTPUTS))
        self._check(self._lib.nbx_model_get_samples(self._h, int(first), int(n), dptr(f), dptr(t),
                                                    None))
        return f, t

    def clear_samples(self):
        self._check(self._lib.nbx_model_clear_samples(self._h))

    def fit_normalization(self):
        self._check(self._lib.nbx_model_fit_normalization(self._h))

    def train(self, iters, batch=4096, lr=1e-3, seed=0):
        losses = np.empty(max(int(iters), 1))
        self._check(self._lib.nbx_model_train(self._h, int(iters), int(batch), float(lr), int(seed),
                                              dptr(losses)))
        return losses[:int(iters)]

    def evaluate(self, first, n):
        mse = np.empty(MLP_OUTPUTS)
        base = np.empty(MLP_OUTPUTS)
        self._check(self._lib.nbx_model_evaluate(self._h, int(first), int(n), dptr(mse), dptr(base)))
        return mse, base
