"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so) -- test infrastructure.

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# NBX_ORACLE_SO selects another build of the same source (the ASan/UBSan one, `make -C
# oracle sanitize`)
ORACLE_SO = os.environ.get("NBX_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")
if not os.path.isabs(ORACLE_SO):
    ORACLE_SO = os.path.join(ROOT, ORACLE_SO)

BC_INFINITE, BC_CUBIC, BC_PERIODIC = 0, 1, 2
TERM_LJ, TERM_ELEC, TERM_GRAV, TERM_RF, TERM_MAG = 1, 2, 3, 4, 5
THERM_NONE, THERM_ANDERSEN, THERM_BERENDSEN, THERM_LANGEVIN, THERM_NOSEHOOVER = 0, 1, 2, 3, 4


class OrcBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("b", C.c_double * 6)]


class OrcTerm(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("p", C.c_double * 4)]


class OrcSystem(C.Structure):
    _fields_ = [("n", C.c_int64), ("water", C.c_int32), ("nterms", C.c_int32),
                ("ms", C.POINTER(C.c_double)), ("qs", C.POINTER(C.c_double)),
                ("bc", OrcBC), ("terms", OrcTerm * 8), ("water_lj", OrcTerm),
                ("water_elec", OrcTerm), ("rOH", C.c_double), ("aHOH", C.c_double),
                ("kbond", C.c_double), ("kangle", C.c_double), ("therm", C.c_int32),
                ("pad2", C.c_int32), ("therm_T", C.c_double), ("therm_p", C.c_double),
                ("kb", C.c_double), ("mm", C.POINTER(C.c_double)), ("nh_zeta", C.c_double)]


_DP = C.POINTER(C.c_double)
_LP = C.POINTER(C.c_int64)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = C.CDLL(ORACLE_SO)
    sp = C.POINTER(OrcSystem)
    bp = C.POINTER(OrcBC)
    sig = {
        "orc_interparticle_distance": (None, [_DP, _DP, bp, _DP, _DP, _DP]),
        "orc_lj_accel": (None, [_DP, _DP, C.c_int64, _LP, C.c_int64, _DP, C.c_double,
                                C.c_double, C.c_double, bp]),
        "orc_elec_accel": (None, [_DP, _DP, C.c_int64, C.c_int64, _DP, _DP, C.c_int32,
                                  C.c_double, C.c_double, bp]),
        "orc_grav_accel": (None, [_DP, _DP, C.c_int64, C.c_int64, _DP, C.c_double]),
        "orc_md_temperature": (C.c_double, [_DP, _DP, C.c_double, C.c_int64, C.c_int64]),
        "orc_rhs": (None, [sp, _DP, _DP, _DP]),
        "orc_vv_run": (C.c_int64, [sp, _DP, _DP, C.c_double, C.c_int64, C.c_uint64, C.c_int64,
                                   _DP, _DP]),
        "orc_vv_run_nh": (C.c_int64, [sp, _DP, _DP, C.c_double, C.c_int64, C.c_uint64,
                                      C.c_int64, _DP, _DP, _DP]),
        "orc_em_run": (C.c_int64, [sp, _DP, _DP, C.c_double, C.c_int64, C.c_uint64, C.c_int64,
                                   _DP, _DP]),
        "orc_kinetic_energy": (C.c_double, [_DP, _DP, C.c_int64]),
        "orc_potential_energy": (C.c_double, [sp, _DP]),
        "orc_grav_potential": (C.c_double, [_DP, C.c_int64, _DP, C.c_double]),
        "orc_lj_potential": (C.c_double, [_DP, _LP, C.c_int64, C.c_double, C.c_double,
                                          C.c_double, bp]),
        "orc_bond_potential": (C.c_double, [_DP, C.c_int64, C.c_double, C.c_double]),
        "orc_angle_potential": (C.c_double, [_DP, C.c_int64, C.c_double, C.c_double]),
        "orc_rdf": (None, [_DP, C.c_int64, C.c_int64, _LP, C.c_int64, C.c_double, C.c_int64,
                           _DP, _DP]),
        "orc_philox4x32": (None, [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)]),
        "orc_lr_features": (None, [_DP, _DP, _DP, C.c_int64, C.c_double, _DP, _DP]),
        "orc_mag_accel": (None, [_DP, _DP, C.c_int64, C.c_int64, _DP, _DP, C.c_double]),
        "orc_mag_potential": (C.c_double, [_DP, C.c_int64, _DP, C.c_double]),
        "orc_u01": (C.c_double, [C.c_uint32, C.c_uint32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def dp(a):
    return a.ctypes.data_as(_DP)


def c3n(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def make_bc(kind=BC_INFINITE, box=None) -> OrcBC:
    bc = OrcBC()
    bc.kind = kind
    if kind == BC_CUBIC:
        bc.b[0] = float(box)
    elif kind == BC_PERIODIC:
        for k in range(6):
            bc.b[k] = float(box[k])
    return bc


class System:
    """An orc_system with its arrays kept alive."""

    def __init__(self, masses, charges=None, water=False, bc=None, kb=1.0):
        self.ms = c3n(masses)
        self.qs = None if charges is None else c3n(charges)
        s = OrcSystem()
        s.n = len(self.ms)
        s.water = 1 if water else 0
        s.ms = dp(self.ms)
        s.qs = dp(self.qs) if self.qs is not None else C.POINTER(C.c_double)()
        s.bc = bc if bc is not None else make_bc()
        s.kb = kb
        s.mm = C.POINTER(C.c_double)()
        self.mm = None
        self.s = s

    def moments(self, mm):
        """Magnetic moments (n x 3) for MAG terms."""
        self.mm = c3n(mm)
        self.s.mm = dp(self.mm)
        return self

    def add(self, kind, *p):
        t = self.s.terms[self.s.nterms]
        t.kind = kind
        for k, x in enumerate(p):
            t.p[k] = float(x)
        self.s.nterms += 1
        return self

    def water_params(self, lj, elec_kind, elec, rOH, aHOH, kbond, kangle):
        self.s.water_lj.kind = TERM_LJ
        for k, x in enumerate(lj):
            self.s.water_lj.p[k] = x
        self.s.water_elec.kind = elec_kind
        for k, x in enumerate(elec):
            self.s.water_elec.p[k] = x
        self.s.rOH, self.s.aHOH, self.s.kbond, self.s.kangle = rOH, aHOH, kbond, kangle
        return self

    def thermostat(self, kind, T, p):
        self.s.therm, self.s.therm_T, self.s.therm_p = kind, T, p
        return self

    def rhs(self, u, v=None) -> np.ndarray:
        lib = load()
        u = c3n(u)
        v = np.zeros_like(u) if v is None else c3n(v)
        dv = np.empty_like(u)
        lib.orc_rhs(C.byref(self.s), dp(u), dp(v), dp(dv))
        return dv

    def vv_run(self, u, v, dt, nsteps, seed=0, save_every=0):
        lib = load()
        u = c3n(u).copy()
        v = c3n(v).copy()
        nf = 1 + nsteps // save_every if save_every > 0 else 0
        uf = np.empty((max(nf, 1),) + u.shape)
        vf = np.empty((max(nf, 1),) + u.shape)
        lib.orc_vv_run(C.byref(self.s), dp(u), dp(v), dt, nsteps, seed, save_every,
                       dp(uf) if nf else None, dp(vf) if nf else None)
        return u, v, (uf[:nf], vf[:nf])

    def vv_run_nh_state(self, u, v, dt, nsteps):
        """Final (u, v) and the Nose-Hoover (zeta, xi) after nsteps steps."""
        lib = load()
        u = c3n(u).copy()
        v = c3n(v).copy()
        nh = np.zeros(2)
        lib.orc_vv_run_nh(C.byref(self.s), dp(u), dp(v), dt, nsteps, 0, 0, None, None, dp(nh))
        return u, v, (float(nh[0]), float(nh[1]))

    def em_run(self, u, v, dt, nsteps, seed=0, save_every=0):
        lib = load()
        u = c3n(u).copy()
        v = c3n(v).copy()
        nf = 1 + nsteps // save_every if save_every > 0 else 0
        uf = np.empty((max(nf, 1),) + u.shape)
        vf = np.empty((max(nf, 1),) + u.shape)
        lib.orc_em_run(C.byref(self.s), dp(u), dp(v), dt, nsteps, seed, save_every,
                       dp(uf) if nf else None, dp(vf) if nf else None)
        return u, v, (uf[:nf], vf[:nf])

    def potential_energy(self, u) -> float:
        return load().orc_potential_energy(C.byref(self.s), dp(c3n(u)))

    def kinetic_energy(self, v) -> float:
        return load().orc_kinetic_energy(dp(c3n(v)), dp(self.ms), len(self.ms))


def interparticle_distance(ri, rj, bc: OrcBC):
    lib = load()
    ri, rj = c3n(ri), c3n(rj)
    rij = np.empty(3)
    r, r2 = C.c_double(), C.c_double()
    lib.orc_interparticle_distance(dp(ri), dp(rj), C.byref(bc), dp(rij), C.byref(r), C.byref(r2))
    return rij, r.value, r2.value


def grav_potential(u, ms, G) -> float:
    return load().orc_grav_potential(dp(c3n(u)), len(ms), dp(c3n(ms)), G)


def rdf(frames, L, indxs=None, maxbin=1000):
    lib = load()
    f = c3n(frames)
    nfr, n = f.shape[0], f.shape[1]
    idx = np.arange(n, dtype=np.int64) if indxs is None else np.asarray(indxs, np.int64)
    rs, gr = np.empty(maxbin), np.empty(maxbin)
    lib.orc_rdf(dp(f), nfr, n, idx.ctypes.data_as(_LP), len(idx), L, maxbin, dp(rs), dp(gr))
    return rs, gr


def philox(c, k):
    out = (C.c_uint32 * 4)()
    load().orc_philox4x32(*[int(x) & 0xFFFFFFFF for x in c], *[int(x) & 0xFFFFFFFF for x in k],
                          out)
    return list(out)


def lr_features(u, a, v, L):
    """Learned-stride descriptor of every atom: (feats (n, 35), frames (n, 6))."""
    lib = load()
    u, a, v = c3n(u), c3n(a), c3n(v)
    n = u.shape[0]
    f = np.empty((n, 35))
    fr = np.empty((n, 6))
    lib.orc_lr_features(dp(u), dp(a), dp(v), n, float(L), dp(f), dp(fr))
    return f, fr


def mag_potential(u, mm, mu_4pi) -> float:
    u, mm = c3n(u), c3n(mm)
    return load().orc_mag_potential(dp(u), u.shape[0], dp(mm), float(mu_4pi))
