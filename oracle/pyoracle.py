"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check the engine against.  The engine never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from implicit_dpd_amd.capi import MdpdChemParams, MdpdConfigSpec, MdpdObservables, MdpdParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

P = C.POINTER
VP = C.c_void_p
_U32P = P(C.c_uint32)
_DP = P(C.c_double)

_PROTOS = {
    "oracle_saru_batch": (None, [C.c_uint32, _U32P, _U32P, _U32P, _U32P, _DP, _DP]),
    "oracle_pair_nrml": (C.c_double, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "oracle_calc_inter_coef": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _DP,
                                      C.c_double, C.c_double, C.c_uint32, C.c_double,
                                      P(MdpdParams)]),
    "oracle_pairs": (C.c_uint64, [P(MdpdParams), C.c_uint32, _DP, C.c_double, C.c_int32,
                                  _U32P, _U32P, C.c_uint64]),
    "oracle_density": (None, [P(MdpdParams), C.c_uint32, _DP, _U32P, _DP, _DP, _U32P]),
    "oracle_force": (None, [P(MdpdParams), C.c_uint32, _DP, _DP, _U32P, _U32P, _DP,
                            C.c_uint32, _DP, _DP]),
    "oracle_force_abs": (None, [P(MdpdParams), C.c_uint32, _DP, _DP, _U32P, _U32P, _DP,
                                C.c_uint32, _DP]),
    "oracle_bonded": (None, [P(MdpdParams), C.c_uint32, _U32P, _DP, _DP, _DP, _DP]),
    "oracle_drift": (None, [P(MdpdParams), C.c_uint32, _DP, _DP, _DP, _DP, _DP]),
    "oracle_kick": (None, [P(MdpdParams), C.c_uint32, _DP, _DP, _DP]),
    "oracle_remove_cm_drift": (None, [C.c_uint32, _DP]),
    "oracle_observe": (None, [P(MdpdParams), C.c_uint32, C.c_uint32, C.c_uint32, _DP, _DP,
                              _DP, _DP, P(MdpdObservables)]),
    "oracle_head_conversion": (C.c_uint32, [P(MdpdParams), C.c_uint32, _U32P, _U32P, _U32P,
                                            C.c_uint32]),
    "oracle_sys_new": (VP, [P(MdpdParams), C.c_uint32, C.c_uint32]),
    "oracle_sys_free": (None, [VP]),
    "oracle_sys_set": (C.c_int, [VP, C.c_uint32, _DP, _DP, _DP, _DP, _U32P, _U32P, _U32P, _U32P]),
    "oracle_sys_set_topology": (C.c_int, [VP, C.c_uint32, _U32P, _U32P, C.c_uint32]),
    "oracle_sys_set_clock": (None, [VP, C.c_uint32]),
    "oracle_sys_get": (C.c_uint32, [VP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _U32P, _U32P,
                                    _U32P, _U32P]),
    "oracle_sys_num_amp": (C.c_uint32, [VP]),
    "oracle_sys_get_topology": (None, [VP, _U32P]),
    "oracle_sys_drift": (None, [VP]),
    "oracle_sys_density": (None, [VP]),
    "oracle_sys_force": (None, [VP]),
    "oracle_sys_bonded": (None, [VP, _DP]),
    "oracle_sys_kick": (None, [VP]),
    "oracle_sys_remove_cm_drift": (None, [VP]),
    "oracle_sys_head_conversion": (C.c_uint32, [VP]),
    "oracle_sys_chem_event": (C.c_int, [VP, P(MdpdChemParams), P(C.c_int32)]),
    "oracle_sys_set_mt": (None, [VP, C.c_uint32, C.c_int32]),
    "oracle_sys_mt_draws": (C.c_uint64, [VP]),
    "oracle_mt_u32": (None, [C.c_uint32, C.c_uint32, _U32P]),
    "oracle_sys_observe": (None, [VP, C.c_uint32, P(MdpdObservables)]),
    "oracle_sys_run": (C.c_int, [VP, C.c_uint32]),
    "oracle_num_threads": (C.c_int, []),
    "oracle_set_num_threads": (None, [C.c_int]),
    "oracle_config_count": (C.c_int, [P(MdpdConfigSpec), _U32P]),
    "oracle_make_config": (C.c_int, [P(MdpdConfigSpec), C.c_uint32, _DP, _DP, _U32P, _U32P, _U32P,
                                     _U32P, _DP]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def _u(a):
    return a.ctypes.data_as(_U32P) if a is not None else None


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def saru_batch(s1, s2, s3):
    s1, s2, s3 = u32(s1), u32(s2), u32(s3)
    n = len(s1)
    o_u = np.zeros(8 * n, np.uint32)
    o_uni = np.zeros(2 * n, np.float64)
    o_n = np.zeros(n, np.float64)
    lib().oracle_saru_batch(n, _u(s1), _u(s2), _u(s3), _u(o_u), _d(o_uni), _d(o_n))
    return o_u.reshape(n, 8), o_uni.reshape(n, 2), o_n


def pair_nrml(id_i, id_j, seed3):
    return lib().oracle_pair_nrml(int(id_i), int(id_j), int(seed3))


def calc_inter_coef(chi, kappa, rho_co, vbb, cf_r6, dt, reo=1.4, all_unit=4, arc=0.9):
    p = MdpdParams()
    c6 = f64(cf_r6)
    lib().oracle_calc_inter_coef(chi, kappa, rho_co, vbb, _d(c6), dt, reo, all_unit, arc,
                                 C.byref(p))
    return p


def pairs(p, pos, cutoff=1.0, brute=False):
    pos = f64(pos)
    n = len(pos)
    cap = max(1024, int(n * 64))
    oi = np.zeros(cap, np.uint32)
    oj = np.zeros(cap, np.uint32)
    cnt = lib().oracle_pairs(C.byref(p), n, _d(pos), cutoff, int(brute), _u(oi), _u(oj), cap)
    if cnt > cap:
        oi = np.zeros(cnt, np.uint32)
        oj = np.zeros(cnt, np.uint32)
        cnt = lib().oracle_pairs(C.byref(p), n, _d(pos), cutoff, int(brute), _u(oi), _u(oj), cnt)
    return oi[:cnt].copy(), oj[:cnt].copy()


def density(p, pos, prop, with_cm=False):
    pos = f64(pos)
    prop = u32(prop)
    n = len(pos)
    dens = np.zeros((n, 3))
    cm = np.zeros((n, 6)) if with_cm else None
    cnt = np.zeros((n, 2), np.uint32) if with_cm else None
    lib().oracle_density(C.byref(p), n, _d(pos), _u(prop), _d(dens), _d(cm), _u(cnt))
    return (dens, cm, cnt) if with_cm else dens


def force(p, pos, vel, ids, prop, dens, seed3):
    pos, vel, dens = f64(pos), f64(vel), f64(dens)
    ids, prop = u32(ids), u32(prop)
    n = len(pos)
    acc = np.zeros((n, 3))
    press = np.zeros((n, 3))
    lib().oracle_force(C.byref(p), n, _d(pos), _d(vel), _u(ids), _u(prop), _d(dens), seed3,
                       _d(acc), _d(press))
    return acc, press


def force_abs(p, pos, vel, ids, prop, dens, seed3):
    pos, vel, dens = f64(pos), f64(vel), f64(dens)
    ids, prop = u32(ids), u32(prop)
    n = len(pos)
    out = np.zeros(n)
    lib().oracle_force_abs(C.byref(p), n, _d(pos), _d(vel), _u(ids), _u(prop), _d(dens), seed3,
                           _d(out))
    return out


def bonded(p, topo, cf_bend, pos, acc=None):
    topo = u32(topo)
    cf_bend = f64(cf_bend)
    pos = f64(pos)
    n_amp = len(topo) // 4
    acc = np.zeros_like(pos) if acc is None else f64(acc).copy()
    vir = np.zeros(3)
    lib().oracle_bonded(C.byref(p), n_amp, _u(topo), _d(cf_bend), _d(pos), _d(acc), _d(vir))
    return acc, vir


def drift(p, pos, vel, vel_buf, acc, delta):
    pos, vel, vel_buf, acc, delta = (f64(x).copy() for x in (pos, vel, vel_buf, acc, delta))
    lib().oracle_drift(C.byref(p), len(pos), _d(pos), _d(vel), _d(vel_buf), _d(acc), _d(delta))
    return pos, vel, vel_buf, delta


def kick(p, vel, vel_buf, acc):
    vel = f64(vel).copy()
    lib().oracle_kick(C.byref(p), len(vel), _d(vel), _d(f64(vel_buf)), _d(f64(acc)))
    return vel


def remove_cm_drift(vel):
    vel = f64(vel).copy()
    lib().oracle_remove_cm_drift(len(vel), _d(vel))
    return vel


def observe(p, vel, press, delta, bonded_vir, amp_num, sol_num):
    o = MdpdObservables()
    vel, press, delta = f64(vel), f64(press), f64(delta)
    bv = f64(bonded_vir)
    lib().oracle_observe(C.byref(p), len(vel), amp_num, sol_num, _d(vel), _d(press), _d(delta),
                         _d(bv), C.byref(o))
    return o


class System:
    """Whole-system oracle (one step = main.cpp:167-196)."""

    def __init__(self, p, cap, amp_cap=0):
        self.p = p
        self.cap = cap
        self.h = lib().oracle_sys_new(C.byref(p), cap, max(amp_cap, 1))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_sys_free(self.h)
            self.h = None

    def set(self, pos, vel, ids, prop, amp_id=None, unit=None, vel_buf=None, acc=None):
        pos, vel = f64(pos), f64(vel)
        n = len(pos)
        rc = lib().oracle_sys_set(self.h, n, _d(pos), _d(vel),
                                  _d(f64(vel_buf)) if vel_buf is not None else None,
                                  _d(f64(acc)) if acc is not None else None,
                                  _u(u32(ids)), _u(u32(prop)),
                                  _u(u32(amp_id)) if amp_id is not None else None,
                                  _u(u32(unit)) if unit is not None else None)
        assert rc == 0

    def set_topology(self, topo, core_amp_id=()):
        topo = u32(topo)
        core = u32(core_amp_id) if len(core_amp_id) else np.zeros(1, np.uint32)
        rc = lib().oracle_sys_set_topology(self.h, len(topo) // 4, _u(topo), _u(core),
                                           len(core_amp_id))
        assert rc == 0

    def set_clock(self, t):
        lib().oracle_sys_set_clock(self.h, t)

    def get(self):
        n = self.cap
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "vel_buf", "acc", "press", "dens",
                                              "delta")}
        ui = {k: np.zeros(n, np.uint32) for k in ("id", "prop", "amp_id", "unit")}
        m = lib().oracle_sys_get(self.h, *[_d(out[k]) for k in ("pos", "vel", "vel_buf", "acc",
                                                                "press", "dens", "delta")],
                                 *[_u(ui[k]) for k in ("id", "prop", "amp_id", "unit")])
        res = {k: v[:m].copy() for k, v in out.items()}
        res.update({k: v[:m].copy() for k, v in ui.items()})
        return res

    def topology(self):
        na = lib().oracle_sys_num_amp(self.h)
        t = np.zeros(4 * max(na, 1), np.uint32)
        lib().oracle_sys_get_topology(self.h, _u(t))
        return t[: 4 * na]

    def drift(self):
        lib().oracle_sys_drift(self.h)

    def density(self):
        lib().oracle_sys_density(self.h)

    def force(self):
        lib().oracle_sys_force(self.h)

    def bonded(self):
        v = np.zeros(3)
        lib().oracle_sys_bonded(self.h, _d(v))
        return v

    def kick(self):
        lib().oracle_sys_kick(self.h)

    def remove_cm_drift(self):
        lib().oracle_sys_remove_cm_drift(self.h)

    def head_conversion(self):
        return lib().oracle_sys_head_conversion(self.h)

    def set_mt(self, seed, enable=True):
        """Reference-MT mode of chem_event: one MT19937 stream seeded like
        ChemManager(seed), consumed in the reference's order."""
        lib().oracle_sys_set_mt(self.h, seed, 1 if enable else 0)

    def mt_draws(self):
        return int(lib().oracle_sys_mt_draws(self.h))

    def chem_event(self, cp):
        stop = C.c_int32(0)
        k = lib().oracle_sys_chem_event(self.h, C.byref(cp), C.byref(stop))
        return k, stop.value

    def observe(self, sol_num=0):
        o = MdpdObservables()
        lib().oracle_sys_observe(self.h, sol_num, C.byref(o))
        return o

    def run(self, nsteps):
        return lib().oracle_sys_run(self.h, nsteps)


def num_threads():
    return lib().oracle_num_threads()


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


def make_config(spec):
    """oracle_make_config: (status, dict of pos, vel [n, 3], ids, prop,
    amp_id, unit, kin_temp)."""
    n = C.c_uint32()
    rc = lib().oracle_config_count(C.byref(spec), C.byref(n))
    if rc:
        return rc, None
    n = n.value
    pos = np.full((n, 3), np.nan)
    vel = np.zeros((n, 3))
    u = [np.full(n, 0xDEADBEEF, np.uint32) for _ in range(4)]
    kt = C.c_double()
    rc = lib().oracle_make_config(C.byref(spec), n, _d(pos), _d(vel), *[_u(a) for a in u], C.byref(kt))
    return rc, dict(pos=pos, vel=vel, ids=u[0], prop=u[1], amp_id=u[2], unit=u[3], kin_temp=kt.value)


def mt_u32(seed, n):
    """The first n genrand_int32 outputs of MT19937 after init_genrand(seed)."""
    out = np.zeros(n, np.uint32)
    lib().oracle_mt_u32(seed, n, _u(out))
    return out
