"""Python handle on the engine's C ABI (libmdpd.so), for tests, bench and tools.

The engine itself is the HIP library; this module only marshals numpy arrays
through ctypes.  Method names follow the reference's driver calls
(drift_and_predict, CalcDensity, CalcForceEpEpDPD, ForceBonded,
kick, remove_cmdrift_global, Observer; src/main.cpp:167-208).
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import MdpdChemParams, MdpdObservables, MdpdParams, MdpdTuning, check

_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)


def _d(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _u(a):
    return None if a is None else a.ctypes.data_as(_U32P)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _nsteps(n):
    # mdpd_run takes an unsigned count: a negative one would wrap to ~4e9 steps
    n = int(n)
    if n < 0:
        raise ValueError(f"negative step count {n}")
    return n


def observables_dict(o):
    return {
        "kin_temp": np.array(o.kin_temp[:]),
        "pressure": np.array(o.pressure[:]),
        "diffusion": np.array(o.diffusion[:]),
        "bonded_vir": np.array(o.bonded_vir[:]),
        "pair_vir": np.array(o.pair_vir[:]),
        "n_pairs": float(o.n_pairs),
        "n_beads": o.n_beads,
        "n_amp": o.n_amp,
        "time": o.time,
    }


class Engine:
    """One engine context: a single system, or an ensemble of replicas."""

    def __init__(self, params, capacity=None, n_beads=None):
        L = capi.lib()
        self.L = L
        h = C.c_void_p()
        if n_beads is None:
            self.params = [params]
            check(L.mdpd_create(C.byref(params), int(capacity), C.byref(h)))
            self.n_rep = 1
        else:
            plist = params if isinstance(params, (list, tuple)) else [params] * len(n_beads)
            arr = (MdpdParams * len(plist))(*plist)
            nb = np.ascontiguousarray(n_beads, dtype=np.uint32)
            check(L.mdpd_create_ensemble(arr, _u(nb), len(plist), C.byref(h)))
            self.params = list(plist)
            self.n_rep = len(plist)
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.mdpd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return check(rc, self.h)

    # ---- state
    def upload(self, pos, vel, ids, prop, amp_id=None, unit=None, acc=None):
        pos, vel = _f64(pos), _f64(vel)
        n = len(pos)
        self._c(self.L.mdpd_upload_soa(self.h, n, _d(pos), _d(vel),
                                       _d(_f64(acc)) if acc is not None else None,
                                       _u(_u32(ids)), _u(_u32(prop)),
                                       _u(_u32(amp_id)) if amp_id is not None else None,
                                       _u(_u32(unit)) if unit is not None else None))

    def upload_fpdpd(self, records):
        buf = np.ascontiguousarray(records)
        self._c(self.L.mdpd_upload_fpdpd(self.h, buf.ctypes.data_as(C.c_void_p), len(buf),
                                         buf.dtype.itemsize))

    def download_fpdpd(self, n=None):
        from .fpdpd import FPDPD_DTYPE
        n = self.num_beads() if n is None else n
        out = np.zeros(n, dtype=FPDPD_DTYPE)
        self._c(self.L.mdpd_download_fpdpd(self.h, out.ctypes.data_as(C.c_void_p), n,
                                           FPDPD_DTYPE.itemsize))
        return out

    def num_beads(self):
        n = C.c_uint32()
        self._c(self.L.mdpd_num_beads(self.h, C.byref(n)))
        return n.value

    def download(self):
        n = self.num_beads()
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "vel_buf", "acc", "press", "dens",
                                              "delta")}
        ui = {k: np.zeros(n, np.uint32) for k in ("id", "prop", "amp_id", "unit")}
        self._c(self.L.mdpd_download_soa(self.h, n, *[_d(out[k]) for k in (
            "pos", "vel", "vel_buf", "acc", "press", "dens", "delta")],
            *[_u(ui[k]) for k in ("id", "prop", "amp_id", "unit")]))
        out.update(ui)
        return out

    def set_topology(self, topo, core_amp_id=()):
        topo = _u32(topo)
        core = _u32(core_amp_id) if len(core_amp_id) else None
        self._c(self.L.mdpd_set_topology(self.h, _u(topo) if len(topo) else None,
                                         len(topo) // 4, _u(core), len(core_amp_id)))

    def make_config(self, spec):
        """Generates the configuration of `spec` (config_maker.py) on the
        device as this context's state, 4-bead amphiphiles as its chains;
        returns the kinetic temperature."""
        kt = C.c_double()
        self._c(self.L.mdpd_make_config(self.h, C.byref(spec), C.byref(kt)))
        return kt.value

    def set_clock(self, time, pair_seed):
        self._c(self.L.mdpd_set_clock(self.h, int(time), int(pair_seed)))

    def get_clock(self):
        t = C.c_uint32()
        self._c(self.L.mdpd_get_clock(self.h, C.byref(t)))
        return t.value

    def tuning(self):
        t = MdpdTuning()
        self._c(self.L.mdpd_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._c(self.L.mdpd_set_tuning(self.h, C.byref(t)))

    # ---- phases (main.cpp:167-196)
    def drift(self):
        self._c(self.L.mdpd_drift_predict(self.h))

    def rebuild(self):
        self._c(self.L.mdpd_rebuild_cells(self.h))

    def density(self):
        self._c(self.L.mdpd_density(self.h))

    def force(self):
        self._c(self.L.mdpd_force(self.h))

    def bonded(self):
        v = np.zeros(3)
        self._c(self.L.mdpd_bonded(self.h, _d(v)))
        return v

    def kick(self):
        self._c(self.L.mdpd_kick(self.h))

    def remove_cm_drift(self):
        self._c(self.L.mdpd_remove_cm_drift(self.h))

    def head_conversion(self):
        self._c(self.L.mdpd_head_conversion(self.h))

    def chem_event(self, p_thresld, eps, influ_rad, influ_hei, influ_dep, influ_grd,
                   max_amp_num, core_ptcl_id):
        ids = _u32(core_ptcl_id)
        cp = MdpdChemParams(p_thresld, eps, influ_rad, influ_hei, influ_dep, influ_grd,
                            max_amp_num, len(ids), _u(ids) if len(ids) else None)
        stop = C.c_int(0)
        self._c(self.L.mdpd_chem_event(self.h, C.byref(cp), C.byref(stop)))
        return stop.value

    def set_chemistry(self, p_thresld, eps, influ_rad, influ_hei, influ_dep, influ_grd,
                      max_amp_num, core_ptcl_id, beg_chem=0, enable=True):
        """RandomChemEvent inside mdpd_run for steps with time >= beg_chem."""
        ids = _u32(core_ptcl_id)
        self._chem_ids = ids
        cp = MdpdChemParams(p_thresld, eps, influ_rad, influ_hei, influ_dep, influ_grd,
                            max_amp_num, len(ids), _u(ids) if len(ids) else None)
        self._c(self.L.mdpd_set_chemistry(self.h, C.byref(cp), int(beg_chem), 1 if enable else 0))

    def set_chem_mt(self, seed=0, enable=True):
        """Reference-MT draws for the chemistry: one MT19937 stream from
        init_genrand(seed), consumed in the reference's order (ChemManager's
        PS::MT; the reference's single-process run seeds it with 0).
        enable=False returns to the keyed draws."""
        self._c(self.L.mdpd_set_chem_mt(self.h, int(seed) & 0xffffffff, 1 if enable else 0))

    def chem_mt_draws(self):
        """Outputs of the reference-MT stream consumed since set_chem_mt."""
        v = C.c_uint64()
        self._c(self.L.mdpd_chem_mt_draws(self.h, C.byref(v)))
        return v.value

    def chem_status(self):
        """(stopped, in-run events, amphiphiles inserted in runs)"""
        st, ev, ins = C.c_int32(), C.c_uint32(), C.c_uint32()
        self._c(self.L.mdpd_chem_status(self.h, C.byref(st), C.byref(ev), C.byref(ins)))
        return bool(st.value), ev.value, ins.value

    def run(self, nsteps):
        self._c(self.L.mdpd_run(self.h, _nsteps(nsteps)))

    def prepare_run(self, nsteps):
        """Records the step graphs the next run(nsteps) launches, so that
        no graph capture falls inside a timed run."""
        self._c(self.L.mdpd_prepare_run(self.h, _nsteps(nsteps)))

    def run_timed(self, nsteps):
        """Runs nsteps (captured graphs) and returns the device time in ms."""
        ms = C.c_double()
        self._c(self.L.mdpd_run_timed(self.h, _nsteps(nsteps), C.byref(ms)))
        return ms.value

    def skin_stats(self):
        """(steps so far that rebuilt cells and skin list, skin-list entries
        of the last build) in skin mode."""
        n, m = C.c_uint64(), C.c_uint64()
        self._c(self.L.mdpd_skin_stats(self.h, C.byref(n), C.byref(m)))
        return n.value, m.value

    def skin_builds(self):
        return self.skin_stats()[0]

    def sync(self):
        self._c(self.L.mdpd_sync(self.h))

    def observe(self):
        o = MdpdObservables()
        self._c(self.L.mdpd_observe(self.h, C.byref(o)))
        return observables_dict(o)

    def observe_replicas(self):
        arr = (MdpdObservables * self.n_rep)()
        self._c(self.L.mdpd_observe_replicas(self.h, arr))
        return [observables_dict(o) for o in arr]

    def density_profile(self, axis, nbins):
        out = np.zeros((self.n_rep, nbins))
        self._c(self.L.mdpd_density_profile(self.h, axis, nbins, _d(out)))
        return out

    def write_xyz_async(self, path, time, append=False):
        """Trajectory frame in the reference's xyz format, written by a host
        thread while the caller goes on; io_wait() waits for it."""
        self._c(self.L.mdpd_write_xyz_async(self.h, str(path).encode(), int(time), 1 if append else 0))

    def io_wait(self):
        self._c(self.L.mdpd_io_wait(self.h))

    def checkpoint_save(self, path):
        self._c(self.L.mdpd_checkpoint_save(self.h, str(path).encode()))

    def checkpoint_load(self, path):
        self._c(self.L.mdpd_checkpoint_load(self.h, str(path).encode()))

    def rdf(self, rmax, nbins):
        """Ordered-pair counts per distance bin [n_rep][nbins] (mdpd_rdf)."""
        out = np.zeros((self.n_rep, nbins))
        self._c(self.L.mdpd_rdf(self.h, float(rmax), int(nbins), _d(out)))
        return out

    def chain_histograms(self, nbins, bond_max):
        """[n_rep][4][nbins]: first bond, other bonds, bond-angle cosine,
        end-to-end (mdpd_chain_histograms)."""
        out = np.zeros((self.n_rep, 4, nbins))
        self._c(self.L.mdpd_chain_histograms(self.h, int(nbins), float(bond_max), _d(out)))
        return out

    def count_map(self, g, type_mask=0xFFFFFFFF):
        """Bead counts per voxel [n_rep][gz][gy][gx] (mdpd_count_map)."""
        out = np.zeros((self.n_rep, g[2], g[1], g[0]))
        self._c(self.L.mdpd_count_map(self.h, int(g[0]), int(g[1]), int(g[2]), int(type_mask),
                                      _d(out)))
        return out

    def pairs(self, cutoff=1.0):
        n = C.c_uint64()
        self._c(self.L.mdpd_neighbor_pairs(self.h, cutoff, None, None, 0, C.byref(n)))
        m = n.value
        oi = np.zeros(max(m, 1), np.uint32)
        oj = np.zeros(max(m, 1), np.uint32)
        self._c(self.L.mdpd_neighbor_pairs(self.h, cutoff, _u(oi), _u(oj), m, C.byref(n)))
        return oi[:m], oj[:m]

    # ---- measurement
    def set_profiling(self, on):
        self._c(self.L.mdpd_set_profiling(self.h, int(bool(on))))

    def phase_times(self):
        ms = np.zeros(32)
        n = C.c_int32()
        names = C.create_string_buffer(512)
        self._c(self.L.mdpd_phase_times(self.h, _d(ms), 32, C.byref(n), names, 512))
        raw = names.raw.split(b"\0")
        return {raw[k].decode(): float(ms[k]) for k in range(n.value)}

    def autotune(self, nsteps=5, skin=False):
        """Fastest block sizes (and, with skin=True, neighbour-list skin) on
        the current state; the state is left as it was."""
        t = MdpdTuning()
        flags = capi.MDPD_TUNE_SKIN if skin else 0
        self._c(self.L.mdpd_autotune_ex(self.h, int(nsteps), flags, C.byref(t)))
        return t


def saru_device(s1, s2, s3):
    s1, s2, s3 = _u32(s1), _u32(s2), _u32(s3)
    n = len(s1)
    u = np.zeros(8 * n, np.uint32)
    f = np.zeros(n, np.float32)
    check(capi.lib().mdpd_saru_device(_u(s1), _u(s2), _u(s3), n, _u(u),
                                      f.ctypes.data_as(C.POINTER(C.c_float))))
    return u.reshape(n, 8), f


def bandwidth_probe(nbytes, reps=20):
    g = C.c_double()
    check(capi.lib().mdpd_bandwidth_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def read_run_param(path):
    p = MdpdParams()
    ri = capi.MdpdRunInput()
    check(capi.lib().mdpd_read_run_param(str(path).encode(), C.byref(p), C.byref(ri)))
    return p, ri


def device_name():
    b = C.create_string_buffer(256)
    check(capi.lib().mdpd_device_name(b, 256))
    return b.value.decode()


def read_xyz(path):
    """io_util::ReadXYZForm through the library's parallel reader: returns
    (FPDPD records, time).  Raises MdpdError on the reference's abort cases."""
    from .fpdpd import FPDPD_DTYPE
    L = capi.lib()
    n, t = C.c_uint32(), C.c_uint32()
    p = str(path).encode()
    check(L.mdpd_read_xyz_header(p, C.byref(n), C.byref(t)))
    rec = np.zeros(n.value, dtype=FPDPD_DTYPE)
    buf = rec.ctypes.data_as(C.c_void_p) if n.value else None
    check(L.mdpd_read_xyz(p, buf, n.value, FPDPD_DTYPE.itemsize, C.byref(n), C.byref(t)))
    return rec, t.value
