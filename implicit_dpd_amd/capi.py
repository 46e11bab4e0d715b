"""ctypes view of the engine's C ABI (include/mdpd_capi.h).

The engine is the HIP shared library ``implicit_dpd_amd/libmdpd.so`` built by
``implicit_dpd_amd/build.py``.  There is no fallback: if the library is not
built, ``lib()`` raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdpd.so")

# status codes (mdpd_capi.h)
MDPD_OK = 0
MDPD_ERR_ARG = 1
MDPD_ERR_HIP = 2
MDPD_ERR_CAPACITY = 3
MDPD_ERR_NONFINITE = 4
MDPD_ERR_OUT_OF_BOX = 5
MDPD_ERR_NLIST_OVERFLOW = 6
MDPD_ERR_PARSE = 7
MDPD_ERR_STATE = 8
MDPD_ERR_NO_DEVICE = 9
MDPD_ERR_UNSUPPORTED = 10
MDPD_ERR_CONFIG = 11

# initial-configuration modes (config_maker.cpp GenConfig)
CFG_FLAT, CFG_ASYM_FLAT, CFG_SPHERE, CFG_CYLIND, CFG_RANDOM = 0, 1, 2, 3, 4

HYPHIL, HYPHOB, SOLVENT = 0, 1, 2
FPDPD_STRIDE = 232
UNIT_NONE = 0xFFFFFFFF
MDPD_TUNE_SKIN = 1

D33 = (C.c_double * 3) * 3
D333 = ((C.c_double * 3) * 3) * 3


class MdpdParams(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3),
        ("dt", C.c_double),
        ("n_species", C.c_int32),
        ("n_dens", C.c_int32),
        ("envelope", C.c_int32),
        ("pad0", C.c_int32),
        ("rd", C.c_double),
        ("crho", C.c_double),
        ("rdm", C.c_double),
        ("arc", C.c_double),
        ("a0", D33),
        ("c1", D33),
        ("mb", D333),
        ("sigma", D33),
        ("gamma", D33),
        ("cf_s", C.c_double),
        ("cf_b", C.c_double),
        ("cf_b_rigid", C.c_double),
        ("bond_leng", C.c_double),
        ("pair_seed", C.c_uint32),
        ("chem_seed", C.c_uint32),
        ("conv_p", C.c_double),
        ("conv_from", C.c_int32),
        ("conv_to", C.c_int32),
        ("cm_corr", C.c_uint32),
        ("pad1", C.c_uint32),
    ]


class MdpdChemParams(C.Structure):
    _fields_ = [
        ("p_thresld", C.c_double),
        ("eps", C.c_double),
        ("influ_rad", C.c_double),
        ("influ_hei", C.c_double),
        ("influ_dep", C.c_double),
        ("influ_grd", C.c_double),
        ("max_amp_num", C.c_uint32),
        ("n_core", C.c_uint32),
        ("core_ptcl_id", C.POINTER(C.c_uint32)),
    ]


class MdpdObservables(C.Structure):
    _fields_ = [
        ("kin_temp", C.c_double * 4),
        ("pressure", C.c_double * 3),
        ("diffusion", C.c_double * 3),
        ("bonded_vir", C.c_double * 3),
        ("pair_vir", C.c_double * 3),
        ("n_pairs", C.c_double),
        ("n_beads", C.c_uint32),
        ("n_amp", C.c_uint32),
        ("time", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class MdpdTuning(C.Structure):
    _fields_ = [
        ("block_scan", C.c_int32),
        ("block_force", C.c_int32),
        ("cell_div", C.c_int32),
        ("nlist_width", C.c_int32),
        ("reorder_every", C.c_int32),
        ("use_graph", C.c_int32),
        ("verlet_skin", C.c_float),
        ("verlet_width", C.c_int32),
    ]


class MdpdRunInput(C.Structure):
    _fields_ = [
        ("init_amp_num", C.c_uint32),
        ("sol_num", C.c_uint32),
        ("all_time", C.c_uint32),
        ("step_mic", C.c_uint32),
        ("step_mac", C.c_uint32),
        ("max_amp_num", C.c_uint32),
        ("p_thresld", C.c_double),
        ("eps", C.c_double),
        ("influ_rad", C.c_double),
        ("influ_hei", C.c_double),
        ("influ_dep", C.c_double),
        ("influ_grd", C.c_double),
        ("influ_vol", C.c_double),
        ("upside_rad", C.c_double),
        ("chi", C.c_double),
        ("kappa", C.c_double),
        ("rho_co", C.c_double),
        ("vbb", C.c_double),
        ("cf_r_in", C.c_double * 6),
        ("n_core", C.c_uint32),
        ("core_amp_id", C.c_uint32 * 64),
    ]


class MdpdConfigSpec(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("with_solvent", C.c_int32),
        ("box", C.c_double * 3),
        ("temperature", C.c_double),
        ("amp_num", C.c_uint32),
        ("head_unit", C.c_uint32),
        ("tail_unit", C.c_uint32),
        ("sol_num", C.c_uint32),
        ("bond_leng", C.c_double),
        ("lip_len", C.c_double),
        ("sph_rad", C.c_double),
        ("upper_the", C.c_double),
        ("cyl_l", C.c_double),
        ("cyl_r", C.c_double),
        ("in_out_rat", C.c_double),
        ("seed", C.c_uint32),
        ("pad", C.c_uint32),
    ]


P = C.POINTER
VP = C.c_void_p
_U32P = P(C.c_uint32)
_DP = P(C.c_double)
_FP = P(C.c_float)

# name -> (restype, argtypes)
PROTOTYPES = {
    "mdpd_create": (C.c_int, [P(MdpdParams), C.c_uint32, P(VP)]),
    "mdpd_create_ensemble": (C.c_int, [P(MdpdParams), _U32P, C.c_int32, P(VP)]),
    "mdpd_destroy": (C.c_int, [VP]),
    "mdpd_last_error": (C.c_char_p, [VP]),
    "mdpd_global_error": (C.c_char_p, []),
    "mdpd_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "mdpd_upload_fpdpd": (C.c_int, [VP, VP, C.c_uint32, C.c_size_t]),
    "mdpd_download_fpdpd": (C.c_int, [VP, VP, C.c_uint32, C.c_size_t]),
    "mdpd_upload_soa": (C.c_int, [VP, C.c_uint32, _DP, _DP, _DP, _U32P, _U32P, _U32P, _U32P]),
    "mdpd_download_soa": (C.c_int, [VP, C.c_uint32, _DP, _DP, _DP, _DP, _DP, _DP, _DP,
                                    _U32P, _U32P, _U32P, _U32P]),
    "mdpd_num_beads": (C.c_int, [VP, _U32P]),
    "mdpd_set_topology": (C.c_int, [VP, _U32P, C.c_uint32, _U32P, C.c_uint32]),
    "mdpd_set_clock": (C.c_int, [VP, C.c_uint32, C.c_uint32]),
    "mdpd_get_clock": (C.c_int, [VP, _U32P]),
    "mdpd_set_tuning": (C.c_int, [VP, P(MdpdTuning)]),
    "mdpd_get_tuning": (C.c_int, [VP, P(MdpdTuning)]),
    "mdpd_drift_predict": (C.c_int, [VP]),
    "mdpd_rebuild_cells": (C.c_int, [VP]),
    "mdpd_density": (C.c_int, [VP]),
    "mdpd_force": (C.c_int, [VP]),
    "mdpd_bonded": (C.c_int, [VP, _DP]),
    "mdpd_kick": (C.c_int, [VP]),
    "mdpd_remove_cm_drift": (C.c_int, [VP]),
    "mdpd_head_conversion": (C.c_int, [VP]),
    "mdpd_chem_event": (C.c_int, [VP, P(MdpdChemParams), P(C.c_int)]),
    "mdpd_observe": (C.c_int, [VP, P(MdpdObservables)]),
    "mdpd_observe_replicas": (C.c_int, [VP, P(MdpdObservables)]),
    "mdpd_density_profile": (C.c_int, [VP, C.c_int32, C.c_int32, _DP]),
    "mdpd_set_chemistry": (C.c_int, [VP, C.POINTER(MdpdChemParams), C.c_uint32, C.c_int32]),
    "mdpd_chem_status": (C.c_int, [VP, C.POINTER(C.c_int32), _U32P, _U32P]),
    "mdpd_set_chem_mt": (C.c_int, [VP, C.c_uint32, C.c_int32]),
    "mdpd_chem_mt_draws": (C.c_int, [VP, P(C.c_uint64)]),
    "mdpd_rdf": (C.c_int, [VP, C.c_double, C.c_int32, _DP]),
    "mdpd_rdf_noalloc": (C.c_int, [VP, C.c_double, C.c_int32, _DP]),
    "mdpd_write_xyz_async": (C.c_int, [VP, C.c_char_p, C.c_uint32, C.c_int32]),
    "mdpd_io_wait": (C.c_int, [VP]),
    "mdpd_read_xyz_header": (C.c_int, [C.c_char_p, _U32P, _U32P]),
    "mdpd_read_xyz": (C.c_int, [C.c_char_p, VP, C.c_uint32, C.c_size_t, _U32P, _U32P]),
    "mdpd_checkpoint_save": (C.c_int, [VP, C.c_char_p]),
    "mdpd_checkpoint_load": (C.c_int, [VP, C.c_char_p]),
    "mdpd_chain_histograms": (C.c_int, [VP, C.c_int32, C.c_double, _DP]),
    "mdpd_count_map": (C.c_int, [VP, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _DP]),
    "mdpd_run": (C.c_int, [VP, C.c_uint32]),
    "mdpd_sync": (C.c_int, [VP]),
    "mdpd_neighbor_pairs": (C.c_int, [VP, C.c_double, _U32P, _U32P, C.c_uint64, P(C.c_uint64)]),
    "mdpd_saru_device": (C.c_int, [_U32P, _U32P, _U32P, C.c_uint32, _U32P, _FP]),
    "mdpd_set_profiling": (C.c_int, [VP, C.c_int32]),
    "mdpd_phase_times": (C.c_int, [VP, _DP, C.c_int32, P(C.c_int32), C.c_char_p, C.c_size_t]),
    "mdpd_autotune": (C.c_int, [VP, C.c_uint32, P(MdpdTuning)]),
    "mdpd_autotune_ex": (C.c_int, [VP, C.c_uint32, C.c_uint32, P(MdpdTuning)]),
    "mdpd_bandwidth_probe": (C.c_int, [C.c_uint64, C.c_int32, P(C.c_double)]),
    "mdpd_read_run_param": (C.c_int, [C.c_char_p, P(MdpdParams), P(MdpdRunInput)]),
    "mdpd_run_timed": (C.c_int, [VP, C.c_uint32, _DP]),
    "mdpd_prepare_run": (C.c_int, [VP, C.c_uint32]),
    # initial configurations
    "mdpd_read_config_param": (C.c_int, [C.c_char_p, P(MdpdConfigSpec)]),
    "mdpd_config_count": (C.c_int, [P(MdpdConfigSpec), _U32P]),
    "mdpd_make_config_host": (C.c_int, [P(MdpdConfigSpec), C.c_uint32, _DP, _DP, _U32P, _U32P,
                                        _U32P, _U32P, _DP]),
    "mdpd_make_config": (C.c_int, [VP, P(MdpdConfigSpec), _DP]),
    "mdpd_skin_stats": (C.c_int, [VP, P(C.c_uint64), P(C.c_uint64)]),
    # equation-of-state surrogate
    "mdpd_mlp_num_params": (C.c_int, [_U32P]),
    "mdpd_mlp_create": (C.c_int, [C.c_uint32, C.c_uint32, P(VP)]),
    "mdpd_mlp_destroy": (C.c_int, [VP]),
    "mdpd_mlp_last_error": (C.c_char_p, [VP]),
    "mdpd_mlp_set_params": (C.c_int, [VP, _FP, C.c_uint32]),
    "mdpd_mlp_get_params": (C.c_int, [VP, _FP, C.c_uint32]),
    "mdpd_mlp_forward": (C.c_int, [VP, _FP, C.c_uint32, _FP]),
    "mdpd_mlp_loss_grad": (C.c_int, [VP, _FP, _FP, C.c_uint32, _DP, _FP]),
    "mdpd_mlp_train": (C.c_int, [VP, _FP, _FP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                 C.c_uint32, _FP]),
}

_lib = None


def lib():
    """Load libmdpd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmdpd.so is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MdpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"MDPD error {code}: {msg}")
        self.code = code


def check(code, ctx=None):
    if code != MDPD_OK:
        L = lib()
        if ctx is not None:
            msg = L.mdpd_last_error(ctx)
        else:
            msg = L.mdpd_global_error()
        raise MdpdError(code, msg.decode() if msg else "")
    return code
