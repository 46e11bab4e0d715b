"""Initial configurations: the role of the reference's config_maker program
(src/config_maker.cpp), generated on the GPU.

``read_config_param(dir)`` reads config_param.txt like ConfigMaker::LoadParam
(:455-467); ``make_config(spec)`` generates the beads on the device and
returns them in the reference's bead order; ``Engine.make_config(spec)``
generates straight into an engine context.  ``python -m
implicit_dpd_amd.config_maker DIR [--seed S]`` is the reference's
``config_maker DIR``: it writes DIR/init_config.xyz (io_util::WriteXYZForm)
for the run driver to read.

Modes (GenConfig :342-384): flat, asym_flat, sphere, cylind, random.  The
reference seeds MT19937 with time(NULL); here each amphiphile, solvent bead
and velocity has a keyed Saru stream, so a seed gives one configuration.
"""
import argparse
import ctypes as C
import os
import sys

import numpy as np

from . import capi
from .capi import MdpdConfigSpec, check

MODES = {"flat": capi.CFG_FLAT, "asym_flat": capi.CFG_ASYM_FLAT, "sphere": capi.CFG_SPHERE,
         "cylind": capi.CFG_CYLIND, "random": capi.CFG_RANDOM}

_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)


def spec(mode, box, amp_num, temperature=1.0, head_unit=1, tail_unit=3, bond_leng=0.5,
         lip_len=0.5, sol_num=0, with_solvent=False, sph_rad=0.0, upper_the=0.0, cyl_l=0.0,
         cyl_r=0.0, in_out_rat=0.0, seed=1):
    """A config spec with the tags of config_param.txt (MatchingTagValues
    :386-418); `mode` is a name of MODES or a CFG_* value."""
    s = MdpdConfigSpec()
    s.mode = MODES[mode.lower()] if isinstance(mode, str) else int(mode)
    s.with_solvent = 1 if with_solvent else 0
    for a in range(3):
        s.box[a] = float(box[a])
    s.temperature = float(temperature)
    s.amp_num = int(amp_num)
    s.head_unit = int(head_unit)
    s.tail_unit = int(tail_unit)
    s.sol_num = int(sol_num)
    s.bond_leng = float(bond_leng)
    s.lip_len = float(lip_len)
    s.sph_rad = float(sph_rad)
    s.upper_the = float(upper_the)
    s.cyl_l = float(cyl_l)
    s.cyl_r = float(cyl_r)
    s.in_out_rat = float(in_out_rat)
    s.seed = int(seed) & 0xFFFFFFFF
    return s


def read_config_param(dirname, seed=1):
    s = MdpdConfigSpec()
    check(capi.lib().mdpd_read_config_param(str(dirname).encode(), C.byref(s)))
    s.seed = int(seed) & 0xFFFFFFFF
    return s


def config_count(s):
    n = C.c_uint32()
    check(capi.lib().mdpd_config_count(C.byref(s), C.byref(n)))
    return n.value


def make_config(s):
    """Beads of spec `s` made on the device: dict of pos, vel [n, 3] FP64 and
    ids, prop, amp_id, unit; 'kin_temp' = sum v.v / 3n."""
    n = config_count(s)
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    u = [np.zeros(n, np.uint32) for _ in range(4)]
    kt = C.c_double()
    check(capi.lib().mdpd_make_config_host(
        C.byref(s), n, pos.ctypes.data_as(_DP), vel.ctypes.data_as(_DP),
        *[a.ctypes.data_as(_U32P) for a in u], C.byref(kt)))
    return dict(pos=pos, vel=vel, ids=u[0], prop=u[1], amp_id=u[2], unit=u[3], kin_temp=kt.value)


def to_fpdpd(cfg):
    """FPDPD records (vel_buf = vel, acc = 0, GenVeloc :67-70)."""
    from .fpdpd import FPDPD_DTYPE
    rec = np.zeros(len(cfg["pos"]), FPDPD_DTYPE)
    rec["pos"] = cfg["pos"]
    rec["vel"] = cfg["vel"]
    rec["vel_buf"] = cfg["vel"]
    rec["id"] = cfg["ids"]
    rec["prop"] = cfg["prop"]
    rec["amp_id"] = cfg["amp_id"]
    rec["unit"] = cfg["unit"]
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description="config_maker: DIR/config_param.txt -> DIR/init_config.xyz")
    ap.add_argument("dir")
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    from .fpdpd import write_xyz
    s = read_config_param(a.dir, a.seed)
    cfg = make_config(s)
    write_xyz(os.path.join(a.dir, "init_config.xyz"), to_fpdpd(cfg), 0)
    print(f"Initial configuration is generated at {a.dir}")
    inv = {v: k for k, v in MODES.items()}
    print(f"Temperature {s.temperature}\nbox_leng {s.box[0]} {s.box[1]} {s.box[2]}\n"
          f"amp_num {s.amp_num}\nhead_unit {s.head_unit}\ntail_unit {s.tail_unit}\n"
          f"all_unit {s.head_unit + s.tail_unit}\nbond_leng {s.bond_leng}\nmode {inv[s.mode]}\n"
          f"kinetic temperature {cfg['kin_temp']:.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
