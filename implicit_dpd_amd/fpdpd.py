"""The reference's host particle record and xyz text format.

FPDPD (include/ptcl_class.hpp:29-36) is 232 bytes: u32 id, prop, amp_id,
unit; F64vec pos, delta_sumr, vel, vel_buf, acc, press, nei_cm_pos[2];
F64 density[3].  The xyz format is the 17-column ASCII of
FPDPD::writeAscii / readAscii (ptcl_class.hpp:60-76) under the two-line
header of io_util::WriteXYZForm (io_util.hpp:25-31).
"""
import numpy as np

FPDPD_DTYPE = np.dtype([
    ("id", "<u4"), ("prop", "<u4"), ("amp_id", "<u4"), ("unit", "<u4"),
    ("pos", "<f8", 3), ("delta_sumr", "<f8", 3), ("vel", "<f8", 3), ("vel_buf", "<f8", 3),
    ("acc", "<f8", 3), ("press", "<f8", 3), ("nei_cm_pos", "<f8", (2, 3)),
    ("density", "<f8", 3),
])
assert FPDPD_DTYPE.itemsize == 232

# Parameter::atom_type (parameter.hpp:207-209)
ATOM_TYPE = "ONCSPZXONCSPZXONCSPZX"


def write_xyz(path, rec, time):
    """io_util::WriteXYZForm + FPDPD::writeAscii (%.15g columns)."""
    with open(path, "w") as f:
        f.write(f"{len(rec)}\n")
        f.write(f"time {time}\n")
        for r in rec:
            cols = [ATOM_TYPE[int(r["prop"])]]
            cols += ["%.15g" % x for x in r["pos"]]
            cols += ["%u" % r[k] for k in ("id", "prop", "amp_id", "unit")]
            cols += ["%.15g" % x for x in r["vel"]]
            cols += ["%.15g" % x for x in r["vel_buf"]]
            cols += ["%.15g" % x for x in r["acc"]]
            f.write(" ".join(cols) + "\n")


def read_xyz(path):
    """io_util::ReadXYZForm: returns (records, time)."""
    with open(path) as f:
        n = int(f.readline())
        head = f.readline().split()
        if len(head) < 2 or head[0] != "time":
            raise ValueError("The file format of init_config.xyz may be old.")
        t = int(head[1])
        rec = np.zeros(n, dtype=FPDPD_DTYPE)
        for i in range(n):
            c = f.readline().split()
            if len(c) < 17:
                raise ValueError("# of lines is not equal to the information specified in file header.")
            rec["pos"][i] = [float(x) for x in c[1:4]]
            rec["id"][i], rec["prop"][i], rec["amp_id"][i], rec["unit"][i] = (int(x) for x in c[4:8])
            rec["vel"][i] = [float(x) for x in c[8:11]]
            rec["vel_buf"][i] = [float(x) for x in c[11:14]]
            rec["acc"][i] = [float(x) for x in c[14:17]]
        rest = f.read().split()
        if rest:
            raise ValueError("# of lines is not equal to the information specified in file header.")
    return rec, t
