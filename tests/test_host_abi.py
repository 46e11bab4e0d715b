"""Host-side pieces of the engine that need no GPU: the C ABI library loads
and exports every function include/mdpd_capi.h declares, the run_param.txt
parser follows parameter.hpp (stof rounding, missing-tag error, unknown tags
ignored), and the xyz format round-trips."""
import math
import os
import re

import numpy as np
import pytest

from implicit_dpd_amd import capi, configs
from implicit_dpd_amd.engine import read_run_param
from implicit_dpd_amd.fpdpd import FPDPD_DTYPE, read_xyz, write_xyz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(mdpd_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python prototypes cover the same set
    assert set(names) <= set(capi.PROTOTYPES) | {"mdpd_run_input"}


def test_parser_matches_oracle_coefficients(oracle):
    # the reference's own test input (test/mpi_bonded/result_mpi/31x31/run_param.txt)
    p, ri = read_run_param(os.path.join(GOLD, "mpi_bonded_31x31"))
    assert tuple(p.box) == (31.0, 30.0, 31.0)
    assert ri.init_amp_num == 9600 and ri.sol_num == 0 and ri.max_amp_num == 10000
    assert ri.n_core == 3 and list(ri.core_amp_id[:3]) == [0, 1, 10]
    f32 = lambda x: float(np.float32(x))
    q = oracle.calc_inter_coef(f32(30.0), f32(50.0), f32(25.0), 0.1, [f32(3.0)] * 6, f32(0.005),
                               reo=1.4)
    for a, b in (("c1", 2), ("sigma", 2), ("gamma", 2), ("mb", 3)):
        x = np.array(getattr(p, a)) if b == 2 else np.array([[list(r) for r in m] for m in p.mb])
        y = np.array(getattr(q, a)) if b == 2 else np.array([[list(r) for r in m] for m in q.mb])
        assert np.array_equal(np.asarray(x, float), np.asarray(y, float)), a
    assert p.dt == f32(0.005) and p.dt != 0.005  # std::stof rounding (parameter.hpp:339)
    assert p.cf_s == 200.0 and p.cf_b == 4.0 and p.cf_b_rigid == 40.0
    # influ_vol from influ_rad/influ_grd (CalcInfluVol, parameter.hpp:58-72)
    up = ri.influ_rad - ri.influ_grd * ri.influ_hei
    assert math.isclose(ri.influ_vol, math.pi * up * up * (ri.influ_hei + ri.influ_dep), rel_tol=1e-12)


def test_parser_missing_tag_and_comments(tmp_path):
    d = tmp_path / "run"
    configs.write_run_param(str(d), (10, 10, 10), 5)
    txt = (d / "run_param.txt").read_text()
    (d / "run_param.txt").write_text("; a comment line\n   " + txt.replace("kappa 50.0", "kappa 50.0 ; trailing\nunknown_tag 1 2"))
    p, ri = read_run_param(str(d))
    assert ri.kappa == 50.0
    (d / "run_param.txt").write_text(txt.replace("chi 30.0\n", ""))
    with pytest.raises(capi.MdpdError) as e:
        read_run_param(str(d))
    assert e.value.code == capi.MDPD_ERR_PARSE and "chi" in str(e.value)


def test_influ_vol_branch(tmp_path):
    d = tmp_path / "run"
    configs.write_run_param(str(d), (10, 10, 10), 5)
    txt = (d / "run_param.txt").read_text().replace("influ_rad 5.0\n", "influ_vol 162\nupside_rad 3.9\n")
    (d / "run_param.txt").write_text(txt.replace("influ_grd 0.0\n", ""))
    p, ri = read_run_param(str(d))
    hs = ri.influ_hei + ri.influ_dep
    v = 12.0 * ri.influ_vol / (math.pi * hs) - 3.0 * ri.upside_rad ** 2
    down = (math.sqrt(v) - ri.upside_rad) * 0.5
    assert math.isclose(ri.influ_grd, (down - ri.upside_rad) / hs, rel_tol=1e-12)


def test_xyz_roundtrip(tmp_path):
    s = configs.random_chains(10, (5.0, 5.0, 5.0), 1, n_solvent=3)
    rec = np.zeros(43, FPDPD_DTYPE)
    rec["pos"], rec["vel"], rec["vel_buf"] = s["pos"], s["vel"], s["vel"]
    rec["id"], rec["prop"], rec["amp_id"], rec["unit"] = s["ids"], s["prop"], s["amp_id"], s["unit"]
    write_xyz(tmp_path / "c.xyz", rec, 17)
    r2, t = read_xyz(tmp_path / "c.xyz")
    assert t == 17
    for k in ("pos", "vel", "vel_buf", "acc"):
        assert np.allclose(r2[k], rec[k], rtol=1e-14, atol=0)
    for k in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(r2[k], rec[k])


def test_driver_binary_links_and_checks_arguments():
    """The stand-alone driver (C++ facade over the ABI) is built in-tree and
    refuses a missing target directory like main.cpp:78-82."""
    import subprocess
    drv = os.path.join(ROOT, "implicit_dpd_amd", "bin", "mdpd_main")
    assert os.path.exists(drv)
    r = subprocess.run([drv], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "argv[1] = target directory." in r.stderr
