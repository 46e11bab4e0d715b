"""GPU parity of the initial-configuration generator (src/config_maker.cpp)
against the oracle's sequential restatement, same spec and seed.

Bars: bead ids, types, amphiphile ids and units bit-exact; flat, asym_flat,
random and every solvent placement bit-exact (only +, *, /, sqrt and floor
in FP64 without contraction, correctly rounded on both sides); sphere and
cylinder positions within 1e-12 (device sin/cos of the ring angle against
the host's, a few ulps of a 30-long box), compared by minimum image since a
bead on a box face may wrap to the other face; velocities within 1e-12
(device log/cos/sqrt against the host's, and a compensated drift sum
against the reference's plain one); the kinetic temperature within 1e-12;
the reference's failures give MDPD_ERR_CONFIG on both sides.
"""
import numpy as np
import pytest

from implicit_dpd_amd import config_maker as cm
from implicit_dpd_amd import configs
from implicit_dpd_amd.capi import MDPD_ERR_CONFIG, MDPD_ERR_UNSUPPORTED, MdpdError

pytestmark = pytest.mark.gpu

BOX = (30.0, 30.0, 30.0)
CASES = [
    ("flat", dict(amp_num=8000, seed=5)),
    ("flat", dict(amp_num=4000, seed=6, with_solvent=True, sol_num=20000)),
    ("asym_flat", dict(amp_num=8000, seed=7, in_out_rat=0.29)),
    ("random", dict(amp_num=8000, seed=8)),
    ("sphere", dict(amp_num=7000, seed=2, sph_rad=9.0, upper_the=180.0)),
    ("sphere", dict(amp_num=6000, seed=3, sph_rad=9.0, upper_the=120.0, with_solvent=True, sol_num=4000)),
    ("cylind", dict(amp_num=7000, seed=3, cyl_r=6.0, cyl_l=30.0)),
    ("cylind", dict(amp_num=9500, seed=4, cyl_r=6.0, cyl_l=30.0, with_solvent=True, sol_num=3000)),
]


def _mi(d, L):
    L = np.asarray(L)
    return d - L * np.round(d / L)


def _compare(g, r, mode, box):
    for k in ("ids", "prop", "amp_id", "unit"):
        assert np.array_equal(g[k], r[k]), k
    if mode in ("sphere", "cylind"):
        na = int((g["unit"] != 0xFFFFFFFF).sum())
        assert np.abs(_mi(g["pos"][:na] - r["pos"][:na], box)).max() < 1e-12
        assert np.array_equal(g["pos"][na:], r["pos"][na:])  # solvent
    else:
        assert np.array_equal(g["pos"], r["pos"])
    assert np.abs(g["vel"] - r["vel"]).max() < 1e-12
    assert abs(g["kin_temp"] - r["kin_temp"]) < 1e-12


@pytest.mark.parametrize("mode,kw", CASES, ids=[f"{m}-{i}" for i, (m, _) in enumerate(CASES)])
def test_config_matches_oracle(oracle, mode, kw):
    s = cm.spec(mode, BOX, **kw)
    rc, r = oracle.make_config(s)
    assert rc == 0
    g = cm.make_config(s)
    _compare(g, r, mode, BOX)


def test_config_errors_match_oracle(oracle):
    cases = [
        cm.spec("sphere", BOX, 11284, sph_rad=9.0, upper_the=180.0, seed=3),  # one past the rings
        cm.spec("sphere", BOX, 10, sph_rad=15.5, upper_the=180.0, seed=3),    # radius > L/2
        cm.spec("cylind", BOX, 5220, cyl_r=6.0, cyl_l=30.0, seed=3),          # outer leaflet alone
        cm.spec("random", BOX, 100, with_solvent=True, sol_num=10),
        cm.spec("random", (10.0,) * 3, 4, seed=1),                            # |kT - 1| >= 1e-2
    ]
    for s in cases:
        rc, _ = oracle.make_config(s)
        assert rc == MDPD_ERR_CONFIG
        with pytest.raises(MdpdError) as ei:
            cm.make_config(s)
        assert ei.value.code == MDPD_ERR_CONFIG


def test_make_config_into_engine_and_run(engine_mod, oracle):
    """The context variant holds the same beads as the host variant and a
    4-bead chain topology: five steps of the membrane model from the
    generated bilayer track the oracle run from the same beads."""
    box = (24.0, 24.0, 16.0)
    # two leaflets at the reference's area per lipid, 0.333 (input_generator.py:69-76)
    n_amp = int(2 * box[0] * box[2] / 0.333)
    s = cm.spec("flat", box, n_amp, seed=12)
    p, _ = configs.membrane(n_chain=1, box=box)
    p.conv_p = 0.0
    n = cm.config_count(s)
    e = engine_mod.Engine(p, capacity=n + 100)
    e.set_tuning(nlist_width=160)  # the sheet is dense before it relaxes
    kt = e.make_config(s)
    g = e.download()
    h = cm.make_config(s)
    assert kt == h["kin_temp"]
    for k, kk in (("pos", "pos"), ("vel", "vel"), ("vel_buf", "vel"), ("id", "ids"), ("prop", "prop"),
                  ("amp_id", "amp_id"), ("unit", "unit")):
        assert np.array_equal(g[k], h[kk]), k
    assert (g["acc"] == 0).all() and (g["delta"] == 0).all()
    topo = np.arange(4 * n_amp, dtype=np.uint32)
    e.set_clock(1, p.pair_seed)
    o = oracle.System(p, n + 64, n_amp + 1)
    o.set(h["pos"], h["vel"], h["ids"], h["prop"], h["amp_id"], h["unit"])
    o.set_topology(topo)
    o.set_clock(1)
    e.run(5)
    o.run(5)
    gg, rr = e.download(), o.get()
    assert np.abs(_mi(gg["pos"] - rr["pos"], box)).max() < 1e-6
    go, ro = e.observe(), o.observe()
    assert np.allclose(go["kin_temp"], ro.kin_temp[:], rtol=1e-6)
    assert np.allclose(go["bonded_vir"], ro.bonded_vir[:], rtol=1e-5, atol=1e-6)
    # a second spec into the same context replaces the state
    s2 = cm.spec("random", box, 2000, seed=16)
    e.make_config(s2)
    assert e.num_beads() == 8000
    assert np.array_equal(e.download()["pos"], cm.make_config(s2)["pos"])
    e.run(2)
    # chains other than 4 beads are refused by the engine
    with pytest.raises(MdpdError) as ei:
        e.make_config(cm.spec("random", box, 100, head_unit=2, tail_unit=3))
    assert ei.value.code == MDPD_ERR_UNSUPPORTED
    e.close()


def test_membrane_bilayer_full_size(oracle):
    """262,144 H1T3 amphiphiles (1,048,576 beads) as a flat bilayer in the
    membrane box: the whole configuration against the oracle."""
    s = cm.spec("flat", (100.0, 100.0, 40.0), 262144, seed=21)
    rc, r = oracle.make_config(s)
    assert rc == 0
    g = cm.make_config(s)
    _compare(g, r, "flat", (100.0, 100.0, 40.0))
