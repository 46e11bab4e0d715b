"""Pins of the oracle's initial-configuration restatement (src/config_maker.cpp).

What fixes each mode independently of the oracle's own code:
  * the keyed draws: a numpy Saru, itself checked against the u32 vectors of
    the reference's own saruprng.hpp (tests/golden/saru_ref_vectors.txt);
  * velocities: Box-Muller over those uniforms, minus their mean (GenVeloc
    :60-72 with RemoveCMDrift :51-58);
  * flat: chains along -+y from heads at 0.5 Ly +- (3 b + 1e-5), leaflets
    alternating (MakeFlatSheet :103-138); asym_flat: each block of 100
    amphiphiles puts its first (int)(100 in_out_rat) up;
  * random: straight chains of bond b along a positive-octant direction;
  * sphere / cylinder: every bead at the radius its leaflet and unit fix
    (head at R +- 2q, q = all_unit b / 2, stepping b inward / outward), ring
    counts floor(2 pi r sin(theta) / lip_len) per ring in the reference's
    ring order, and the reference's failures (too many amphiphiles, radius
    too large, the cylinder's outer leaflet filled alone);
  * solvent: inside the region its mode names;
  * config_param.txt parsing: a file of the reference's test tree.
"""
import math
import os

import numpy as np
import pytest

from implicit_dpd_amd import config_maker as cm
from implicit_dpd_amd.capi import HYPHIL, HYPHOB, MDPD_ERR_CONFIG, SOLVENT, UNIT_NONE

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TAG_AMP, TAG_SOL, TAG_VEL = 0x414D5048, 0x534F4C56, 0x56454C4F
M32 = np.uint64(0xFFFFFFFF)


# ---------------------------------------------------------------- numpy Saru
def _asr(x, s):
    return (x.astype(np.uint32).view(np.int32) >> s).view(np.uint32).astype(np.uint64)


def _m(x):
    return x & M32


def saru_np(s1, s2, s3):
    """Saru(seed1, seed2, seed3) of saruprng.hpp:208-255 on uint64 lanes."""
    s1, s2, s3 = (np.asarray(v, np.uint64) & M32 for v in (s1, s2, s3))
    s3 = _m(s3 ^ _m(s1 << np.uint64(7)) ^ (s2 >> np.uint64(6)))
    s2 = _m(s2 + ((s1 >> np.uint64(4)) ^ (s3 >> np.uint64(15))))
    s1 = _m(s1 ^ _m(_m(s2 << np.uint64(9)) + _m(s3 << np.uint64(8))))
    s3 = _m(s3 ^ _m(np.uint64(0xA5366B4D) * ((s2 >> np.uint64(11)) ^ _m(s1 << np.uint64(1)))))
    s2 = _m(s2 + _m(np.uint64(0x72BE1579) * (_m(s1 << np.uint64(4)) ^ (s3 >> np.uint64(16)))))
    s1 = _m(s1 ^ _m(np.uint64(0x3F38A6ED) * ((s3 >> np.uint64(5)) ^ _asr(s2, 22))))
    s2 = _m(s2 + _m(s1 * s3))
    s1 = _m(s1 + (s3 ^ (s2 >> np.uint64(2))))
    s2 = _m(s2 ^ _asr(s2, 17))
    st = _m(np.uint64(0x79dedea3) * (s1 ^ _asr(s1, 14)))
    ws = _m((st + s2) ^ _asr(st, 8))
    st = _m(st + _m(ws * (ws ^ np.uint64(0xdddf97f5))))
    ws = _m(np.uint64(0xABCB96F7) + (ws >> np.uint64(1)))
    return [st, ws]


def saru_next(s):
    st, ws = s
    st = _m(np.uint64(0x4beb5d59) * st + np.uint64(0x2600e1f7))
    ws = _m(ws + np.uint64(0x8009d14b) + (_asr(ws, 31) & np.uint64(0xda879add)))
    v = _m((st ^ (st >> np.uint64(26))) + ws)
    s[0], s[1] = st, ws
    return _m((v ^ (v >> np.uint64(20))) * np.uint64(0x6957f5a7))


def uni_np(s, open_=True):
    a = saru_next(s) >> np.uint64(6)
    b = saru_next(s) >> np.uint64(6)
    k = a.astype(np.float64) * 67108864.0 + b.astype(np.float64)
    return (k + 0.5 if open_ else k) / 4503599627370496.0


def test_numpy_saru_matches_reference_vectors():
    rows = [l.split() for l in open(os.path.join(GOLD, "saru_ref_vectors.txt")) if l.startswith("U ")]
    seeds = np.array([[int(x) for x in r[1:4]] for r in rows], np.uint64)
    want = np.array([[int(x) for x in r[4:12]] for r in rows], np.uint64)
    s = saru_np(seeds[:, 0], seeds[:, 1], seeds[:, 2])
    got = np.stack([saru_next(s) for _ in range(8)], axis=1)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- helpers
def _run(oracle, **kw):
    mode = kw.pop("mode")
    box = kw.pop("box")
    amp = kw.pop("amp_num")
    rc, d = oracle.make_config(cm.spec(mode, box, amp, **kw))
    return rc, d


def _mi(d, L):
    L = np.asarray(L)
    return d - L * np.round(d / L)


def _chains(d, n_amp, all_unit=4):
    return d["pos"][: n_amp * all_unit].reshape(n_amp, all_unit, 3)


def _expected_vel(n, seed, T=1.0):
    s = saru_np(np.arange(n), TAG_VEL, seed)
    v = np.zeros((n, 3))
    for c in range(3):
        u1 = uni_np(s)
        u2 = uni_np(s)
        v[:, c] = math.sqrt(T) * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return v - v.mean(axis=0)


def _common_checks(d, n_amp, box, all_unit=4, head_unit=1):
    n = len(d["pos"])
    na = n_amp * all_unit
    assert np.array_equal(d["ids"], np.arange(n))
    u = np.arange(na) % all_unit
    assert np.array_equal(d["unit"][:na], u)
    assert np.array_equal(d["amp_id"][:na], np.arange(na) // all_unit)
    assert np.array_equal(d["prop"][:na], np.where(u < head_unit, HYPHIL, HYPHOB))
    assert (d["prop"][na:] == SOLVENT).all()
    assert (d["unit"][na:] == UNIT_NONE).all() and (d["amp_id"][na:] == UNIT_NONE).all()
    assert ((d["pos"] >= 0.0) & (d["pos"] <= np.asarray(box))).all()


# ---------------------------------------------------------------- modes
def test_flat_bilayer_geometry_and_velocities(oracle):
    box, n_amp, b, seed = (30.0, 24.0, 20.0), 6000, 0.5, 7
    rc, d = _run(oracle, mode="flat", box=box, amp_num=n_amp, seed=seed)
    assert rc == 0
    _common_checks(d, n_amp, box)
    ch = _chains(d, n_amp)
    # x, z: the amphiphile's uniform base point, the same for its beads
    assert (ch[:, :, [0, 2]] == ch[:, :1, [0, 2]]).all()
    sa = saru_np(np.arange(n_amp), TAG_AMP, seed)
    ux, _, uz = uni_np(sa), uni_np(sa), uni_np(sa)
    assert np.array_equal(ch[:, 0, 0], box[0] * ux) and np.array_equal(ch[:, 0, 2], box[2] * uz)
    # heads at 0.5 Ly +- (3 b + 1e-5), up for even amphiphiles; tails toward the mid plane
    sign = np.where(np.arange(n_amp) % 2 == 0, 1.0, -1.0)
    assert np.abs(ch[:, 0, 1] - (0.5 * box[1] + (3 * b + 1e-5) * sign)).max() < 1e-12
    dy = np.diff(ch[:, :, 1], axis=1)
    assert np.abs(dy + sign[:, None] * b).max() < 1e-12
    assert np.abs(d["vel"] - _expected_vel(len(d["pos"]), seed)).max() < 1e-12
    assert np.abs(d["vel"].mean(axis=0)).max() < 1e-15
    assert abs(d["kin_temp"] - (d["vel"] ** 2).sum() / (3 * len(d["pos"]))) < 1e-12


@pytest.mark.parametrize("ratio,rat", [(0.3, 30), (0.29, 28), (0.0, 0), (0.5, 50)])
def test_asym_flat_leaflet_split_per_hundred(oracle, ratio, rat):
    n_amp = 1000
    rc, d = _run(oracle, mode="asym_flat", box=(30.0, 30.0, 20.0), amp_num=n_amp, in_out_rat=ratio,
                 seed=3)
    assert rc == 0
    up = _chains(d, n_amp)[:, 0, 1] > 15.0
    a = np.arange(n_amp)
    if rat:
        # (int)(100 * 0.29) = 28: the reference truncates the product
        assert np.array_equal(up, a % 100 < rat)
    else:
        # only the multiples of 100 toggle: whole blocks alternate
        assert np.array_equal(up, (a // 100) % 2 == 0)


def test_random_chains_are_straight_in_positive_octant(oracle):
    box, n_amp, seed = (16.0, 16.0, 16.0), 4096, 11
    rc, d = _run(oracle, mode="random", box=box, amp_num=n_amp, seed=seed)
    assert rc == 0
    _common_checks(d, n_amp, box)
    ch = _chains(d, n_amp)
    bond = _mi(np.diff(ch, axis=1), box)
    assert np.abs(np.linalg.norm(bond, axis=2) - 0.5).max() < 1e-12
    assert (bond > -1e-12).all()
    assert np.abs(bond - bond[:, :1]).max() < 1e-12
    sa = saru_np(np.arange(n_amp), TAG_AMP, seed)
    base = np.stack([box[c] * uni_np(sa, False) for c in range(3)], axis=1)
    nv = np.stack([uni_np(sa, False) for _ in range(3)], axis=1)
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    assert np.abs(_mi(ch[:, 0] - base, box)).max() < 1e-12
    assert np.abs(bond[:, 0] - 0.5 * nv).max() < 1e-12


def _ring_counts(rad, thetas, lip):
    return [max(int(math.floor(2 * math.pi * rad * math.sin(t) / lip + 1e-9)), 0) for t in thetas]


def test_sphere_radii_rings_and_solvent(oracle):
    box, R, b, lip, up = (30.0, 30.0, 30.0), 9.0, 0.5, 0.5, 180.0
    q = 0.5 * 4 * b
    out_rad, in_rad = R + q, R - q
    d_out = math.radians(up) / int(math.radians(up) / (lip / out_rad))
    d_in = math.radians(up) / int(math.radians(up) / (lip / in_rad))
    n_out = sum(_ring_counts(out_rad, [d_out * i for i in range(int(math.radians(up) / (lip / out_rad)))], lip))
    n_in = sum(_ring_counts(in_rad, [d_in * i for i in range(int(math.radians(up) / (lip / in_rad)))], lip))
    n_amp = n_out + n_in // 2
    rc, d = _run(oracle, mode="sphere", box=box, amp_num=n_amp, sph_rad=R, upper_the=up,
                 with_solvent=True, sol_num=2000, seed=2)
    assert rc == 0
    _common_checks(d, n_amp, box)
    c = np.asarray(box) / 2
    r = np.linalg.norm(_chains(d, n_amp) - c, axis=2)
    u = np.arange(4) * b
    outer = np.abs(r - (R + 2 * q - u)).max(axis=1) < 1e-11
    inner = np.abs(r - (R - 2 * q + u)).max(axis=1) < 1e-11
    assert (outer | inner).all()
    assert outer[:n_out].all() and inner[n_out:].all()
    rs = np.linalg.norm(d["pos"][4 * n_amp:] - c, axis=1)
    assert (rs < R - q).all()
    # one more amphiphile than the three passes hold, and a radius past L/2
    cap = n_out + n_in + sum(_ring_counts(in_rad, [d_in * (i + 0.5) for i in range(int(math.radians(up) / (lip / in_rad)))], lip))
    # (seed 3: a draw whose kinetic temperature passes the 1e-2 check)
    assert _run(oracle, mode="sphere", box=box, amp_num=cap, sph_rad=R, upper_the=up, seed=3)[0] == 0
    assert _run(oracle, mode="sphere", box=box, amp_num=cap + 1, sph_rad=R, upper_the=up, seed=3)[0] == MDPD_ERR_CONFIG
    assert _run(oracle, mode="sphere", box=box, amp_num=10, sph_rad=15.5, upper_the=up)[0] == MDPD_ERR_CONFIG


def test_cylinder_radii_rings_and_reference_failure(oracle):
    box, R, b, lip = (30.0, 30.0, 30.0), 6.0, 0.5, 0.5
    q = 0.5 * 4 * b
    z_elem = int(box[2] / lip)
    per_out = _ring_counts(R + q, [math.pi / 2], lip)[0]
    per_in = _ring_counts(R - q, [math.pi / 2], lip)[0]
    n_out = z_elem * per_out
    n_amp = n_out + 3 * per_in + 5
    rc, d = _run(oracle, mode="cylind", box=box, amp_num=n_amp, cyl_r=R, cyl_l=30.0,
                 with_solvent=True, sol_num=1000, seed=4)
    assert rc == 0
    _common_checks(d, n_amp, box)
    ch = _chains(d, n_amp)
    r = np.linalg.norm(ch[:, :, :2] - 15.0, axis=2)
    u = np.arange(4) * b
    assert np.abs(r[:n_out] - (R + 2 * q - u)).max() < 1e-11
    assert np.abs(r[n_out:] - (R - 2 * q + u)).max() < 1e-11
    # ring k of each leaflet sits at z = k lip_len
    zk = np.concatenate([np.repeat(np.arange(z_elem), per_out), np.repeat(np.arange(4), per_in)[: n_amp - n_out]])
    assert np.abs(_mi(ch[:, 0, 2] - zk * lip, box[2])).max() < 1e-12
    rs = np.linalg.norm(d["pos"][4 * n_amp:, :2] - 15.0, axis=1)
    assert (rs < R - q).all()
    # the reference's axis passes have no guard on entry: when the outer
    # leaflet alone holds amp_num, the inner pass lays one amphiphile past
    # the end and the run fails (MakeCylindSheet :248-260)
    assert _run(oracle, mode="cylind", box=box, amp_num=n_out, cyl_r=R, cyl_l=30.0, seed=3)[0] == MDPD_ERR_CONFIG
    assert _run(oracle, mode="cylind", box=box, amp_num=n_out + 1, cyl_r=R, cyl_l=30.0, seed=3)[0] == 0


def test_flat_solvent_in_lower_half(oracle):
    box = (20.0, 20.0, 20.0)
    rc, d = _run(oracle, mode="flat", box=box, amp_num=2000, with_solvent=True, sol_num=6000, seed=9)
    assert rc == 0
    s = d["pos"][8000:]
    assert (s[:, 1] < 10.0).all() and (s >= 0).all()
    assert s[:, 0].max() > 19.0 and s[:, 1].max() > 9.0


def test_errors_of_config_maker(oracle):
    # with_solvent refused in random / asym_flat, and a kinetic temperature
    # more than 1e-2 from the target
    assert _run(oracle, mode="random", box=(10.0,) * 3, amp_num=100, with_solvent=True, sol_num=10)[0] == MDPD_ERR_CONFIG
    assert _run(oracle, mode="asym_flat", box=(10.0,) * 3, amp_num=100, with_solvent=True, sol_num=10)[0] == MDPD_ERR_CONFIG
    bad = [s for s in range(1, 20) if _run(oracle, mode="random", box=(10.0,) * 3, amp_num=4, seed=s)[0] == MDPD_ERR_CONFIG]
    assert bad  # 16 beads: |kT - 1| >= 1e-2 almost always


# ---------------------------------------------------------------- input file
def test_read_config_param_reference_file(tmp_path):
    # test/simple_dpdfluid/test_result/rho1.0/config_param.txt of the reference
    src = os.path.join(GOLD, "config_param_rho1.0.txt")
    d = tmp_path / "run"
    d.mkdir()
    (d / "config_param.txt").write_text(open(src).read())
    s = cm.read_config_param(str(d), seed=3)
    assert (s.mode, s.amp_num, s.head_unit, s.tail_unit) == (cm.MODES["random"], 1024, 1, 3)
    assert list(s.box) == [16.0, 16.0, 16.0] and s.temperature == 1.0 and s.bond_leng == 0.5
    assert s.lip_len == 0.5 and s.with_solvent == 0 and s.seed == 3
    assert cm.config_count(s) == 4096
    # a file of the older format (no head_unit) is refused like the reference's Matching
    (d / "config_param.txt").write_text("Temperature 1.0\nbox_leng 35.0 35.0 35.0\namp_num 4680\nlip_len 0.5\nmode flat\n")
    with pytest.raises(Exception, match="head_unit"):
        cm.read_config_param(str(d))
    # stof rounding of float tags, with_solvent by presence, mode case-insensitive
    (d / "config_param.txt").write_text(
        "Temperature 1.0\nbox_leng 30 30 30\namp_num 10 ;comment\nmode SPHERE\nhead_unit 1\ntail_unit 3\n"
        "bond_leng 0.5\nsph_rad 10.3\nupper_the 180\nwith_solvent\nsol_num 77\nlip_len 0.5\n")
    s = cm.read_config_param(str(d))
    assert s.mode == cm.MODES["sphere"] and s.with_solvent == 1 and s.sol_num == 77
    assert s.sph_rad == float(np.float32(10.3)) and s.amp_num == 10
