"""Pins of the oracle's RandomChemEvent (chemmanager.hpp:244-281 with the
LOCAL_CHEM_EVENT target test DetermineTargetIdWithCorePos, :116-164) against
geometry fixed by construction, brute-force cone counts and binomial
statistics (no GPU).

- Target test (:116-164): the core head's membrane normal h2t points from
  the head to its hydrophobic neighbour CM; the head is moved along h2t by
  the projection of (hydrophilic CM - head); an amphiphile is a candidate
  when its head lies in the truncated cone rad <= influ_grd * depth +
  influ_rad, -influ_hei <= depth <= influ_dep around that apex, and is a
  target when its draw is <= p_thresld (a draw is always in (0, 1), so
  p_thresld = 1 takes every candidate and 0 none).  Several cores each add
  their targets, so a head inside two cones is inserted twice.
- Insertion (:35-67, :166-173, :260-270): for target head h with unit-2
  bead t, h_hat = (t - h)/|t - h| (minimum image), a unit tangent t_hat
  perpendicular to h_hat, bead u of the new amphiphile at
  wrap(h + eps t_hat + u bond_leng h_hat); ids N + 4i + u, unit u, head
  Hyphil then three Hyphob, amphiphile id n_amp + i, appended to the
  topology (the MPI path's AppendTopol, bdf_calculator.hpp:442-543).

The draws are keyed Saru streams in place of the reference's PS::MT
sequence (absent here), so only what does not depend on the draw values is
pinned exactly; the draws themselves are pinned through Saru's own vectors.
"""
import math

import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.capi import MdpdChemParams

from test_oracle_model_pins import cm_brute

HYPHIL, HYPHOB, SOLVENT = configs.HYPHIL, configs.HYPHOB, configs.SOLVENT
TAG_TANG = 0x5BD1E995


def _minimg(d, box):
    return d - box * np.round(d / box)


def _params(box, chem_seed=77):
    from oracle import pyoracle
    p = pyoracle.calc_inter_coef(30.0, 50.0, 25.0, 0.1, [3.0] * 6, 0.005)
    for a in range(3):
        p.box[a] = box[a]
    p.cf_s, p.cf_b, p.cf_b_rigid, p.bond_leng = 50.0, 10.0, 10.0, 0.5
    p.pair_seed, p.chem_seed = 5, chem_seed
    return p


def _chem(core_ids, p_thresld, eps=0.2, rad=2.0, grd=0.5, hei=1.0, dep=3.0, max_amp=10 ** 6):
    ids = np.ascontiguousarray(core_ids, dtype=np.uint32)
    cp = MdpdChemParams(p_thresld, eps, rad, hei, dep, grd, max_amp, len(ids),
                        ids.ctypes.data_as(MdpdChemParams._fields_[8][1]))
    cp._keep = ids
    return cp


# -------------------------------------------------------- constructed case
SHIFT = np.array([6.5, 0.0, 0.0])   # moves the core next to the x face
BOX = np.array([16.0, 16.0, 16.0])
H = np.array([8.0, 8.0, 8.0])       # core head before the shift
# (x, y, z) heads of the test amphiphiles before the shift, with the apex at
# (8, 8, 8.3) and h2t = -z: depth = 8.3 - z, rad = in-plane distance from
# (8, 8); cone rad <= 0.5 depth + 2, -1 <= depth <= 3
HEADS = [
    ((10.5, 8.0, 6.8), True),    # depth 1.5, rad 2.5 <= 2.75
    ((8.0, 10.8, 6.8), False),   # depth 1.5, rad 2.8 > 2.75
    ((6.8, 8.0, 9.2), True),     # depth -0.9 >= -1, rad 1.2 <= 1.55
    ((8.0, 8.5, 9.4), False),    # depth -1.1 < -1
    ((4.6, 8.0, 5.4), True),     # depth 2.9 <= 3, rad 3.4 <= 3.45
    ((8.0, 8.0, 5.2), False),    # depth 3.1 > 3
    ((8.0, 5.8, 7.8), True),     # depth 0.5, rad 2.2 <= 2.25
    ((5.7, 8.0, 7.8), False),    # depth 0.5, rad 2.3 > 2.25
]


def _constructed():
    """Amphiphile 0 is the core: head at H, tail straight down (-z), so its
    hydrophobic neighbours within 1.2 are units 1 and 2 (CM (8, 8, 7.25),
    h2t = -z).  A free Hyphil bead at (8.3, 8, 8.6) makes the hydrophilic CM
    (8.15, 8, 8.3): the projection moves the apex to (8, 8, 8.3) and drops
    the in-plane offset.  Test amphiphiles have their tails pointing away
    from H, so no other bead enters the core head's CM sums."""
    pos, prop = [], []
    for u in range(4):
        pos.append(H + [0, 0, -0.5 * u])
        prop.append(HYPHIL if u == 0 else HYPHOB)
    dirs = []
    for h, _ in HEADS:
        h = np.array(h)
        e = (h - H) / np.linalg.norm(h - H)
        dirs.append(e)
        # units 1, 2 along e, unit 3 bent off it (h_hat is taken to unit 2)
        kink = np.cross(e, [0.0, 0.0, 1.0] if abs(e[2]) < 0.9 else [1.0, 0.0, 0.0])
        kink /= np.linalg.norm(kink)
        for u, q in enumerate((0.0, 0.5 * e, 1.0 * e, 1.0 * e + 0.5 * kink)):
            pos.append(h + q)
            prop.append(HYPHIL if u == 0 else HYPHOB)
    pos.append(np.array([8.3, 8.0, 8.6]))
    prop.append(HYPHIL)
    pos = np.array(pos) + SHIFT
    pos -= np.floor(pos / BOX) * BOX
    n = len(pos)
    n_amp = 1 + len(HEADS)
    amp_id = np.concatenate([np.repeat(np.arange(n_amp), 4), [0xFFFFFFFF]]).astype(np.uint32)
    unit = np.concatenate([np.tile(np.arange(4), n_amp), [0xFFFFFFFF]]).astype(np.uint32)
    return dict(pos=pos, prop=np.array(prop, np.uint32), ids=np.arange(n, dtype=np.uint32),
                amp_id=amp_id, unit=unit, topo=np.arange(4 * n_amp, dtype=np.uint32),
                vel=np.zeros((n, 3)), dirs=[np.array([0.0, 0.0, -1.0])] + dirs)


def _run_event(oracle, s, cp, time=40, extra=64, core_amp=(0,)):
    p = _params(BOX if "box" not in s else s["box"])
    n = len(s["pos"])
    n_amp = len(s["topo"]) // 4
    o = oracle.System(p, n + 4 * extra, n_amp + extra)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"], core_amp_id=list(core_amp))
    o.set_clock(time)
    k, stop = o.chem_event(cp)
    return p, o, k, stop


def test_constructed_cone_targets_and_geometry(oracle):
    s = _constructed()
    # the CM sums the target test starts from, by brute force
    hl, hb, _, _ = cm_brute(s["pos"], s["prop"], BOX)
    assert np.allclose(hb[0], (H + [0, 0, -0.75] + SHIFT) % BOX, atol=1e-12, rtol=0)
    assert np.allclose(hl[0], (H + [0.15, 0, 0.3] + SHIFT) % BOX, atol=1e-12, rtol=0)
    n0, amp0 = len(s["pos"]), len(s["topo"]) // 4
    eps = 0.2
    p, o, k, stop = _run_event(oracle, s, _chem([0], 1.0, eps=eps))
    # the core itself (depth 0.3, rad 0) and the four heads marked inside
    expect = [0] + [1 + i for i, (_, inside) in enumerate(HEADS) if inside]
    assert k == len(expect) and stop == 0
    r = o.get()
    topo = o.topology()
    assert len(r["pos"]) == n0 + 4 * k
    for i, a in enumerate(expect):
        head = s["pos"][4 * a]
        h_hat = s["dirs"][a]
        new = slice(n0 + 4 * i, n0 + 4 * i + 4)
        assert np.array_equal(r["id"][new], n0 + 4 * i + np.arange(4))
        assert np.array_equal(r["prop"][new], [HYPHIL, HYPHOB, HYPHOB, HYPHOB])
        assert np.array_equal(r["amp_id"][new], np.full(4, amp0 + i))
        assert np.array_equal(r["unit"][new], np.arange(4))
        assert np.array_equal(topo[4 * (amp0 + i):4 * (amp0 + i) + 4], n0 + 4 * i + np.arange(4))
        d = _minimg(r["pos"][new] - head, BOX)          # offsets from the target head
        t = d[0] / eps                                  # the unit tangent
        assert math.isclose(np.linalg.norm(t), 1.0, rel_tol=1e-12)
        assert abs(t @ h_hat) < 1e-12
        for u in range(4):
            assert np.allclose(d[u], eps * t + u * 0.5 * h_hat, atol=1e-12, rtol=0)
        assert np.all((r["pos"][new] >= 0) & (r["pos"][new] < BOX))
        assert np.all(np.isfinite(r["vel"][new]))
        assert np.all(r["delta"][new] == 0) and np.all(r["vel_buf"][new] == 0)
    # the old beads are untouched
    for key in ("pos", "vel"):
        assert np.array_equal(r[key][:n0], s[key])


def test_tangent_follows_keyed_draw(oracle):
    """t_hat = normalise(b - h_hat (b . h_hat)) with b = (cos th, sin th, 0)
    and th = 2 pi u, u the first uni draw of Saru(new amphiphile id, time,
    chem_seed ^ TAG_TANG) (GetMembraneTangRandomVec, chemmanager.hpp:166-173,
    with the keyed stream in place of genrand_real1)."""
    s = _constructed()
    n0, amp0 = len(s["pos"]), len(s["topo"]) // 4
    p, o, k, _ = _run_event(oracle, s, _chem([0], 1.0, eps=0.3), time=1234)
    r = o.get()
    expect = [0] + [1 + i for i, (_, inside) in enumerate(HEADS) if inside]
    seeds = np.arange(amp0, amp0 + k, dtype=np.uint32)
    _, uni, _ = oracle.saru_batch(seeds, np.full(k, 1234, np.uint32),
                                  np.full(k, p.chem_seed ^ TAG_TANG, np.uint32))
    for i, a in enumerate(expect):
        th = 2 * math.pi * uni[i, 0]
        b = np.array([math.cos(th), math.sin(th), 0.0])
        h = s["dirs"][a]
        t = b - h * (b @ h)
        t /= np.linalg.norm(t)
        d = _minimg(r["pos"][n0 + 4 * i] - s["pos"][4 * a], BOX)
        assert np.allclose(d, 0.3 * t, atol=1e-12, rtol=0)


def test_threshold_zero_and_two_cores(oracle):
    s = _constructed()
    n0 = len(s["pos"])
    _, o, k, stop = _run_event(oracle, s, _chem([0], 0.0))
    assert k == 0 and stop == 0 and len(o.get()["pos"]) == n0
    # the same core listed twice: every candidate is a target once per core,
    # in core order, then amphiphile order
    _, o, k, _ = _run_event(oracle, s, _chem([0, 0], 1.0))
    expect = [0] + [1 + i for i, (_, inside) in enumerate(HEADS) if inside]
    assert k == 2 * len(expect)
    r = o.get()
    heads = [s["pos"][4 * a] for a in expect + expect]
    for i, h in enumerate(heads):
        assert math.isclose(np.linalg.norm(_minimg(r["pos"][n0 + 4 * i] - h, BOX)), 0.2, rel_tol=1e-12)


def test_stop_flag_at_max_amp_num(oracle):
    s = _constructed()
    amp0 = len(s["topo"]) // 4
    _, _, k, stop = _run_event(oracle, s, _chem([0], 1.0, max_amp=amp0 + 5))
    assert k == 5 and stop == 1
    _, _, k, stop = _run_event(oracle, s, _chem([0], 1.0, max_amp=amp0 + 6))
    assert k == 5 and stop == 0


# -------------------------------------------------------- random membrane
def _cone_candidates(s, box, core_amp, rad0, grd, hei, dep):
    """Brute-force candidate amphiphiles of one core (numpy)."""
    hl, hb, _, _ = cm_brute(s["pos"], s["prop"], box)
    hi = s["topo"][4 * core_amp]
    ch, ct, ccm = s["pos"][hi], hb[hi], hl[hi]
    h2t = _minimg(ct - ch, box)
    h2t /= np.linalg.norm(h2t)
    apex = ch + (_minimg(ccm - ch, box) @ h2t) * h2t
    apex -= np.floor(apex / box) * box
    heads = s["pos"][s["topo"][0::4]]
    c2p = _minimg(heads - apex, box)
    depth = c2p @ h2t
    rad = np.linalg.norm(c2p - depth[:, None] * h2t, axis=1)
    return np.nonzero((rad <= grd * depth + rad0) & (depth >= -hei) & (depth <= dep))[0]


def _membrane(seed):
    box = np.array([12.0, 12.0, 10.0])
    rng = np.random.default_rng(seed)
    n_amp = 700
    # a rough bilayer: heads on two sheets, tails pointing inwards
    xy = rng.random((n_amp, 2)) * box[:2]
    up = rng.random(n_amp) < 0.5
    z = np.where(up, 6.6, 3.4) + rng.normal(scale=0.2, size=n_amp)
    head = np.column_stack([xy, z])
    d = np.column_stack([rng.normal(scale=0.2, size=(n_amp, 2)), np.where(up, -1.0, 1.0)])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (head[:, None, :] + 0.5 * np.arange(4)[None, :, None] * d[:, None, :])
    pos[:, 1:] += rng.normal(scale=0.15, size=(n_amp, 3, 3))  # bent chains
    pos = pos.reshape(-1, 3)
    pos -= np.floor(pos / box) * box
    n = len(pos)
    return dict(box=box, pos=pos, vel=np.zeros((n, 3)), ids=np.arange(n, dtype=np.uint32),
                prop=np.tile(np.array([HYPHIL, HYPHOB, HYPHOB, HYPHOB], np.uint32), n_amp),
                amp_id=np.repeat(np.arange(n_amp, dtype=np.uint32), 4),
                unit=np.tile(np.arange(4, dtype=np.uint32), n_amp),
                topo=np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("seed", [1, 2])
def test_candidates_equal_brute_force_cone(oracle, seed):
    s = _membrane(seed)
    box = s["box"]
    cone = dict(rad0=2.5, grd=0.3, hei=1.5, dep=2.0)
    core = [3, 100]
    cand = [_cone_candidates(s, box, c, **cone) for c in core]
    assert all(len(c) > 10 for c in cand)
    core_ids = [int(s["topo"][4 * c]) for c in core]
    cp = _chem(core_ids, 1.0, eps=0.25, rad=cone["rad0"], grd=cone["grd"], hei=cone["hei"],
               dep=cone["dep"])
    n0 = len(s["pos"])
    p, o, k, _ = _run_event(oracle, s, cp, extra=4 * 400, core_amp=core)
    expect = np.concatenate(cand)
    assert k == len(expect)
    r = o.get()
    # each new amphiphile's head sits eps from its target's head
    for i, a in enumerate(expect):
        d = np.linalg.norm(_minimg(r["pos"][n0 + 4 * i] - s["pos"][4 * a], box))
        assert math.isclose(d, 0.25, rel_tol=1e-10), (i, a, d)
        h_hat = _minimg(s["pos"][4 * a + 2] - s["pos"][4 * a], box)
        h_hat /= np.linalg.norm(h_hat)
        step = _minimg(r["pos"][n0 + 4 * i + 3] - r["pos"][n0 + 4 * i], box)
        assert np.allclose(step, 1.5 * h_hat, atol=1e-10)
    # Box-Muller velocities of the keyed draws: mean 0, variance kT = 1
    v = r["vel"][n0:].ravel()
    assert abs(v.mean()) < 5 / math.sqrt(len(v)) and abs(v.var() - 1) < 5 * math.sqrt(2 / len(v))


def test_target_count_is_binomial(oracle):
    s = _membrane(3)
    box = s["box"]
    cone = dict(rad0=4.0, grd=0.0, hei=2.0, dep=3.0)
    cand = _cone_candidates(s, box, 10, **cone)
    m = len(cand)
    assert m > 40
    core_id = int(s["topo"][40])
    total, trials, pthr = 0, 12, 0.3
    for t in range(trials):
        cp = _chem([core_id], pthr, rad=cone["rad0"], grd=0.0, hei=cone["hei"], dep=cone["dep"])
        _, _, k, _ = _run_event(oracle, s, cp, time=500 + t, extra=m + 4, core_amp=[10])
        assert 0 <= k <= m
        total += k
    mean, sd = trials * m * pthr, math.sqrt(trials * m * pthr * (1 - pthr))
    assert abs(total - mean) < 5 * sd, (total, mean, sd)


# ------------------------------------------------------ reference-MT mode
def test_mt19937_published_values(oracle):
    """The oracle's MT19937 (the generator behind PS::MT): the 10000th
    genrand_int32 output after init_genrand(5489) is 4123659995 (the check
    value of the published generator), and the raw outputs equal numpy's
    legacy MT19937 (RandomState seeds with init_genrand)."""
    a = oracle.mt_u32(5489, 10000)
    assert int(a[-1]) == 4123659995
    for seed in (0, 1, 77, 12345):
        ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=2000, dtype=np.uint32)
        assert np.array_equal(oracle.mt_u32(seed, 2000), ref)


def oracle_mt(seed, n):
    from oracle import pyoracle
    return pyoracle.mt_u32(seed, n)


def test_reference_mt_mode_order_and_draws(oracle):
    """RandomChemEvent with one MT19937 stream, as ChemManager(seed) drives
    it (chemmanager.hpp:116-173, 244-281): one genrand_real1 per in-cone
    head, in core then amphiphile order, compared with p_thresld; per target
    one genrand_real1 for the tangent angle and, per new bead and axis, two
    genrand_real3 for GenThermalVeloc (:18-20).  The targets, their order
    and the insertion geometry are those of the keyed mode wherever the
    draws do not decide (p_thresld = 1 and 0); the tangent angles and the
    velocities are recomputed here from the same stream."""
    s = _constructed()
    n0, amp0 = len(s["pos"]), len(s["topo"]) // 4
    expect = [0] + [1 + i for i, (_, inside) in enumerate(HEADS) if inside]
    seed = 2024
    for pth in (1.0, 0.0):
        _, ok_, kk, _ = _run_event(oracle, s, _chem([0], pth))
        p = _params(BOX)
        o = oracle.System(p, n0 + 4 * 64, amp0 + 64)
        o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        o.set_topology(s["topo"], core_amp_id=[0])
        o.set_clock(40)
        o.set_mt(seed)
        k, stop = o.chem_event(_chem([0], pth, eps=0.2))
        assert k == kk == (len(expect) if pth == 1.0 else 0)
        assert o.mt_draws() == len(expect) + k * (1 + 4 * 3 * 2)
        r, rk = o.get(), ok_.get()
        for key in ("id", "prop", "amp_id", "unit"):
            assert np.array_equal(r[key], rk[key]), key
        if k == 0:
            continue
        u32 = oracle_mt(seed, o.mt_draws()).astype(np.float64)
        q = len(expect)  # the cone draws come first
        for i, a in enumerate(expect):
            th = 2 * math.pi * (u32[q] * (1.0 / 4294967295.0))
            q += 1
            b = np.array([math.cos(th), math.sin(th), 0.0])
            h = s["dirs"][a]
            t = b - h * (b @ h)
            t /= np.linalg.norm(t)
            d = _minimg(r["pos"][n0 + 4 * i:n0 + 4 * i + 4] - s["pos"][4 * a], BOX)
            for uu in range(4):
                assert np.allclose(d[uu], 0.2 * t + uu * 0.5 * h, atol=1e-12, rtol=0)
            for uu in range(4):
                for ax in range(3):
                    r3a = (u32[q] + 0.5) * (1.0 / 4294967296.0)
                    r3b = (u32[q + 1] + 0.5) * (1.0 / 4294967296.0)
                    q += 2
                    v = math.sqrt(-2.0 * math.log(r3a)) * math.cos(2.0 * math.pi * r3b)
                    assert math.isclose(r["vel"][n0 + 4 * i + uu, ax], v, rel_tol=1e-14, abs_tol=1e-15)
        assert q == o.mt_draws()


def test_reference_mt_mode_threshold_statistics(oracle):
    """With p_thresld = 0.3 the MT mode takes each in-cone head with its own
    genrand_real1 <= 0.3: the targets are exactly the in-cone heads whose
    draw (in stream order) passes, recomputed here from the stream."""
    s = _constructed()
    expect = [0] + [1 + i for i, (_, inside) in enumerate(HEADS) if inside]
    hits = 0
    for seed in range(40):
        p = _params(BOX)
        n0, amp0 = len(s["pos"]), len(s["topo"]) // 4
        o = oracle.System(p, n0 + 4 * 64, amp0 + 64)
        o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        o.set_topology(s["topo"], core_amp_id=[0])
        o.set_clock(40)
        o.set_mt(seed)
        k, _ = o.chem_event(_chem([0], 0.3))
        u = oracle_mt(seed, len(expect)).astype(np.float64) * (1.0 / 4294967295.0)
        chosen = [a for a, x in zip(expect, u) if x <= 0.3]
        assert k == len(chosen)
        r = o.get()
        for i, a in enumerate(chosen):
            d = _minimg(r["pos"][n0 + 4 * i] - s["pos"][4 * a], BOX)
            assert abs(np.linalg.norm(d) - 0.2) < 1e-12   # eps from the chosen head
        hits += k
    # 40 events x 5 candidates at p = 0.3: mean 60, sd 6.5
    assert 30 < hits < 90


# A thermal velocity component is sqrt(-2 T log(u)) cos(2 pi u') with u, u'
# two genrand_real3 calls inside one expression (chemmanager.hpp:18-20),
# whose evaluation order C++ leaves open.  The oracle's MT mode takes the
# logarithm's draw first; this probe, an expression of that shape compiled
# with the reference's Makefile compiler and flags (g++ -O3 -std=c++11
# -ffast-math -funroll-loops), shows that g++ does the same.
_ORDER_PROBE = r"""
#include <cmath>
#include <cstdio>
static int n_calls = 0;
__attribute__((noinline)) double draw() { return 0.3 + 0.3 * n_calls++; }
static inline double thermal(double kt) {
  return std::sqrt(-2.0 * kt * std::log(draw())) * std::cos(2.0 * M_PI * draw());
}
int main() { std::printf("%.17g\n", thermal(1.0)); return 0; }
"""


@pytest.mark.parametrize("flags", [["-O3", "-std=c++11", "-ffast-math", "-funroll-loops"],
                                   ["-O0", "-std=c++11"]])
def test_gcc_draws_the_logarithm_first(tmp_path, flags):
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "order.cpp"
    src.write_text(_ORDER_PROBE)
    exe = tmp_path / "order"
    subprocess.run(["g++"] + flags + [str(src), "-o", str(exe)], check=True)
    v = float(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    log_first = math.sqrt(-2.0 * math.log(0.3)) * math.cos(2.0 * math.pi * 0.6)
    cos_first = math.sqrt(-2.0 * math.log(0.6)) * math.cos(2.0 * math.pi * 0.3)
    assert abs(v - log_first) < 1e-12 and abs(v - cos_first) > 0.5, (v, log_first, cos_first)
