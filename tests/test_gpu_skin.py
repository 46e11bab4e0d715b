"""Skin mode (mdpd_tuning.verlet_skin > 0): cells at least 1 + skin wide, a
neighbour list of radius 1 + skin, and cells + list rebuilt on the device
only in steps where some bead has moved more than skin / 2 since the last
build.

Parity bar: the pair sets the density and force passes see are those of the
FP64 oracle after every run (bit-exact, also after steps without a rebuild);
trajectories agree with the oracle and with the every-step path to the FP32
rounding of the pair terms; graph replay is bit-identical to eager launches.
"""
import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.capi import MDPD_ERR_ARG, MdpdError

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    box = (8.0, 8.0, 8.0)
    p = configs.groot_warren_params(box, pair_seed=100)
    out.append(("groot_warren", p, configs.random_fluid(1536, box, 1)))
    box = (9.0, 8.0, 10.0)
    p = configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=3)
    out.append(("warren", p, configs.random_fluid(4320, box, 2)))
    box = (10.0, 10.0, 8.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=4)
    p.cf_s, p.cf_b = 50.0, 10.0
    out.append(("warren3", p, configs.random_chains(600, box, 3)))
    return out


SYSTEMS = _systems()
IDS = [x[0] for x in SYSTEMS]


def _engine(engine_mod, p, s, skin, time=1, extra=64):
    e = engine_mod.Engine(p, capacity=len(s["pos"]) + extra)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(s["topo"]):
        e.set_topology(s["topo"])
    e.set_clock(time, p.pair_seed)
    if skin:
        e.set_tuning(verlet_skin=skin, verlet_width=256)
    return e


def _check_pairs(e, oracle, p):
    g = e.download()
    for cut in sorted({1.0, p.rd}):
        gi, gj = e.pairs(cut)
        oi, oj = oracle.pairs(p, g["pos"], cut, brute=True)
        assert len(gi) == len(oi), (cut, len(gi), len(oi))
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), cut


@pytest.mark.parametrize("name,p,s", SYSTEMS, ids=IDS)
@pytest.mark.parametrize("skin", [0.06, 0.4])
def test_skin_pair_sets_exact(engine_mod, oracle, name, p, s, skin):
    """The production pair list (cutoff 1, from the skin list) and the cell
    search (cutoff rd, cells of the last build, beads up to skin / 2 away)
    equal the FP64 brute-force pair sets after every step, whether the step
    rebuilt or reused an older build."""
    e = _engine(engine_mod, p, s, 0.0)
    e.run(40)  # past the violent first steps of the random start
    e.set_tuning(verlet_skin=skin, verlet_width=256)
    rebuilt = []
    for k in range(8):
        b = e.skin_builds()
        e.run(1)  # a run builds only when the displacement test asks for it
        rebuilt.append(e.skin_builds() - b)
        _check_pairs(e, oracle, p)
    b = e.skin_builds()
    e.run(12)
    _check_pairs(e, oracle, p)
    assert rebuilt[0] == 1  # the tuning change invalidated the last build
    if skin > 0.1:
        assert sum(rebuilt) < len(rebuilt), rebuilt  # some steps reused a build
    else:
        assert e.skin_builds() - b >= 3
    e.close()


@pytest.mark.parametrize("name,p,s", SYSTEMS, ids=IDS)
def test_skin_steps_track_oracle(engine_mod, oracle, name, p, s):
    """Five steps in skin mode against the oracle (same bar as the
    every-step path), then 40 steps against the every-step path."""
    n = len(s["pos"])
    e = _engine(engine_mod, p, s, 0.3)
    o = oracle.System(p, n + 64, len(s["topo"]) // 4 + 1)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(s["topo"]):
        o.set_topology(s["topo"])
    o.set_clock(1)
    e.run(5)
    o.run(5)
    g = e.download()
    r = o.get()
    box = np.array(p.box[:])
    dx = g["pos"] - r["pos"]
    dx -= np.round(dx / box) * box
    assert np.abs(dx).max() < 1e-6, np.abs(dx).max()
    assert np.abs(g["vel"] - r["vel"]).max() < 1e-3
    go = e.observe()
    ro = o.observe()
    assert np.allclose(go["kin_temp"], ro.kin_temp[:], rtol=1e-6)
    assert np.allclose(go["pressure"], ro.pressure[:], rtol=1e-5, atol=1e-6)
    e.close()

    a = _engine(engine_mod, p, s, 0.0)
    b = _engine(engine_mod, p, s, 0.1)
    a.run(40)
    b.run(40)
    assert b.skin_builds() > 1
    ga, gb = a.download(), b.download()
    dx = ga["pos"] - gb["pos"]
    dx -= np.round(dx / box) * box
    # FP32 rounding differences of the pair terms, grown over 40 steps
    assert np.abs(dx).max() < 1e-5, np.abs(dx).max()
    assert np.array_equal(ga["id"], gb["id"]) and np.array_equal(ga["prop"], gb["prop"])
    oa, ob = a.observe(), b.observe()
    assert np.allclose(oa["kin_temp"], ob["kin_temp"], rtol=1e-4)
    a.close()
    b.close()


def test_skin_graph_matches_eager(engine_mod):
    name, p, s = SYSTEMS[2]
    res = []
    for graph in (1, 1, 0):
        e = _engine(engine_mod, p, s, 0.15, time=3)
        e.set_tuning(use_graph=graph)
        e.run(25)
        res.append((e.download(), e.skin_builds()))
        e.close()
    assert res[0][1] == res[1][1] == res[2][1] and res[0][1] > 1
    for k in ("pos", "vel", "acc", "dens"):
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
        assert np.array_equal(res[0][0][k], res[2][0][k]), k


def test_skin_chem_event_after_run(engine_mod, oracle):
    """A chemistry event right after a skin-mode run: the neighbour
    centre-of-mass search uses the cells of the last build widened by the
    skin; the outcome equals the oracle's on the same state."""
    box = (8.0, 9.0, 8.0)
    n_amp = 500
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    n = len(s["pos"])
    cap = n + 4 * 400
    core_ids = np.array([0, 40, 80], np.uint32)
    kw = dict(p_thresld=0.5, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.0,
              max_amp_num=n_amp + 300)
    e = _engine(engine_mod, p, s, 0.3, time=123, extra=4 * 400)
    e.set_topology(s["topo"], core_amp_id=[0, 10, 20])
    e.run(12)
    g0 = e.download()
    t = 123 + 12
    stop = e.chem_event(core_ptcl_id=core_ids, **kw)
    g = e.download()
    o = oracle.System(p, cap, n_amp + 400)
    o.set(g0["pos"], g0["vel"], g0["id"], g0["prop"], g0["amp_id"], g0["unit"])
    o.set_topology(s["topo"], core_amp_id=[0, 10, 20])
    o.set_clock(t)
    cp = engine_mod.MdpdChemParams(kw["p_thresld"], kw["eps"], kw["influ_rad"], kw["influ_hei"],
                                   kw["influ_dep"], kw["influ_grd"], kw["max_amp_num"], 3,
                                   engine_mod._u(core_ids))
    k, ostop = o.chem_event(cp)
    r = o.get()
    assert stop == ostop and k > 0
    assert len(g["pos"]) == len(r["pos"]) == n + 4 * k
    for key in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(g[key], r[key]), key
    assert np.abs(g["pos"] - r["pos"]).max() < 1e-12
    # and the run goes on (the next run builds again for the new beads)
    e.run(3)
    _check_pairs(e, oracle, p)
    e.close()


def test_skin_small_box(engine_mod, oracle):
    """A box edge of three list radii (beads crossing the periodic boundary
    between builds keep coordinates continuous with their build cell), and
    the size limit."""
    box = (5.0, 6.0, 6.0)
    p = configs.groot_warren_params(box, pair_seed=1)
    s = configs.random_fluid(540, box, 1)
    e = _engine(engine_mod, p, s, 0.0)
    e.set_tuning(verlet_skin=0.6)  # 5 >= 3 (1 + 0.6): accepted
    with pytest.raises(MdpdError) as ex:
        e.set_tuning(verlet_skin=0.7)  # 5 < 3 (1 + 0.7)
    assert ex.value.code == MDPD_ERR_ARG
    with pytest.raises(MdpdError):
        e.set_tuning(verlet_skin=0.01)
    # the failed calls left the accepted tuning in place
    assert abs(e.tuning().verlet_skin - 0.6) < 1e-7
    reused = 0
    for k in range(16):
        b = e.skin_builds()
        e.run(1)
        reused += e.skin_builds() == b
        _check_pairs(e, oracle, p)
    assert reused > 0
    e.close()
