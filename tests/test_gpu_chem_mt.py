"""Reference-MT mode of the chemistry (mdpd_set_chem_mt): the events draw
from one MT19937 stream in the reference's order (ChemManager's PS::MT,
chemmanager.hpp:18-20,140-161,167,176-179).  The device against the
oracle's reference-MT mode, whose generator is pinned by mt19937ar's
published outputs and whose draw order and counts are pinned in
tests/test_oracle_chem_pins.py; and the events inside mdpd_run against the
same events called one by one.

Tolerances as for the keyed events: ids, types and counts exact, the MT
stream's draw count exact, positions to 1e-12 and the new beads'
Box-Muller velocities to 1e-13 relative (the device's and the host's FP64
log / cos / sin differ in the last bits)."""
import numpy as np
import pytest

from implicit_dpd_amd import configs

pytestmark = pytest.mark.gpu

KW1 = dict(p_thresld=0.5, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.0)
KW2 = dict(p_thresld=0.3, eps=0.15, influ_rad=2.5, influ_hei=1.0, influ_dep=3.0, influ_grd=0.1)
CORE1 = [0, 10, 20]
CORE2 = [5, 10]


def _pair(engine_mod, oracle, extra, n_amp=500, seed=0, time=123):
    box = (8.0, 9.0, 8.0)
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    n = len(s["pos"])
    cap = n + 4 * extra
    e = engine_mod.Engine(p, capacity=cap)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"], core_amp_id=CORE1)
    e.set_clock(time, p.pair_seed)
    e.set_chem_mt(seed)
    o = oracle.System(p, cap, n_amp + extra)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"], core_amp_id=CORE1)
    o.set_clock(time)
    o.set_mt(seed)
    return p, s, e, o


def _cp(engine_mod, kw, core_amp, max_amp):
    ids = np.array([4 * a for a in core_amp], np.uint32)
    return ids, engine_mod.MdpdChemParams(kw["p_thresld"], kw["eps"], kw["influ_rad"],
                                          kw["influ_hei"], kw["influ_dep"], kw["influ_grd"],
                                          max_amp, len(ids), engine_mod._u(ids))


@pytest.mark.parametrize("seed", [0, 1234])
def test_mt_events_match_oracle(engine_mod, oracle, seed):
    """Two events in a row on one stream (the second sees the first's new
    chains), each thousands of draws long, so the stream crosses several
    twists of the 624-word state inside an event and between events.  Seed
    0 is the reference's single-process seed, 1234 the one of its own
    chemistry test (test/mpi_chem/chem_test.cpp:111)."""
    n_amp = 500
    p, s, e, o = _pair(engine_mod, oracle, extra=700, n_amp=n_amp, seed=seed)
    n = len(s["pos"])
    total = 0
    for kw, core in ((KW1, CORE1), (KW2, CORE2)):
        ids, cp = _cp(engine_mod, kw, core, n_amp + 10 ** 6)
        e.rebuild()
        e.density()
        stop = e.chem_event(core_ptcl_id=ids, max_amp_num=n_amp + 10 ** 6, **kw)
        k, ostop = o.chem_event(cp)
        assert k > 0 and stop == ostop == 0
        total += k
        g, r = e.download(), o.get()
        assert len(g["pos"]) == len(r["pos"]) == n + 4 * total
        for key in ("id", "prop", "amp_id", "unit"):
            assert np.array_equal(g[key], r[key]), key
        assert np.abs(g["pos"] - r["pos"]).max() < 1e-12
        assert np.array_equal(g["vel"][:n], r["vel"][:n])
        assert np.allclose(g["vel"][n:], r["vel"][n:], rtol=1e-13, atol=1e-14)
        assert e.chem_mt_draws() == o.mt_draws()
        assert e.observe()["n_amp"] == n_amp + total
    assert o.mt_draws() > 3 * 624
    e.close()


def test_mt_refused_event_consumes_only_the_cone_draws(engine_mod, oracle):
    """An event whose targets exceed the capacity inserts nothing; like the
    oracle's refused event, it has drawn the cone tests and nothing after."""
    n_amp = 500
    p, s, e, o = _pair(engine_mod, oracle, extra=4, n_amp=n_amp, seed=3)
    ids, cp = _cp(engine_mod, KW1, CORE1, n_amp + 100)
    e.rebuild()
    e.density()
    with pytest.raises(RuntimeError, match="capacity"):
        e.chem_event(core_ptcl_id=ids, max_amp_num=n_amp + 100, **KW1)
    k, _ = o.chem_event(cp)
    assert k == -1
    assert 0 < o.mt_draws() == e.chem_mt_draws()
    assert e.observe()["n_amp"] == n_amp
    e.close()


KWR = dict(p_thresld=0.03, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.1)


def _run_system(engine_mod):
    box = (8.0, 9.0, 8.0)
    n_amp = 500
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    e = engine_mod.Engine(p, capacity=len(s["pos"]) + 4 * 600)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"], core_amp_id=CORE1)
    return p, s, e, np.array([4 * a for a in CORE1], np.uint32)


def test_mt_run_equals_explicit_events(engine_mod):
    """mdpd_run with the reference-MT events inside its captured steps gives
    the state and stream position of run(1) + mdpd_chem_event per step, bit
    for bit; switching the mode on after keyed runs re-records the steps."""
    t0, beg = 196, 198
    p, s, a, core = _run_system(engine_mod)
    a.set_clock(t0, p.pair_seed)
    a.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=beg, **KWR)
    a.run(3)                      # keyed events, graphs recorded
    a.set_chem_mt(11)
    a.run(5)
    a.run(4)
    ga = a.download()
    _, events, inserted = a.chem_status()
    assert events == 12 - (beg - t0) and inserted > 0
    _, _, b, _ = _run_system(engine_mod)
    b.set_clock(t0, p.pair_seed)
    for t in range(t0, t0 + 12):
        if t == t0 + 3:
            b.set_chem_mt(11)
        b.run(1)
        if t >= beg:
            b.set_clock(t, p.pair_seed)
            b.chem_event(max_amp_num=10 ** 6, core_ptcl_id=core, **KWR)
            b.set_clock(t + 1, p.pair_seed)
    gb = b.download()
    assert a.chem_mt_draws() == b.chem_mt_draws() > 0
    assert a.observe()["n_amp"] == b.observe()["n_amp"]
    for k in ("pos", "vel", "vel_buf", "acc", "delta", "id", "prop", "amp_id", "unit"):
        assert np.array_equal(ga[k], gb[k]), k
    # back to the keyed draws: the stream stops advancing
    a.set_chem_mt(enable=False)
    a.run(2)
    assert a.chem_mt_draws() == 0
    a.close()
    b.close()


def test_mt_checkpoint_carries_the_stream(engine_mod, tmp_path):
    """A checkpoint taken with the reference-MT draws on holds the stream: a
    fresh context that loads it (its chemistry set, its draw mode not)
    continues the run with the events inside bit for bit, at the same
    position of the stream."""
    t0 = 196
    p, s, a, core = _run_system(engine_mod)
    a.set_clock(t0, p.pair_seed)
    a.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=t0, **KWR)
    a.set_chem_mt(13)
    a.run(4)
    ck = tmp_path / "mt.ckp"
    a.checkpoint_save(ck)
    d0 = a.chem_mt_draws()
    a.run(6)
    ga = a.download()
    _, _, b, _ = _run_system(engine_mod)
    b.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=t0, **KWR)
    assert b.chem_mt_draws() == 0
    b.checkpoint_load(ck)
    assert b.chem_mt_draws() == d0 > 0
    b.run(6)
    gb = b.download()
    assert a.chem_mt_draws() == b.chem_mt_draws() > d0
    assert a.observe()["n_amp"] == b.observe()["n_amp"]
    for k in ("pos", "vel", "vel_buf", "acc", "delta", "id", "prop", "amp_id", "unit"):
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("mt", [False, True], ids=["keyed", "reference_mt"])
def test_p_thresld_two_makes_every_head_a_target(engine_mod, oracle, mt):
    """A head outside every cone keeps rnd = 2.0 (chemmanager.hpp:143), so
    with p_thresld >= 2 every head of every core's loop is a target, in cone
    or not: n_core * n_amp insertions, in (core, chain) order, the in-cone
    heads still drawing their (now irrelevant) numbers in the MT mode."""
    box = (6.0, 6.0, 6.0)
    n_amp = 100
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 8, n_solvent=100)
    n = len(s["pos"])
    cap = n + 4 * 2 * n_amp
    e = engine_mod.Engine(p, capacity=cap)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(40, p.pair_seed)
    o = oracle.System(p, cap, 3 * n_amp)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"])
    o.set_clock(40)
    if mt:
        e.set_chem_mt(4)
        o.set_mt(4)
    kw = dict(KW1, p_thresld=2.0)
    ids, cp = _cp(engine_mod, kw, [3, 7], 10 ** 6)
    e.rebuild()
    e.density()
    e.chem_event(core_ptcl_id=ids, max_amp_num=10 ** 6, **kw)
    k, _ = o.chem_event(cp)
    assert k == 2 * n_amp
    g, r = e.download(), o.get()
    assert len(g["pos"]) == len(r["pos"]) == n + 4 * k
    for key in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(g[key], r[key]), key
    assert np.abs(g["pos"] - r["pos"]).max() < 1e-12
    assert np.allclose(g["vel"][n:], r["vel"][n:], rtol=1e-13, atol=1e-14)
    assert e.chem_mt_draws() == (o.mt_draws() if mt else 0)
    if mt:
        assert o.mt_draws() > 25 * k   # the in-cone heads' draws came first
    e.close()
