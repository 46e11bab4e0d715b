"""The density pass's neighbour centre-of-mass sums on the device
(FPDPD::nei_cm_pos, nbdf_calculator.hpp:18-58 with the write-back of
ptcl_class.hpp:53-57) against the oracle, whose sums are pinned by brute
force and constructed cases in tests/test_oracle_model_pins.py."""
import numpy as np
import pytest

from implicit_dpd_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nsteps", [0, 6])
def test_nei_cm_pos_matches_oracle(engine_mod, oracle, nsteps):
    box = (9.0, 10.0, 8.0)
    p, _ = configs.reference_params(box, pair_seed=7, init_amp_num=450, dt=0.005)
    s = configs.random_chains(450, box, 13, n_solvent=500)
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(10, p.pair_seed)
    if nsteps:
        e.run(nsteps)
    else:
        e.rebuild()
    rec = e.download_fpdpd()
    e.close()
    pos = np.ascontiguousarray(rec["pos"])
    _, cm, cnt = oracle.density(p, pos, rec["prop"], with_cm=True)
    got = rec["nei_cm_pos"]
    # hydrophilic CM (self-seeded) is always defined
    assert np.allclose(got[:, 0], cm[:, :3], rtol=0, atol=1e-12)
    # hydrophobic CM: 0/0 where no Hyphob bead is within 1.2, as in the reference
    nan_o = np.isnan(cm[:, 3]) | (cnt[:, 1] == 0)
    assert np.array_equal(np.isnan(got[:, 1, 0]), nan_o)
    assert np.allclose(got[~nan_o, 1], cm[~nan_o, 3:], rtol=0, atol=1e-12)
    assert (cnt[:, 0] > 1).sum() > n // 4 and (cnt[:, 1] > 1).sum() > n // 4


# ---------------------------------------------------------------- in-run chemistry
KW = dict(p_thresld=0.03, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.1)
CORE_AMP = [0, 10, 20]


def _chem_system(engine_mod, extra=600, n_amp=500):
    box = (8.0, 9.0, 8.0)
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n + 4 * extra)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"], core_amp_id=CORE_AMP)
    core_ids = np.array([4 * a for a in CORE_AMP], np.uint32)  # their head beads
    return p, s, e, core_ids


def _state(e):
    g = e.download()
    return g, e.observe()["n_amp"], e.get_clock()


def _same(a, b):
    for k in ("pos", "vel", "vel_buf", "acc", "delta", "id", "prop", "amp_id", "unit"):
        assert np.array_equal(a[k], b[k]), k


def test_run_chemistry_equals_explicit_events(engine_mod):
    """mdpd_run with the chemistry inside its captured steps (events at steps
    with time >= beg_chem, after the CM removal at time 200) gives the state
    of the per-step sequence run(1) + mdpd_chem_event at that step's time,
    bit for bit."""
    t0, nsteps, beg = 196, 10, 198
    p, s, a, core = _chem_system(engine_mod)
    n_amp0 = len(s["topo"]) // 4
    a.set_clock(t0, p.pair_seed)
    a.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=beg, **KW)
    a.run(6)
    a.run(4)
    ga, na, ta = _state(a)
    stopped, events, inserted = a.chem_status()
    assert not stopped and events == nsteps - (beg - t0) and ta == t0 + nsteps
    assert na == n_amp0 + inserted and inserted > 0
    assert len(ga["pos"]) == len(s["pos"]) + 4 * inserted
    _, _, b, _ = _chem_system(engine_mod)
    b.set_clock(t0, p.pair_seed)
    for t in range(t0, t0 + nsteps):
        b.run(1)
        if t >= beg:
            b.set_clock(t, p.pair_seed)
            assert b.chem_event(max_amp_num=10 ** 6, core_ptcl_id=core, **KW) == 0
            b.set_clock(t + 1, p.pair_seed)
    gb, nb, tb = _state(b)
    assert nb == na and tb == ta
    _same(ga, gb)
    a.close()
    b.close()


@pytest.mark.parametrize("mt", [False, True], ids=["keyed", "reference_mt"])
def test_run_chemistry_stops_like_the_reference_loop(engine_mod, mt):
    """An event that brings amp_num to max_amp_num ends mdpd_run after its
    step (main.cpp:188 breaks before time++): the state and clock equal the
    explicit loop's at its break, and the next run is refused until the
    chemistry is set again; with keyed draws and with the reference's
    MT19937 stream (which then stands at the same draw in both)."""
    t0 = 300

    def system():
        out = _chem_system(engine_mod)
        if mt:
            out[2].set_chem_mt(9)
        return out

    p, s, a, core = system()
    n_amp0 = len(s["topo"]) // 4
    # where does the explicit loop stop?
    _, _, b, _ = system()
    b.set_clock(t0, p.pair_seed)
    counts = []
    for t in range(t0, t0 + 12):
        b.run(1)
        b.set_clock(t, p.pair_seed)
        b.chem_event(max_amp_num=10 ** 6, core_ptcl_id=core, **KW)
        counts.append(b.observe()["n_amp"])
        b.set_clock(t + 1, p.pair_seed)
    grow = [c for c in counts if c > n_amp0]
    assert len(set(grow)) >= 3, counts
    target = sorted(set(grow))[1]          # reached at the second or a later event
    k_stop = counts.index(target)
    b.close()
    # the explicit loop with that max_amp_num
    _, _, b, _ = system()
    b.set_clock(t0, p.pair_seed)
    for t in range(t0, t0 + 12):
        b.run(1)
        b.set_clock(t, p.pair_seed)
        if b.chem_event(max_amp_num=target, core_ptcl_id=core, **KW):
            break
        b.set_clock(t + 1, p.pair_seed)
    assert t == t0 + k_stop
    gb, nb, tb = _state(b)
    assert nb == target and tb == t0 + k_stop
    a.set_clock(t0, p.pair_seed)
    a.set_chemistry(max_amp_num=target, core_ptcl_id=core, beg_chem=0, **KW)
    a.run(12)
    stopped, events, _ = a.chem_status()
    assert stopped and events == k_stop + 1
    ga, na, ta = _state(a)
    assert na == nb and ta == tb
    _same(ga, gb)
    assert a.chem_mt_draws() == b.chem_mt_draws()
    with pytest.raises(RuntimeError):
        a.run(1)
    # switched off, the run goes on from the stopped state
    a.set_chemistry(max_amp_num=target, core_ptcl_id=core, enable=False, **KW)
    a.run(2)
    assert a.get_clock() == tb + 2
    a.close()
    b.close()


def test_chemistry_beyond_capacity_is_refused(engine_mod):
    p, s, e, core = _chem_system(engine_mod, extra=1)
    n0 = len(s["pos"])
    e.set_clock(50, p.pair_seed)
    e.run(2)
    kw = dict(KW, p_thresld=1.0)      # every head in a cone: far more than one chain
    with pytest.raises(RuntimeError):
        e.chem_event(max_amp_num=10 ** 6, core_ptcl_id=core, **kw)
    assert e.num_beads() == n0
    e.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=0, **kw)
    with pytest.raises(RuntimeError):
        e.run(3)
    assert e.num_beads() == n0 and e.chem_status()[0]
    e.close()


def test_pair_sets_after_insertions_match_brute_force(engine_mod, oracle):
    """The neighbour set after particles are appended (the reference's
    test/increase_n, increase_test.cpp:120-190: +20 particles per iteration
    over 20 iterations, the list compared with a brute-force one): here
    twenty steps with RandomChemEvent inside, each appending chains; after
    every one the cell list of the grown system gives the FP64 brute-force
    pair sets at 1 and rd, bit for bit."""
    p, s, e, core = _chem_system(engine_mod)
    n0 = len(s["pos"])
    e.set_clock(500, p.pair_seed)
    # chains inserted next to their templates crowd the cones: room for
    # the denser neighbourhoods (the overflow check stops the run otherwise)
    e.set_tuning(nlist_width=192)
    e.set_chemistry(max_amp_num=10 ** 6, core_ptcl_id=core, beg_chem=0, **KW)
    sizes = []
    for it in range(20):
        e.run(1)
        e.rebuild()
        g = e.download()
        sizes.append(len(g["pos"]))
        for cut in sorted({1.0, p.rd}):
            gi, gj = e.pairs(cut)
            oi, oj = oracle.pairs(p, g["pos"], cut, brute=True)
            assert len(gi) == len(oi), (it, cut)
            assert np.array_equal(gi, oi) and np.array_equal(gj, oj), (it, cut)
    assert sizes[-1] > n0 and len(set(sizes)) > 5, sizes
    e.close()
