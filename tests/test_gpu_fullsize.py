"""GPU parity at the benchmark configurations' full sizes.

- Membrane (1,048,576 beads as stacked bilayers, the headline system; and
  the same model on random chains), MDPD slab (198,246) and
  the 64-replica ensemble (about 2.2M beads), after 20 steps of the
  production path: the pair sets of the step (cutoff 1, the pair list the
  force pass walks) and of rd equal the FP64 oracle's cell-list pair sets
  (bit-exact); densities within 2e-6 relative and forces within 1e-5 of the
  per-bead pair-magnitude sum (floored at one unit of force), against the
  oracle on the same state; the membrane also relaxed (2200 steps).
- RandomChemEvent on the membrane (keyed and reference-MT draws) against
  the oracle's event on the same state.
- Groot-Warren fluid, 24,000 beads, 1000 steps: kinetic temperature within
  1.5% of kT, pressure within 1.5% of the oracle's run of the same system,
  and the excess pressure (p - rho kT) / (a rho^2) in [0.085, 0.105] (the
  Groot-Warren fit gives 0.101 for rho > 2, J. Chem. Phys. 107, 4423
  (1997); at rho = 3, a = 25 simulations give p close to 23.7, i.e. 0.092).
"""
import numpy as np
import pytest

from implicit_dpd_amd import configs

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
# The bar is relative to the bead's sum of pair-term magnitudes down to 1e-3
# (the dense systems' floor): a vapour bead whose only neighbours sit next to
# the cutoff, where the FP32 error of r (~1e-7) times sigma' |xi| would miss
# it, has those pairs' 1 - r taken from FP64 by the force pass.
FORCE_FLOOR = 1e-3


def _sorted_pairs(i, j):
    o = np.lexsort((j, i))
    return i[o], j[o]


def _check_state(oracle, p, g, gi, gj, t, base=0):
    """Pairs (global indices gi, gj of this system's beads), densities and
    forces of one system against the oracle on the same positions."""
    for cut, (ci, cj) in zip((1.0, p.rd), zip(gi, gj)):
        ci, cj = _sorted_pairs(ci - base, cj - base)
        oi, oj = _sorted_pairs(*oracle.pairs(p, g["pos"], cut))
        assert len(ci) == len(oi), (cut, len(ci), len(oi))
        assert np.array_equal(ci, oi) and np.array_equal(cj, oj), cut
    nd = p.n_dens
    dens = oracle.density(p, g["pos"], g["prop"])
    scale = max(1.0, float(np.abs(dens).max()))
    assert np.abs(g["dens"][:, :nd] - dens[:, :nd]).max() < 2e-6 * scale
    acc, _ = oracle.force(p, g["pos"], g["vel"], g["id"], g["prop"], dens, p.pair_seed + t)
    fa = oracle.force_abs(p, g["pos"], g["vel"], g["id"], g["prop"], dens, p.pair_seed + t)
    err = np.abs(g["acc"] - acc).max(axis=1) / np.maximum(fa, FORCE_FLOOR)
    assert err.max() < FORCE_TOL, (err.max(), int(err.argmax()), fa[int(err.argmax())])


def _after_steps(e, steps):
    """The step's own quantities, recomputed by the explicit calls on the
    state `steps` production steps left."""
    e.run(steps)
    e.rebuild()
    e.density()
    e.force()


FULL = {"membrane_stack": configs.membrane_stack, "membrane_gas": configs.membrane,
        "slab": configs.mdpd_slab}


@pytest.mark.parametrize("name,steps", [("membrane_stack", 20), ("membrane_stack", 2200),
                                        ("membrane_gas", 20), ("slab", 20)])
def test_full_size_pairs_density_force_match_oracle(engine_mod, oracle, name, steps):
    """The headline membrane (1,048,576 beads as five stacked bilayers, 3
    species, chains) after 20 production steps and relaxed (2200 steps, the
    state the benchmark times: about 20 pairs per bead), the same model on
    random chains, and the MDPD slab (198,246 beads, liquid + vapour), after
    20 production steps."""
    print(f"[fullsize] {name} after {steps} steps", flush=True)
    p, s = FULL[name]()
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(s["topo"]):
        e.set_topology(s["topo"])
    e.set_clock(0, p.pair_seed)
    _after_steps(e, steps)
    g = e.download()
    pairs = [e.pairs(cut) for cut in (1.0, p.rd)]
    t = e.get_clock()
    e.close()
    _check_state(oracle, p, g, [x[0] for x in pairs], [x[1] for x in pairs], t)


@pytest.mark.parametrize("mt", [False, True], ids=["keyed", "reference_mt"])
def test_membrane_full_size_chem_event_matches_oracle(engine_mod, oracle, mt):
    """RandomChemEvent on the 1,048,576-bead membrane after 20 steps (four
    core amphiphiles, cones of radius 3 and depth -1..3, p_thresld 0.3), with
    the keyed draws and with the reference's MT19937 stream seeded with 0:
    the inserted chains' ids, types and counts exact, positions to 1e-12,
    velocities to 1e-13 relative and the stream's draw count exact, against
    the oracle on the same state."""
    print(f"[fullsize] membrane chemistry event, mt={mt}", flush=True)
    p, s = configs.membrane_stack()
    n, n_amp = len(s["pos"]), len(s["topo"]) // 4
    cap = n + 4 * 4096
    e = engine_mod.Engine(p, capacity=cap)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(0, p.pair_seed)
    e.run(20)
    g0 = e.download()
    t = e.get_clock()
    if mt:
        e.set_chem_mt(0)
    e.rebuild()
    e.density()
    kw = dict(p_thresld=0.3, eps=0.2, influ_rad=3.0, influ_hei=1.0, influ_dep=3.0, influ_grd=0.0)
    core = np.array([0, 4, 8, 12], np.uint32)
    stop = e.chem_event(core_ptcl_id=core, max_amp_num=10 ** 9, **kw)
    g = e.download()
    draws = e.chem_mt_draws()
    e.close()
    o = oracle.System(p, cap, n_amp + 4096)
    o.set(g0["pos"], g0["vel"], g0["id"], g0["prop"], g0["amp_id"], g0["unit"])
    o.set_topology(s["topo"])
    o.set_clock(t)
    if mt:
        o.set_mt(0)
    cp = engine_mod.MdpdChemParams(kw["p_thresld"], kw["eps"], kw["influ_rad"], kw["influ_hei"],
                                   kw["influ_dep"], kw["influ_grd"], 10 ** 9, len(core),
                                   engine_mod._u(core))
    k, ostop = o.chem_event(cp)
    r = o.get()
    assert k > 10 and stop == ostop == 0, k
    assert len(g["pos"]) == len(r["pos"]) == n + 4 * k
    for key in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(g[key], r[key]), key
    assert np.abs(g["pos"] - r["pos"]).max() < 1e-12
    assert np.array_equal(g["vel"][:n], r["vel"][:n])
    assert np.allclose(g["vel"][n:], r["vel"][n:], rtol=1e-13, atol=1e-14)
    assert draws == (o.mt_draws() if mt else 0)


def test_membrane_full_size_observables(engine_mod, oracle):
    """Pair-distance histogram, chain histograms (262,144 chains) and the
    count map of the 1M-bead membrane after 20 steps, exact against numpy
    on the downloaded positions (pairs from the oracle's cell list)."""
    from obs_ref import chain_ref, count_ref, rdf_from_pairs
    p, s = configs.membrane()
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(0, p.pair_seed)
    e.run(20)
    g = e.download()
    rdf = e.rdf(0.75, 75)[0]
    ch = e.chain_histograms(60, 1.5)[0]
    cm = e.count_map((50, 50, 20), 0b011)[0]
    e.close()
    box = np.array(p.box[:])
    i, j = oracle.pairs(p, g["pos"], 1.0)
    assert np.array_equal(rdf, rdf_from_pairs(g["pos"], box, i, j, 0.75, 75))
    ref = chain_ref(g["pos"], s["topo"], box, 60, 1.5)
    for row in range(4):
        assert np.array_equal(ch[row], ref[row]), row
    assert np.array_equal(cm, count_ref(g["pos"], g["prop"], box, (50, 50, 20), 0b011))


def test_ensemble_full_size_replicas_match_oracle(engine_mod, oracle):
    """The 64-replica ensemble (about 2.2M beads in one context, replica
    sizes U[4096, 65536], coefficients and seeds per replica) after 20
    production steps: every replica's pairs, densities and forces against
    the oracle run on that replica alone."""
    params, systems = configs.ensemble()
    sizes = [len(x["pos"]) for x in systems]
    st = configs.concat_systems(systems)
    e = engine_mod.Engine(params, n_beads=sizes)
    e.set_tuning(nlist_width=256)
    e.upload(st["pos"], st["vel"], st["ids"], st["prop"], st["amp_id"], st["unit"])
    e.set_clock(0, 0)
    _after_steps(e, 20)
    g = e.download()
    pairs = [e.pairs(cut) for cut in (1.0, 0.85)]
    t = e.get_clock()
    e.close()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for r, p in enumerate(params):
        b0, b1 = off[r], off[r + 1]
        gr = {k: v[b0:b1] for k, v in g.items()}
        sel = []
        for (i, j), cut in zip(pairs, (1.0, 0.85)):
            m = (i >= b0) & (i < b1)
            i, j = i[m], j[m]
            if cut != 1.0:
                # the context lists one cutoff for all replicas; keep this
                # replica's rd pairs
                d = gr["pos"][i - b0] - gr["pos"][j - b0]
                box = np.array(p.box[:])
                d -= np.round(d / box) * box
                keep = (d * d).sum(axis=1) < p.rd * p.rd
                i, j = i[keep], j[keep]
            sel.append((i, j))
        _check_state(oracle, p, gr, [x[0] for x in sel], [x[1] for x in sel], t, base=b0)


def _run_observe(sim, warm, blocks, every, sample=None):
    """Mean temperature and pressure over `blocks` observations `every`
    steps apart after `warm` steps; `sample(sim)` is called at each
    observation too and its results are returned stacked."""
    sim.run(warm)
    T, P, S = [], [], []
    for _ in range(blocks):
        sim.run(every)
        o = sim.observe()
        kt = o["kin_temp"] if isinstance(o, dict) else o.kin_temp[:]
        pr = o["pressure"] if isinstance(o, dict) else o.pressure[:]
        T.append(float(np.mean(kt[:3])))
        P.append(float(np.mean(pr[:3])))
        if sample is not None:
            S.append(sample(sim))
    return np.mean(T), np.mean(P), (np.array(S) if S else None)


def _agree(a, b, z=5.0, floor=0.0):
    """Two stacks of sampled histograms agree bin by bin within z combined
    standard errors of their means (plus an absolute floor)."""
    ma, mb = a.mean(axis=0), b.mean(axis=0)
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    bad = np.abs(ma - mb) > z * se + floor
    return not bad.any(), (np.abs(ma - mb) / np.maximum(se, 1e-12)).max()


def test_groot_warren_24k_equation_of_state(engine_mod, oracle):
    """Long run on both sides from the same start: temperature, pressure and
    the pair-distance histogram agree within statistical error."""
    from obs_ref import rdf_from_pairs
    p, s = configs.groot_warren_fluid()
    n = len(s["pos"])
    box = np.array(p.box[:])
    rmax, nb = 0.96, 48  # the top edge stays inside the oracle's r < 1 pair list
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"])
    e.set_clock(0, p.pair_seed)
    Tg, Pg, Hg = _run_observe(e, 400, 12, 50, lambda sim: sim.rdf(rmax, nb)[0])
    e.close()
    o = oracle.System(p, n)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"])
    o.set_clock(0)

    def orc_rdf(sim):
        pos = sim.get()["pos"]
        i, j = oracle.pairs(p, pos, 1.0)
        return rdf_from_pairs(pos, box, i, j, rmax, nb)

    To, Po, Ho = _run_observe(o, 400, 12, 50, orc_rdf)
    rho = n / (20.0 ** 3)
    alpha = (Pg - rho) / (25.0 * rho * rho)
    assert abs(Tg - 1.0) < 0.015, Tg
    assert abs(Pg - Po) < 0.015 * Po, (Pg, Po)
    assert 0.085 < alpha < 0.105, (Pg, alpha)
    ok, zmax = _agree(Hg, Ho, floor=2.0)
    assert ok, zmax


def test_chain_system_long_run_histograms(engine_mod, oracle):
    """4096 H1T3 chains (16,384 beads) of the membrane model: bond-length,
    bond-angle and end-to-end histograms of independent long runs on the
    GPU and on the oracle agree within statistical error."""
    from obs_ref import chain_ref
    p, s = configs.membrane(n_chain=4096, box=(25.0, 25.0, 10.0), conv_p=0.0)
    n = len(s["pos"])
    box = np.array(p.box[:])
    nb, bmax = 40, 1.2
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(0, p.pair_seed)
    _, Pg, Hg = _run_observe(e, 300, 10, 40, lambda sim: sim.chain_histograms(nb, bmax)[0])
    e.close()
    o = oracle.System(p, n, len(s["topo"]) // 4 + 1)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"])
    o.set_clock(0)
    _, Po, Ho = _run_observe(o, 300, 10, 40,
                             lambda sim: chain_ref(sim.get()["pos"], s["topo"], box, nb, bmax))
    for row in range(4):
        ok, zmax = _agree(Hg[:, row], Ho[:, row], floor=2.0)
        assert ok, (row, zmax)


def test_slab_long_run_density_profile(engine_mod, oracle):
    """A 24,576-bead MDPD liquid slab (A = -40, B = 25, rd = 0.75) in a
    16 x 16 x 32 box: the z density profile of independent long runs on the
    GPU and on the oracle agrees within statistical error."""
    p, s = configs.mdpd_slab(n=24576, box=(16.0, 16.0, 32.0))
    n = len(s["pos"])
    nb = 32
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"])
    e.set_clock(0, p.pair_seed)
    _, _, Hg = _run_observe(e, 300, 10, 30, lambda sim: sim.density_profile(2, nb)[0])
    e.close()
    o = oracle.System(p, n)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"])
    o.set_clock(0)

    def orc_profile(sim):
        z = sim.get()["pos"][:, 2]
        b = np.clip(np.floor(z * (1.0 / 32.0) * nb).astype(int), 0, nb - 1)
        return np.bincount(b, minlength=nb).astype(float)

    _, _, Ho = _run_observe(o, 300, 10, 30, orc_profile)
    ok, zmax = _agree(Hg, Ho, floor=2.0)
    assert ok, zmax
    # a liquid slab: the middle is dense, the edges of the box nearly empty
    mid = Hg.mean(axis=0)[nb // 4:3 * nb // 4].mean()
    assert mid > 20 * max(Hg.mean(axis=0)[0], 1.0)
