"""Observables beyond the Observer: pair-distance histogram (RDF), chain
bond / angle / end-to-end histograms (analysis/bond_angle/h2e.cpp:64-104)
and the 3-D count map (analysis/trjanalysis/density_map.hpp:13-87).

The bar is exact integer counts: a numpy evaluation of the same FP64
expressions (minimum image by half-box comparisons, squared lengths summed
x, y, z in order, bins against the squared edges (k w)^2, cosine bins
decided without a square root) on the downloaded positions."""
import numpy as np
import pytest

from implicit_dpd_amd import configs
from obs_ref import chain_ref, count_ref, rdf_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("skin", [0.0, 0.3])
def test_observables_single_system(engine_mod, skin):
    box = (10.0, 10.0, 8.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=4)
    p.cf_s, p.cf_b = 50.0, 10.0
    s = configs.random_chains(600, box, 3, n_solvent=400)
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(2, p.pair_seed)
    if skin:
        e.set_tuning(verlet_skin=skin)
    e.run(15)
    g = e.download()
    bx = np.array(box)
    rdf = e.rdf(1.0, 50)[0]
    assert np.array_equal(rdf, rdf_ref(g["pos"], bx, 1.0, 50))
    assert rdf.sum() > 0
    ch = e.chain_histograms(40, 1.2)[0]
    ref = chain_ref(g["pos"], s["topo"], bx, 40, 1.2)
    for row in range(4):
        assert np.array_equal(ch[row], ref[row]), row
    # every angle lands in a bin; bonds longer than bond_max fall outside
    assert ch[2].sum() == 1200 and 0 < ch[0].sum() <= 600 and 0 < ch[1].sum() <= 1200
    for mask in (0xFFFFFFFF, 0b011):
        cm = e.count_map((5, 4, 3), mask)[0]
        assert np.array_equal(cm, count_ref(g["pos"], g["prop"], bx, (5, 4, 3), mask)), mask
    assert e.count_map((5, 4, 3))[0].sum() == n
    e.close()


def test_observables_per_replica(engine_mod):
    """Two replicas of different boxes and sizes in one context: every
    replica's histograms equal the ones of its own beads."""
    params, systems = [], []
    for r, (box, nb) in enumerate((((8.0, 8.0, 8.0), 1536), ((9.0, 8.0, 10.0), 2160))):
        params.append(configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=10 + r))
        systems.append(configs.random_fluid(nb, box, 20 + r))
    sizes = [len(x["pos"]) for x in systems]
    st = configs.concat_systems(systems)
    e = engine_mod.Engine(params, n_beads=sizes)
    e.upload(st["pos"], st["vel"], st["ids"], st["prop"], st["amp_id"], st["unit"])
    e.set_clock(0, 0)
    e.run(6)
    g = e.download()
    rdf = e.rdf(0.9, 30)
    cm = e.count_map((4, 4, 4))
    off = np.concatenate([[0], np.cumsum(sizes)])
    for r, p in enumerate(params):
        pos = g["pos"][off[r]:off[r + 1]]
        bx = np.array(p.box[:])
        assert np.array_equal(rdf[r], rdf_ref(pos, bx, 0.9, 30)), r
        assert np.array_equal(cm[r], count_ref(pos, g["prop"][off[r]:off[r + 1]], bx,
                                                (4, 4, 4), 0xFFFFFFFF)), r
    e.close()
