"""Exact restart: a binary checkpoint taken mid-run and loaded into a fresh
context continues the trajectory bit for bit (positions, velocities,
accelerations, types after head conversions, delta_sumr, observables),
across a centre-of-mass correction step; per replica for an ensemble.
Malformed or mismatched files are refused with a status code."""
import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.capi import MDPD_ERR_ARG, MDPD_ERR_PARSE, MdpdError

pytestmark = pytest.mark.gpu

KEYS = ("pos", "vel", "vel_buf", "acc", "press", "delta", "id", "prop", "amp_id", "unit")


def _chains():
    box = (10.0, 10.0, 8.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=4)
    p.cf_s, p.cf_b = 50.0, 10.0
    p.conv_p, p.conv_from, p.conv_to = 0.02, configs.HYPHIL, configs.SOLVENT
    return p, configs.random_chains(600, box, 3, n_solvent=400)


def test_checkpoint_restart_is_exact(engine_mod, tmp_path):
    p, s = _chains()
    n = len(s["pos"])

    def fresh():
        e = engine_mod.Engine(p, capacity=n + 64)
        e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        e.set_topology(s["topo"])
        e.set_clock(90, p.pair_seed)
        return e

    ck = tmp_path / "state.ckp"
    a = fresh()
    a.run(6)
    a.checkpoint_save(ck)
    a.run(9)  # crosses the CM correction at time 100
    ga, oa = a.download(), a.observe()
    a.close()
    b = engine_mod.Engine(p, capacity=n + 64)
    b.checkpoint_load(ck)
    assert b.get_clock() == 96 and b.num_beads() == n
    b.run(9)
    gb, ob = b.download(), b.observe()
    b.close()
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k
    for k in ("kin_temp", "pressure", "diffusion", "bonded_vir"):
        assert np.array_equal(oa[k], ob[k]), k
    assert oa["n_amp"] == ob["n_amp"] == 600


def test_checkpoint_restart_ensemble(engine_mod, tmp_path):
    params, systems = [], []
    for r, (box, nb) in enumerate((((8.0, 8.0, 8.0), 1536), ((9.0, 8.0, 10.0), 2160))):
        params.append(configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=10 + r))
        systems.append(configs.random_fluid(nb, box, 20 + r))
    sizes = [len(x["pos"]) for x in systems]
    st = configs.concat_systems(systems)
    a = engine_mod.Engine(params, n_beads=sizes)
    a.upload(st["pos"], st["vel"], st["ids"], st["prop"], st["amp_id"], st["unit"])
    a.set_clock(0, 0)
    a.run(5)
    ck = tmp_path / "ens.ckp"
    a.checkpoint_save(ck)
    a.run(7)
    ga = a.download()
    a.close()
    b = engine_mod.Engine(params, n_beads=sizes)
    b.checkpoint_load(ck)
    b.run(7)
    gb = b.download()
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k
    b.close()
    # a context of another layout refuses it
    c = engine_mod.Engine(params[0], capacity=sum(sizes))
    with pytest.raises(MdpdError) as ex:
        c.checkpoint_load(ck)
    assert ex.value.code == MDPD_ERR_ARG
    bad = tmp_path / "bad.ckp"
    bad.write_bytes(b"NOTACKPT" + bytes(64))
    with pytest.raises(MdpdError) as ex:
        c.checkpoint_load(bad)
    assert ex.value.code == MDPD_ERR_PARSE
    trunc = tmp_path / "trunc.ckp"
    trunc.write_bytes(ck.read_bytes()[:200])
    with pytest.raises(MdpdError) as ex:
        c.checkpoint_load(trunc)
    assert ex.value.code in (MDPD_ERR_ARG, MDPD_ERR_PARSE)
    c.close()
