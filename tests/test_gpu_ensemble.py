"""Replica ensembles and state transfer on the GPU.

- Packing invariance: every replica of an ensemble context gives, bit for
  bit, what the same replica gives alone in a single-system context (cells,
  sort order and per-bead sums are per replica), and agrees with the oracle.
- FPDPD records (ptcl_class.hpp:29-36) round-trip through upload/download
  with the upload order restored after the engine's cell-order reordering.
"""
import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.engine import Engine
from implicit_dpd_amd.fpdpd import FPDPD_DTYPE

pytestmark = pytest.mark.gpu


def small_ensemble():
    params, systems = configs.ensemble(seed=5, n_rep=5, nmin=700, nmax=3000)
    return params, systems


def test_ensemble_replicas_equal_single_runs():
    params, systems = small_ensemble()
    allb = configs.concat_systems(systems)
    sizes = [len(s["pos"]) for s in systems]
    e = Engine(params, n_beads=sizes)
    e.upload(allb["pos"], allb["vel"], allb["ids"], allb["prop"])
    e.set_clock(3, 0)
    e.run(6)
    ge = e.download()
    obs = e.observe_replicas()
    off = 0
    for r, (p, s) in enumerate(zip(params, systems)):
        one = Engine(p, capacity=len(s["pos"]))
        one.upload(s["pos"], s["vel"], s["ids"], s["prop"])
        one.set_clock(3, p.pair_seed)
        one.run(6)
        g1 = one.download()
        n = len(s["pos"])
        for k in ("pos", "vel", "acc", "dens"):
            assert np.array_equal(ge[k][off:off + n], g1[k]), (r, k)
        o1 = one.observe()
        assert np.array_equal(obs[r]["kin_temp"], o1["kin_temp"])
        assert np.array_equal(obs[r]["pressure"], o1["pressure"])
        off += n
        one.close()
    e.close()


def test_ensemble_replica_tracks_oracle(oracle):
    params, systems = small_ensemble()
    allb = configs.concat_systems(systems)
    sizes = [len(s["pos"]) for s in systems]
    e = Engine(params, n_beads=sizes)
    e.upload(allb["pos"], allb["vel"], allb["ids"], allb["prop"])
    e.set_clock(1, 0)
    e.rebuild()
    e.density()
    e.force()
    g = e.download()
    off = 0
    for p, s in zip(params, systems):
        n = len(s["pos"])
        dens = oracle.density(p, s["pos"], s["prop"])
        acc, _ = oracle.force(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + 1)
        fa = oracle.force_abs(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + 1)
        err = np.abs(g["acc"][off:off + n] - acc).max(axis=1) / np.maximum(fa, 1e-3)
        assert err.max() < 1e-5
        off += n
    e.close()


def test_fpdpd_roundtrip_after_reordering():
    box = (9.0, 9.0, 9.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, n_species=3, pair_seed=2)
    p.cf_s, p.cf_b = 50.0, 10.0
    s = configs.random_chains(400, box, 8)
    n = len(s["pos"])
    rec = np.zeros(n, FPDPD_DTYPE)
    rec["id"], rec["prop"], rec["amp_id"], rec["unit"] = s["ids"], s["prop"], s["amp_id"], s["unit"]
    rec["pos"], rec["vel"], rec["vel_buf"] = s["pos"], s["vel"], s["vel"]
    rec["delta_sumr"] = 0.25
    e = Engine(p, capacity=n)
    e.upload_fpdpd(rec)
    back = e.download_fpdpd()
    for k in ("id", "prop", "amp_id", "unit", "pos", "vel", "vel_buf", "delta_sumr"):
        assert np.array_equal(back[k], rec[k]), k
    e.set_topology(s["topo"])
    e.run(4)
    out = e.download_fpdpd()
    # beads come back at their upload slots, ids and chain membership intact
    assert np.array_equal(out["id"], rec["id"])
    assert np.array_equal(out["amp_id"], rec["amp_id"])
    assert np.array_equal(out["unit"], rec["unit"])
    # bonded neighbours stay close (the topology followed the reordering)
    d = out["pos"][1::4] - out["pos"][0::4]
    d -= np.round(d / 9.0) * 9.0
    assert np.linalg.norm(d, axis=1).max() < 1.0
    # nei_cm_pos of a bead with no neighbour of a kind is a 0/0 in the
    # reference; the hydrophilic CM always includes the bead itself
    assert np.isfinite(out["nei_cm_pos"][:, 0]).all()
    e.close()
