"""GPU parity of the step path against the FP64 oracle (same seeded inputs).

Tolerances: pair sets and Saru u32 streams bit-exact; densities 2e-6
relative; forces |F_gpu - F_oracle| / sum_j |F_ij| <= 1e-5 per bead (the
pair-magnitude-normalised criterion); drift, kick and bonded FP64 arithmetic
bit-exact given the same FP32-stored acceleration.
"""
import numpy as np
import pytest

from implicit_dpd_amd import configs

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _saru_fixture():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "saru_ref_vectors.txt")
    U, N = [], []
    for line in open(path):
        f = line.split()
        if f[0] == "U":
            U.append([int(x) for x in f[1:]])
        elif f[0] == "N":
            N.append([int(x) for x in f[1:4]] + [float.fromhex(f[4])])
    return np.array(U, dtype=np.uint64), N


def test_saru_device_bitexact(engine_mod):
    U, N = _saru_fixture()
    s = U[:, :3].astype(np.uint32)
    u, _ = engine_mod.saru_device(s[:, 0], s[:, 1], s[:, 2])
    assert np.array_equal(u, U[:, 3:].astype(np.uint32))
    ns = np.array([r[:3] for r in N], dtype=np.uint32)
    ref = np.array([r[3] for r in N])
    _, f = engine_mod.saru_device(ns[:, 0], ns[:, 1], ns[:, 2])
    err = np.abs(f.astype(np.float64) - ref)
    assert err.max() < 2e-6 * max(1.0, np.abs(ref).max()), err.max()


def test_saru_nrml_accuracy_many(engine_mod, oracle):
    rng = np.random.default_rng(5)
    s = rng.integers(0, 2**32, size=(200000, 3), dtype=np.uint64).astype(np.uint32)
    _, f = engine_mod.saru_device(s[:, 0], s[:, 1], s[:, 2])
    _, _, ref = oracle.saru_batch(s[:, 0], s[:, 1], s[:, 2])
    err = np.abs(f.astype(np.float64) - ref)
    # FP32 Box-Muller: absolute error bounded by a few FP32 ulps of |x| <= 6.5
    assert err.max() < 4e-6, err.max()
    assert abs(f.mean()) < 0.01 and abs(f.std() - 1.0) < 0.01


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
    box = (8.0, 9.0, 8.0)
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=500, dt=0.005)
    out.append(("reference", p, configs.random_chains(500, box, 4, n_solvent=300)))
    return out


SYSTEMS = _systems()


def _engine(engine_mod, p, s, time=7):
    e = engine_mod.Engine(p, capacity=len(s["pos"]) + 64)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(s["topo"]):
        e.set_topology(s["topo"])
    e.set_clock(time, p.pair_seed)
    return e


@pytest.mark.parametrize("name,p,s", SYSTEMS, ids=[x[0] for x in SYSTEMS])
def test_pair_sets_bitexact(engine_mod, oracle, name, p, s):
    e = _engine(engine_mod, p, s)
    e.rebuild()
    e.density()
    for cut in sorted({1.0, p.rd}):
        gi, gj = e.pairs(cut)
        oi, oj = oracle.pairs(p, s["pos"], cut, brute=True)
        assert len(gi) == len(oi), (cut, len(gi), len(oi))
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    e.close()


@pytest.mark.parametrize("name,p,s", SYSTEMS, ids=[x[0] for x in SYSTEMS])
def test_density_and_force(engine_mod, oracle, name, p, s):
    time = 11
    e = _engine(engine_mod, p, s, time)
    e.rebuild()
    e.density()
    e.force()
    g = e.download()
    dens = oracle.density(p, s["pos"], s["prop"])
    nd = p.n_dens
    scale = max(1.0, float(np.abs(dens).max()))
    assert np.abs(g["dens"][:, :nd] - dens[:, :nd]).max() < 2e-6 * scale
    acc, press = oracle.force(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + time)
    fa = oracle.force_abs(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + time)
    err = np.abs(g["acc"] - acc).max(axis=1) / np.maximum(fa, 1e-3)
    assert err.max() < FORCE_TOL, (name, err.max(), int(err.argmax()))
    perr = np.abs(g["press"] - press).max(axis=1) / np.maximum(fa, 1e-3)
    assert perr.max() < FORCE_TOL
    # Newton's third law: the pair noise is symmetric, so total force ~ 0
    assert np.abs(g["acc"].sum(axis=0)).max() < 1e-5 * fa.sum()
    e.close()


def test_drift_kick_bitexact(engine_mod, oracle):
    name, p, s = SYSTEMS[1]
    rng = np.random.default_rng(9)
    n = len(s["pos"])
    acc = (rng.normal(size=(n, 3)) * 50).astype(np.float32).astype(np.float64)
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], acc=acc)
    e.drift()
    g = e.download()
    pos, vel, vb, dl = oracle.drift(p, s["pos"], s["vel"], s["vel"], acc, np.zeros((n, 3)))
    assert np.array_equal(g["pos"], pos)
    assert np.array_equal(g["vel"], vel)
    assert np.array_equal(g["vel_buf"], vb)
    assert np.array_equal(g["delta"], dl)
    e.kick()
    g2 = e.download()
    v2 = oracle.kick(p, vel, vb, acc)
    assert np.array_equal(g2["vel"], v2)
    e.remove_cm_drift()
    g3 = e.download()
    v3 = oracle.remove_cm_drift(v2)
    assert np.abs(g3["vel"] - v3).max() < 1e-12
    assert np.abs(g3["vel"].mean(axis=0)).max() < 1e-12
    e.close()


def test_bonded_matches_oracle(engine_mod, oracle):
    name, p, s = SYSTEMS[2]
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    vir = e.bonded()
    g = e.download()
    cf_bend = np.full(len(s["topo"]) // 4, p.cf_b)
    acc, ovir = oracle.bonded(p, s["topo"], cf_bend, s["pos"])
    # FP64 forces rounded once to the FP32 acceleration store
    assert np.abs(g["acc"] - acc).max() <= 1e-6 * np.abs(acc).max()
    assert np.allclose(vir, ovir, rtol=1e-12, atol=1e-12)
    e.close()


@pytest.mark.parametrize("name,p,s", SYSTEMS, ids=[x[0] for x in SYSTEMS])
def test_steps_track_oracle(engine_mod, oracle, name, p, s):
    """Five steps of main.cpp:167-196 on both sides; positions agree to the
    FP32 force error propagated through the integrator."""
    n = len(s["pos"])
    e = _engine(engine_mod, p, s, time=1)
    o = oracle.System(p, n + 64, len(s["topo"]) // 4 + 1)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(s["topo"]):
        o.set_topology(s["topo"])
    o.set_clock(1)
    e.run(5)
    o.run(5)
    g = e.download()
    r = o.get()
    dx = g["pos"] - r["pos"]
    dx -= np.round(dx / np.array(p.box[:])) * np.array(p.box[:])
    # per step: |dx| ~ dt^2 |dF| with |dF| ~ 1e-5 sum|F_ij|, compounded over 5 steps
    assert np.abs(dx).max() < 1e-6, np.abs(dx).max()
    assert np.abs(g["vel"] - r["vel"]).max() < 1e-3
    go = e.observe()
    ro = o.observe()
    assert np.allclose(go["kin_temp"], ro.kin_temp[:], rtol=1e-6)
    assert np.allclose(go["pressure"], ro.pressure[:], rtol=1e-5, atol=1e-6)
    e.close()


def test_run_is_deterministic_and_graph_matches_eager(engine_mod):
    name, p, s = SYSTEMS[2]
    res = []
    for graph in (1, 1, 0):
        e = _engine(engine_mod, p, s, time=3)
        e.set_tuning(use_graph=graph)
        e.run(7)
        res.append(e.download())
        e.close()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(res[0][k], res[1][k])
        assert np.array_equal(res[0][k], res[2][k])


def test_split_runs_and_cm_steps_are_bit_identical(engine_mod):
    """Cell keys of a step come from the previous step's force pass, except
    on a run's first step and after a CM-drift removal, which bin the beads
    in a key pass of their own: one 12-step run, the same 12 steps as runs of
    5 + 4 + 3, and the kernel-by-kernel launches all give the same
    trajectory bit for bit, with a CM-drift removal every third step; so do
    runs whose step graphs were recorded beforehand by mdpd_prepare_run (for
    the run's own length, and for a longer run than the one that follows)."""
    name, p0, s = SYSTEMS[2]
    p = type(p0).from_buffer_copy(p0)
    p.cm_corr = 3
    res = []
    for graph, split, prep in ((1, (12,), 0), (1, (5, 4, 3), 0), (0, (12,), 0), (0, (1, 2, 9), 0),
                               (1, (12,), 1), (1, (5, 4, 3), 1), (1, (7, 5), 9)):
        e = _engine(engine_mod, p, s, time=4)
        e.set_tuning(use_graph=graph)
        for k in split:
            if prep:
                e.prepare_run(k * prep)
            e.run(k)
        res.append(e.download())
        e.close()
    for r in res[1:]:
        for k in ("pos", "vel", "vel_buf", "acc", "delta"):
            assert np.array_equal(res[0][k], r[k]), k
