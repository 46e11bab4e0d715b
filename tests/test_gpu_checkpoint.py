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


def _layout(n, na, nr=1):
    """byte offsets of the arrays of a version-2 checkpoint"""
    off = 8 + 24 + 4 * (nr + 1) + 4 * nr * 4 + 8 * nr
    o = {"fprint": off - 8 * nr, "pos": off}
    o["f32"] = off + 8 * 12 * n
    o["u32"] = o["f32"] + 4 * 6 * n
    for q, k in enumerate(("id", "prop", "amp_id", "unit", "orig", "chain")):
        o[k] = o["u32"] + 4 * q * n
    o["topo"] = o["u32"] + 4 * 6 * n
    o["cf"] = o["topo"] + 16 * na
    o["amp_rep"] = o["cf"] + 8 * na
    o["end"] = o["amp_rep"] + 4 * na
    return o


def test_corrupt_checkpoints_are_refused_before_any_copy(engine_mod, tmp_path):
    """Every word a kernel uses as an index is checked on the host: a file
    with a bad type, chain word, topology entry or slot map, a position
    outside the box or not finite, extra or missing bytes, another box, or
    the old format is refused, and the context keeps the state it had."""
    import struct
    p, s = _chains()
    n, na = len(s["pos"]), 600
    a = engine_mod.Engine(p, capacity=n + 64)
    a.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    a.set_topology(s["topo"])
    a.set_clock(10, p.pair_seed)
    a.run(4)
    ck = tmp_path / "good.ckp"
    a.checkpoint_save(ck)
    raw = ck.read_bytes()
    o = _layout(n, na)
    assert len(raw) == o["end"]
    # the receiving context holds another state, which must survive
    b = engine_mod.Engine(p, capacity=n + 64)
    b.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    b.set_topology(s["topo"])
    b.set_clock(3, p.pair_seed)
    b.run(2)
    before = b.download()

    def u32(buf, off, v):
        buf[off:off + 4] = struct.pack("<I", v)

    chain = np.frombuffer(raw[o["chain"]:o["chain"] + 4 * n], np.uint32)
    cb = np.nonzero(chain != 0xFFFFFFFF)[0][:2]     # two chain beads

    def f64(buf, off, v):
        buf[off:off + 8] = struct.pack("<d", v)

    cases = {
        "type": lambda q: u32(q, o["prop"] + 4 * 5, 7),
        "chain_range": lambda q: u32(q, o["chain"] + 4 * 10, (na + 3) << 2),
        "chain_repeat": lambda q: u32(q, o["chain"] + 4 * int(cb[0]), int(chain[cb[1]])),
        "topo": lambda q: u32(q, o["topo"] + 4 * 7, n + 5),
        "orig": lambda q: q.__setitem__(slice(o["orig"], o["orig"] + 4),
                                        bytes(q[o["orig"] + 4:o["orig"] + 8])),
        "nan": lambda q: f64(q, o["pos"] + 8 * 3, float("nan")),
        "outside": lambda q: f64(q, o["pos"] + 8 * 4, 10.5),
        "amp_rep": lambda q: u32(q, o["amp_rep"] + 4 * 2, 3),
        "fingerprint": lambda q: q.__setitem__(slice(o["fprint"], o["fprint"] + 8), bytes(8)),
        "version1": lambda q: u32(q, 8, 1),
        "flags": lambda q: u32(q, 8 + 20, 2),
        "stream_missing": lambda q: u32(q, 8 + 20, 1),
        "extra": lambda q: q.extend(b"\0" * 4),
        "short": lambda q: q.__delitem__(slice(len(q) - 4, len(q))),
    }
    for name, mutate in cases.items():
        q = bytearray(raw)
        mutate(q)
        bad = tmp_path / f"{name}.ckp"
        bad.write_bytes(bytes(q))
        with pytest.raises(MdpdError) as ex:
            b.checkpoint_load(bad)
        assert ex.value.code in (MDPD_ERR_ARG, MDPD_ERR_PARSE), name
        after = b.download()
        for k in KEYS:
            assert np.array_equal(before[k], after[k]), (name, k)
    # a context of another box refuses the good file
    p2, _ = _chains()
    p2.box[2] = 8.5
    c = engine_mod.Engine(p2, capacity=n + 64)
    with pytest.raises(MdpdError) as ex:
        c.checkpoint_load(ck)
    assert ex.value.code == MDPD_ERR_ARG and "parameters" in str(ex.value)
    # and the good file still loads into b and continues a's run
    b.checkpoint_load(ck)
    a.run(3)
    b.run(3)
    ga, gb = a.download(), b.download()
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k
    for e in (a, b, c):
        e.close()


def test_checkpoint_restores_chem_seed(engine_mod, tmp_path):
    """The head conversion draws on chem_seed: a checkpoint loaded into a
    context created with another chem_seed continues bit for bit."""
    p, s = _chains()
    n = len(s["pos"])
    a = engine_mod.Engine(p, capacity=n)
    a.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    a.set_topology(s["topo"])
    a.set_clock(20, p.pair_seed)
    a.run(3)
    ck = tmp_path / "cs.ckp"
    a.checkpoint_save(ck)
    a.run(12)
    ga = a.download()
    p2, _ = _chains()
    p2.chem_seed = (p.chem_seed + 12345) & 0xFFFFFFFF
    b = engine_mod.Engine(p2, capacity=n)
    b.checkpoint_load(ck)
    b.run(12)
    gb = b.download()
    assert (ga["prop"] != s["prop"]).sum() > 0      # some heads converted
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()


def test_checkpoint_load_after_chemistry_growth(engine_mod, tmp_path):
    """A context whose chemistry appended beads loads an earlier, smaller
    checkpoint and continues bit for bit as another context (also past its
    first chemistry call) that loads the same file; the chemistry then runs
    again from there with the same events in both."""
    box = (8.0, 9.0, 8.0)
    n_amp = 500
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    n = len(s["pos"])
    kw = dict(p_thresld=0.05, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.1,
              max_amp_num=10 ** 6, core_ptcl_id=np.array([0, 40, 80], np.uint32))

    def fresh():
        e = engine_mod.Engine(p, capacity=n + 4 * 600)
        e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        e.set_topology(s["topo"], core_amp_id=[0, 10, 20])
        e.set_clock(50, p.pair_seed)
        return e

    ck = tmp_path / "grow.ckp"
    a = fresh()
    a.run(3)
    a.checkpoint_save(ck)
    a.set_chemistry(beg_chem=0, **kw)
    a.run(6)
    assert a.observe()["n_amp"] > n_amp and a.num_beads() > n
    a.checkpoint_load(ck)
    assert a.observe()["n_amp"] == n_amp and a.num_beads() == n
    b = fresh()
    b.set_chemistry(beg_chem=10 ** 9, **kw)
    b.checkpoint_load(ck)
    for e in (a, b):
        e.set_chemistry(beg_chem=56, **kw)
        e.run(7)
    ga, gb = a.download(), b.download()
    assert a.observe()["n_amp"] == b.observe()["n_amp"] > n_amp
    assert a.get_clock() == b.get_clock() == 60
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k
    a.close()
    b.close()
