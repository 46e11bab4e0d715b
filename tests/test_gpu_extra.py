"""GPU checks of the host driver, the autotuner, head conversion, the
chemistry event and the observables against the FP64 oracle.

Tolerances: the driver's trajectory follows the oracle's to the FP32 force
error carried through the integrator (positions 1e-5 after 10 steps);
autotuned launches are bit-identical to the default ones; head conversion and
the chemistry's bead records and topology are exact (keyed Saru draws, FP64
geometry with the same operation order, positions to 1e-12); bead-count
profiles are exact.
"""
import os
import subprocess

import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.fpdpd import FPDPD_DTYPE, read_xyz, write_xyz

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "implicit_dpd_amd", "bin", "mdpd_main")


def _seq_mean(x):
    """Mean with a sequential sum (the reference's loop order)."""
    return np.array([np.add.accumulate(x[:, a])[-1] for a in range(3)]) / len(x)


def test_host_driver_matches_oracle(engine_mod, oracle, tmp_path):
    """mdpd_main (main.cpp's control flow over the C++ facade) on a
    run_param.txt + init_config.xyz directory against the oracle's steps."""
    box = (8.0, 9.0, 8.0)
    n_amp, n_sol, steps = 500, 300, 10
    s = configs.random_chains(n_amp, box, 4, n_solvent=n_sol)
    d = str(tmp_path)
    configs.write_run_param(d, box, n_amp, sol_num=n_sol, dt=0.005, all_time=steps, step_mic=5,
                            step_mac=3, core=(0,))
    n = len(s["pos"])
    rec = np.zeros(n, dtype=FPDPD_DTYPE)
    rec["pos"], rec["vel"], rec["vel_buf"] = s["pos"], s["vel"], s["vel"]
    rec["id"], rec["prop"], rec["amp_id"], rec["unit"] = s["ids"], s["prop"], s["amp_id"], s["unit"]
    write_xyz(os.path.join(d, "init_config.xyz"), rec, 0)
    print(f"[driver] {DRIVER} {d} --seed 77", flush=True)
    r = subprocess.run([DRIVER, d, "--seed", "77"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "m_seed 77" in r.stdout

    # the oracle from the same file, with LoadParticleConfig's CM adjustments
    p, _ = engine_mod.read_run_param(d)
    p.pair_seed = 77
    r0, t0 = read_xyz(os.path.join(d, "init_config.xyz"))
    assert t0 == 0
    L = np.array(box)
    vel = r0["vel"] - _seq_mean(r0["vel"])
    pos = r0["pos"] + (L * 0.5 - _seq_mean(r0["pos"]))
    pos = pos - np.floor(pos * (1.0 / L)) * L
    o = oracle.System(p, n + 64, n_amp + 1)
    o.set(pos, vel, r0["id"], r0["prop"], r0["amp_id"], r0["unit"], vel_buf=r0["vel_buf"],
          acc=r0["acc"])
    o.set_topology(s["topo"], core_amp_id=[0])
    o.set_clock(0)

    kin = np.loadtxt(os.path.join(d, "kin_temp.txt"))
    prs = np.loadtxt(os.path.join(d, "pressure.txt"))
    dif = np.loadtxt(os.path.join(d, "diffus.txt"), comments="#")
    assert kin.shape == (4, 4) and prs.shape == (4, 3) and dif.shape == (4, 3)
    # macro observations at times 3, 6, 9 follow steps 0..t
    done = 0
    for row, t in enumerate((3, 6, 9), start=1):
        o.run(t + 1 - done)
        done = t + 1
        ro = o.observe(sol_num=n_sol)
        assert np.allclose(kin[row], ro.kin_temp[:], rtol=2e-5), (t, kin[row], ro.kin_temp[:])
        assert np.allclose(prs[row], ro.pressure[:], rtol=1e-4, atol=1e-4), (t, prs[row])
        assert np.allclose(dif[row], ro.diffusion[:], rtol=1e-4, atol=1e-9), (t, dif[row])
    o.run(steps - done)
    fin, tf = read_xyz(os.path.join(d, "fin_config.xyz"))
    assert tf == steps
    ref = o.get()
    order = np.argsort(ref["id"])
    assert np.array_equal(fin["id"], ref["id"][order])
    dx = fin["pos"] - ref["pos"][order]
    dx -= np.round(dx / L) * L
    assert np.abs(dx).max() < 1e-5, np.abs(dx).max()
    assert np.abs(fin["vel"] - ref["vel"][order]).max() < 1e-2
    traj = open(os.path.join(d, "traject.xyz")).read().split("\n")
    assert traj[0] == str(n) and traj[1] == "time 0"
    assert [int(x) for x in open(os.path.join(d, "num_amp.txt")).read().split()] == [n_amp] * 4


@pytest.mark.parametrize("max_extra,mt", [(1000, False), (1000, True), (1, False)])
def test_host_driver_chemistry_matches_oracle(engine_mod, oracle, tmp_path, max_extra, mt):
    """mdpd_main --chem: the chemistry events run inside the driver's
    mdpd_run calls from --beg-chem on; the bead records (ids, types, chains)
    and amphiphile counts equal the oracle's step + RandomChemEvent loop, and
    with max_amp_num one above the start the run ends at the first event
    that inserts, as main.cpp:188 breaks (no observation of that step,
    fin_config.xyz stamped with its time).  With --chem-mt the draws are the
    reference's MT19937 stream seeded with 0, as the oracle's reference-MT
    mode draws them."""
    box = (8.0, 9.0, 8.0)
    n_amp, n_sol, steps, beg = 500, 300, 10, 3
    s = configs.random_chains(n_amp, box, 4, n_solvent=n_sol)
    d = str(tmp_path)
    kw = dict(p_thresld=0.08, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.1)
    configs.write_run_param(d, box, n_amp, sol_num=n_sol, dt=0.005, all_time=steps, step_mic=5,
                            step_mac=3, core=(0,), max_amp_num=n_amp + max_extra, **kw)
    n = len(s["pos"])
    rec = np.zeros(n, dtype=FPDPD_DTYPE)
    rec["pos"], rec["vel"], rec["vel_buf"] = s["pos"], s["vel"], s["vel"]
    rec["id"], rec["prop"], rec["amp_id"], rec["unit"] = s["ids"], s["prop"], s["amp_id"], s["unit"]
    write_xyz(os.path.join(d, "init_config.xyz"), rec, 0)
    cmd = [DRIVER, d, "--seed", "77", "--chem", "--beg-chem", str(beg)] + (["--chem-mt"] if mt else [])
    print("[driver] " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p, _ = engine_mod.read_run_param(d)
    p.pair_seed = 77
    r0, _ = read_xyz(os.path.join(d, "init_config.xyz"))
    L = np.array(box)
    vel = r0["vel"] - _seq_mean(r0["vel"])
    pos = r0["pos"] + (L * 0.5 - _seq_mean(r0["pos"]))
    pos = pos - np.floor(pos * (1.0 / L)) * L
    o = oracle.System(p, n + 4 * 400, n_amp + 400)
    o.set(pos, vel, r0["id"], r0["prop"], r0["amp_id"], r0["unit"], vel_buf=r0["vel_buf"],
          acc=r0["acc"])
    o.set_topology(s["topo"], core_amp_id=[0])
    if mt:
        o.set_mt(0)
    cp = engine_mod.MdpdChemParams(kw["p_thresld"], kw["eps"], kw["influ_rad"], kw["influ_hei"],
                                   kw["influ_dep"], kw["influ_grd"], n_amp + max_extra, 1,
                                   engine_mod._u(np.zeros(1, np.uint32)))
    amp = [n_amp]            # num_amp.txt: the initial step's observation at time 0
    t_end = steps
    for t in range(steps):
        o.set_clock(t)
        o.run(1)
        if t >= beg and t >= 1:
            o.set_clock(t)
            k, stop = o.chem_event(cp)
            if stop:
                t_end = t
                break
        if t % 3 == 0 and t > 0:
            amp.append(int(o.observe(sol_num=n_sol).n_amp))
    fin, tf = read_xyz(os.path.join(d, "fin_config.xyz"))
    assert tf == t_end
    ref = o.get()
    assert len(fin) == len(ref["id"]) > n
    order = np.argsort(ref["id"])
    for k in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(fin[k], ref[k][order]), k
    dx = fin["pos"] - ref["pos"][order]
    dx -= np.round(dx / L) * L
    assert np.abs(dx).max() < 1e-5, np.abs(dx).max()
    got = [int(x) for x in open(os.path.join(d, "num_amp.txt")).read().split()]
    assert got == amp, (got, amp)
    if max_extra == 1:
        assert t_end < steps - 1 and "simulation will stop" in r.stdout


def test_autotuned_launches_are_bit_identical(engine_mod):
    box = (10.0, 10.0, 8.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=4)
    p.cf_s, p.cf_b = 50.0, 10.0
    s = configs.random_chains(600, box, 3)
    res = []
    for mode in ("default", "manual", "autotune", "manual256"):
        e = engine_mod.Engine(p, capacity=len(s["pos"]))
        e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        e.set_topology(s["topo"])
        e.set_clock(5, p.pair_seed)
        if mode == "manual":
            e.set_tuning(block_scan=64, block_force=128, nlist_width=128)
        elif mode == "manual256":
            e.set_tuning(block_scan=256, block_force=256)
        elif mode == "autotune":
            t = e.autotune(3)
            assert t.block_scan in (64, 128, 256) and t.block_force in (64, 128, 256)
            assert e.get_clock() == 5
        e.run(6)
        res.append(e.download())
        e.close()
    for k in ("pos", "vel", "acc", "dens"):
        assert np.array_equal(res[0][k], res[1][k]), k
        assert np.array_equal(res[0][k], res[2][k]), k
        assert np.array_equal(res[0][k], res[3][k]), k


@pytest.mark.parametrize("box", [(10.0, 10.0, 8.0), (3.5, 6.0, 6.0)])
def test_autotune_with_skin_choice(engine_mod, oracle, box):
    """mdpd_autotune_ex with MDPD_TUNE_SKIN: a valid tuning (a skin the box
    allows; the small box refuses 0.2 and 0.3 and those candidates are
    skipped), the state left as it was, and exact pair sets afterwards."""
    p = configs.groot_warren_params(box, pair_seed=2)
    n = int(3 * box[0] * box[1] * box[2])
    s = configs.random_fluid(n, box, 4)
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"])
    e.set_clock(9, p.pair_seed)
    e.run(20)
    before = e.download()
    t = e.autotune(3, skin=True)
    assert t.block_scan in (64, 128, 256) and t.block_force in (64, 128, 256)
    allowed = [0.0] + [x for x in (0.1, 0.2, 0.3) if 3 * (1 + x) <= min(box) + 1e-9]
    assert min(abs(t.verlet_skin - x) for x in allowed) < 1e-6, (t.verlet_skin, allowed)
    assert e.get_clock() == 29
    after = e.download()
    for k in ("pos", "vel", "id", "prop"):
        assert np.array_equal(before[k], after[k]), k
    for _ in range(3):
        e.run(4)
        g = e.download()
        gi, gj = e.pairs(1.0)
        oi, oj = oracle.pairs(p, g["pos"], 1.0, brute=True)
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    e.close()


def test_head_conversion_matches_oracle(engine_mod, oracle):
    box = (12.0, 12.0, 12.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=9)
    p.conv_p, p.conv_from, p.conv_to = 0.05, configs.HYPHIL, configs.SOLVENT
    p.chem_seed = 4242
    s = configs.random_chains(4000, box, 8)
    n = len(s["pos"])
    total = 0
    for t in (17, 18, 1000):
        e = engine_mod.Engine(p, capacity=n)
        e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        e.set_topology(s["topo"])
        e.set_clock(t, p.pair_seed)
        e.head_conversion()
        g = e.download()
        e.close()
        o = oracle.System(p, n, 4001)
        o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
        o.set_topology(s["topo"])
        o.set_clock(t)
        k = o.head_conversion()
        r = o.get()
        assert np.array_equal(g["prop"], r["prop"]), t
        changed = int((g["prop"] != s["prop"]).sum())
        assert changed == k
        # only heads of type conv_from change, into conv_to
        idx = np.nonzero(g["prop"] != s["prop"])[0]
        assert np.all(s["unit"][idx] == 0) and np.all(g["prop"][idx] == configs.SOLVENT)
        total += changed
    # binomial count over the three draws: 3 * 4000 heads at p = 0.05
    m, sd = 3 * 4000 * 0.05, np.sqrt(3 * 4000 * 0.05 * 0.95)
    assert abs(total - m) < 5 * sd, (total, m)


@pytest.mark.parametrize("skin", [0.0, 0.3])
def test_run_head_conversion_matches_oracle(engine_mod, oracle, skin):
    """Head conversion inside mdpd_run (the force pass's epilogue, or its own
    pass in skin mode): the types after six steps equal six keyed
    conversions of the oracle at the same clock values (types do not depend
    on the trajectory)."""
    box = (12.0, 12.0, 12.0)
    p = configs.warren_params(box, A=configs.MEMBRANE_A, B=25.0, rd=0.75, n_species=3, pair_seed=9)
    p.conv_p, p.conv_from, p.conv_to = 0.05, configs.HYPHIL, configs.SOLVENT
    p.chem_seed = 4242
    s = configs.random_chains(4000, box, 8)
    n = len(s["pos"])
    e = engine_mod.Engine(p, capacity=n)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(40, p.pair_seed)
    if skin:
        e.set_tuning(verlet_skin=skin)
    e.run(4)
    e.run(2)
    g = e.download()
    assert e.get_clock() == 46
    e.close()
    o = oracle.System(p, n, 4001)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"])
    for t in range(40, 46):
        o.set_clock(t)
        o.head_conversion()
    r = o.get()
    assert np.array_equal(g["prop"], r["prop"])
    assert int((g["prop"] != s["prop"]).sum()) > 0


def test_chem_event_matches_oracle(engine_mod, oracle):
    box = (8.0, 9.0, 8.0)
    n_amp = 500
    p, _ = configs.reference_params(box, pair_seed=5, init_amp_num=n_amp, dt=0.005)
    s = configs.random_chains(n_amp, box, 4, n_solvent=300)
    n = len(s["pos"])
    cap = n + 4 * 400
    core_ids = np.array([0, 40, 80], np.uint32)  # bead ids of three core heads
    kw = dict(p_thresld=0.5, eps=0.2, influ_rad=3.0, influ_hei=0.8, influ_dep=4.0, influ_grd=0.0,
              max_amp_num=n_amp + 300)
    e = engine_mod.Engine(p, capacity=cap)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"], core_amp_id=[0, 10, 20])
    e.set_clock(123, p.pair_seed)
    e.rebuild()
    e.density()
    stop = e.chem_event(core_ptcl_id=core_ids, **kw)
    g = e.download()
    o = oracle.System(p, cap, n_amp + 400)
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    o.set_topology(s["topo"], core_amp_id=[0, 10, 20])
    o.set_clock(123)
    cp = engine_mod.MdpdChemParams(kw["p_thresld"], kw["eps"], kw["influ_rad"], kw["influ_hei"],
                                   kw["influ_dep"], kw["influ_grd"], kw["max_amp_num"], 3,
                                   engine_mod._u(core_ids))
    k, ostop = o.chem_event(cp)
    r = o.get()
    assert stop == ostop
    assert len(g["pos"]) == len(r["pos"]) == n + 4 * k, (len(g["pos"]), len(r["pos"]), k)
    assert k > 0
    for key in ("id", "prop", "amp_id", "unit"):
        assert np.array_equal(g[key], r[key]), key
    assert np.abs(g["pos"] - r["pos"]).max() < 1e-12
    # old beads untouched; new beads' Box-Muller velocities agree to the
    # last-bit differences of the device and host FP64 log/cos
    assert np.array_equal(g["vel"][:n], r["vel"][:n])
    assert np.allclose(g["vel"][n:], r["vel"][n:], rtol=1e-13, atol=1e-14)
    obs = e.observe()
    assert obs["n_amp"] == n_amp + k
    e.close()


def test_density_profile_counts(engine_mod):
    box = (9.0, 8.0, 10.0)
    p = configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=3)
    s = configs.random_fluid(4320, box, 2, zlim=(2.5, 7.5))
    e = engine_mod.Engine(p, capacity=len(s["pos"]))
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_clock(0, p.pair_seed)
    e.run(3)
    g = e.download()
    for axis, nb in ((2, 20), (0, 9)):
        prof = e.density_profile(axis, nb)[0]
        b = np.floor(g["pos"][:, axis] * (1.0 / box[axis]) * nb).astype(int)
        b = np.clip(b, 0, nb - 1)
        assert np.array_equal(prof, np.bincount(b, minlength=nb).astype(float)), axis
    e.close()
