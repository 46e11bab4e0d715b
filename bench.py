#!/usr/bin/env python3
"""Headline benchmark: MDPD bead-steps/s at 1M beads on one GPU.

Workload: the implicit-solvent amphiphile membrane model, 262,144 H1T3 chains
(1,048,576 beads) in a 100x100x40 box, Warren MDPD with A_ij in [-40,-20] by
bead type, B = 25, rd = 0.75, harmonic bonds k = 100 l0 = 0.5, bending 10,
head conversion p = 1e-4 per step, dt = 0.005.  The state is a membrane: five
stacked bilayers (config_maker's flat sheet, src/config_maker.cpp:103-137,
repeated along z), so the beads sit at the MDPD liquid density inside the
bilayers.  `--workload membrane_gas` times the same model on seeded random
chains spread over the box instead.  One step is the reference's main-loop
body (src/main.cpp:167-196): drift, cell-list rebuild, density pass, force
pass, bonded forces, kick, centre-of-mass drift removal every 100 steps, head
conversion, clock.

The timed region runs K steps as replayed hipGraphs (the production path)
between two stream synchronisations; `value` is beads * K / (wall time of the
region), and HIP events on the engine stream give its device time.  K more
steps launched kernel by kernel with HIP events around every phase give the
per-kernel times for the roofline of the dominant pair pass (density or
force, whichever takes longer); its HBM traffic, VALU counters and the
rocprofv3 kernel-trace average of the same command come from the committed
summaries in profiles/.  Then K steps with RandomChemEvent inside the
captured steps (chemistry on) are timed the same way.  The CPU baseline is
the FP64 oracle (the reference's per-bead gather loops, OpenMP) timed on a
bounded sample of the same workload with every host thread and with one.
"""
import argparse
import atexit
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MDPD bead-steps/s at 1M beads, % HBM roofline; surrogate vs plain calib runs"
HBM_PEAK_GBS = 8000.0
CHEM_ROOM = 4 * 8192   # beads of headroom for the chemistry-on steps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def density_bytes_per_launch(n, n_pairs, nd, ncell):
    """Algorithmic HBM bytes of one density pass: per bead its sorted
    position (16) and cell (4) read, densities (4 nd) and pair count (4)
    written; per pair-list entry 4 bytes written; the cell offsets (4 per
    cell) read once.  Neighbour data reuse is on-chip and not counted."""
    return n * (16 + 4 + 4 * nd + 4) + 4.0 * n_pairs + 4.0 * (ncell + 1)


def skin_density_bytes_per_launch(n, n_pairs, nd, n_skin):
    """Algorithmic HBM bytes of one density pass of skin mode: per bead its
    position (16) and skin-list count (4) read, densities (4 nd) and pair
    count (4) written; 4 bytes per skin-list entry read and per pair-list
    entry written."""
    return n * (16 + 4 + 4 * nd + 4) + 4.0 * n_skin + 4.0 * n_pairs


def force_bytes_per_launch(n, n_pairs, nd, n_chain_beads, n_heads, ncell, keys=True, conv=True):
    """Algorithmic HBM bytes of one force pass with its fused epilogue (the
    production step's variant).  Pair part: per bead its own 32-byte record
    (position + predicted velocity, the density in its spare lane when
    nd = 1, else 4 nd bytes of densities) and pair count (4); per pair-list
    entry 4 bytes (neighbour records are reused on chip).  Epilogue: the
    chain map (4 per bead) and the chain's FP64 positions (32 per chain
    bead, each read once); the kick's vel_buf read (24), velocity (24) and
    FP32 acceleration (12) written; with keys the master position read (24),
    the cell key written (4) and the cell histogram (4 read + 4 written per
    cell); with the head conversion every bead's unit (4) and the heads'
    type and chain id (8)."""
    b = n * (32 + 4 + (4 * nd if nd > 1 else 0)) + 4.0 * n_pairs
    b += 4 * n + 32 * n_chain_beads
    b += (24 + 24 + 12) * n
    if keys:
        b += (24 + 4) * n + 8.0 * ncell
    if conv:
        b += 4 * n + 8 * n_heads
    return b


VALU_LANE_OPS_PEAK = 256 * 64 * 2.4e9   # CUs x lanes x clock: non-packed lane-ops/s


def profiled_counters(workload):
    """Per-launch HBM bytes, VALU counters and kernel-trace durations of each
    kernel from the committed rocprofv3 summary of this workload (profiles/),
    if present."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.txt")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            if line.startswith("#") or "=" not in line:
                continue
            kv = dict(tok.split("=", 1) for tok in line.split())
            k = kv.get("variant", kv.get("kernel"))
            out[k] = {"traffic_bytes": float(kv["traffic_bytes"]), "avg_ns": float(kv["avg_ns"]),
                      "valu_busy": float(kv["valu_busy"]), "valu_util": float(kv["valu_util"]),
                      "valu_insts": float(kv.get("valu_insts", "nan"))}
    return out


def pick(counters, prefix):
    """The profiled variant of a kernel family with the most time (for the
    force pass the production variant, not the virial-step one, whose last
    template argument is true)."""
    c = [(v["avg_ns"], k, v) for k, v in counters.items()
         if k.startswith(prefix) and not (prefix == "k_force" and k.endswith(",true>"))]
    return max(c)[1:] if c else (None, None)


def step_bytes_per_bead(nd, bonded, rho_mean):
    """Algorithmic HBM bytes per bead-step of the whole step (FP64 master
    pos/vel read + written by drift and kick, FP32 force, ids and types,
    sorted FP32 copies written once and read by both pair passes, slot map,
    densities, bonded gather, cell offsets)."""
    return 264 + 8 * nd + (16 if bonded else 0) + 16.0 / rho_mean


def cpu_baseline(p, s, topo, budget_s, max_steps, threads=None):
    """The FP64 oracle on the same system: steps until budget_s has passed
    (at most max_steps) after one warm-up step, with `threads` OpenMP threads
    (all host threads when None)."""
    from oracle import pyoracle as po
    n = len(s["pos"])
    nthr0 = po.num_threads()
    if threads:
        po.set_num_threads(threads)
    o = po.System(p, n, max(len(topo) // 4, 1))
    o.set(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    if len(topo):
        o.set_topology(topo)
    o.set_clock(0)
    o.run(1)  # warm-up step (page faults, thread start)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        o.run(1)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cores = po.num_threads()
    po.set_num_threads(nthr0)
    return {"value": n * steps / dt, "unit": "bead-steps/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (per-bead gather loops, OpenMP) on the same {n}-bead "
                      f"system, {steps} steps after 1 warm-up step, {dt:.1f} s"}


def calibration_record():
    """The committed record of the surrogate-guided calibration against plain
    sweeps (the metric's third clause), made by implicit_dpd_amd.calibrate:
    the JSON block of profiles/calibration_runs.md."""
    path = os.path.join(ROOT, "profiles", "calibration_runs.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if "```json" not in text:
        return None
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return json.loads(block)


class Workload:
    """One benchmark system: parameters, initial state and its description."""

    def __init__(self, name, params, systems, text, bonded):
        self.name, self.params, self.systems, self.text = name, params, systems, text
        self.bonded = bonded
        self.ensemble = isinstance(params, list)
        plist = params if self.ensemble else [params]
        slist = systems if self.ensemble else [systems]
        self.sizes = [len(x["pos"]) for x in slist]
        self.n = sum(self.sizes)
        self.volume = sum(q.box[0] * q.box[1] * q.box[2] for q in plist)
        self.ncell = sum(max(int(math.floor(q.box[0])), 1) * max(int(math.floor(q.box[1])), 1) *
                         max(int(math.floor(q.box[2])), 1) for q in plist)
        self.nd = plist[0].n_dens

    def start(self, Engine, e=None, warmup=1, skin=0.0):
        """Uploads the initial state (into `e`, or a new engine) and runs
        the warm-up steps."""
        from implicit_dpd_amd import configs
        if e is None:
            # room for the amphiphiles the chemistry-on steps insert
            e = (Engine(self.params, n_beads=self.sizes) if self.ensemble
                 else Engine(self.params, capacity=self.n + (CHEM_ROOM if self.bonded else 0)))
            if self.ensemble:
                e.set_tuning(nlist_width=256)  # dense corners of the sampled range
            if skin > 0.0:
                e.set_tuning(verlet_skin=skin, verlet_width=512 if self.ensemble else 256)
        st = configs.concat_systems(self.systems) if self.ensemble else self.systems
        e.upload(st["pos"], st["vel"], st["ids"], st["prop"], st["amp_id"], st["unit"],
                 acc=st.get("acc"))
        if self.bonded:
            e.set_topology(st["topo"])
        e.set_clock(0, 0 if self.ensemble else self.params.pair_seed)
        e.run(max(warmup, 1))
        e.sync()
        return e


def workload(name, chains):
    from implicit_dpd_amd import configs
    model = ("Warren MDPD A_ij in [-40,-20] by type, B=25, rd=0.75, bonds k=100 l0=0.5, bending "
             "10, head conversion p=1e-4/step, dt=0.005")
    if name == "membrane":
        p, s = configs.membrane_stack(n_chain=chains)
        return Workload(name, p, s, f"membrane: {chains} H1T3 chains ({4 * chains} beads) as 5 "
                        f"stacked bilayers, box 100x100x40, {model}", True)
    if name == "membrane_gas":
        p, s = configs.membrane(n_chain=chains)
        return Workload(name, p, s, f"membrane model on random chains: {chains} H1T3 chains "
                        f"({4 * chains} beads) spread over the box 100x100x40, {model}", True)
    if name == "slab":
        p, s = configs.mdpd_slab()
        return Workload(name, p, s, "MDPD slab: 198246 beads in the middle half of a 32x32x64 box, "
                        "Warren A=-40, B=25, rd=0.75, gamma=4.5, kBT=1, dt=0.005", False)
    if name == "groot_warren":
        p, s = configs.groot_warren_fluid()
        return Workload(name, p, s, "Groot-Warren DPD fluid: 24000 beads, L=20, rho=3, a=25, "
                        "gamma=4.5, kBT=1, dt=0.01", False)
    if name == "ensemble":
        p, s = configs.ensemble()
        return Workload(name, p, s, "replica ensemble: 64 slab replicas of U[4096,65536] beads "
                        f"({sum(len(x['pos']) for x in s)} beads), A~U[-60,-20], B~U[10,40], "
                        "rd~U[0.65,0.85], kBT=1, dt=0.005, one context", False)
    raise SystemExit(f"unknown workload {name}")


# steps run from the constructed state before anything is timed or measured:
# the stacked bilayers are built as flat sheets at 0.5 bond length and relax
# (the pairs per bead fall from 24 to 16 within 40 steps and settle near 20
# after ~1000), so the headline is taken on the relaxed membrane
RELAX_STEPS = {"membrane": 2000}


def relax(w, Engine, steps, skin=0.0, state=None):
    """Runs the workload's constructed state for `steps` production steps in
    a new engine and returns that engine; the result (positions,
    velocities, accelerations, types) becomes the workload's initial state
    for the CPU baseline.  With `state`, a checkpoint file of that relaxed
    state: loaded when it exists (no relaxation steps run), else written."""
    if state and os.path.exists(state):
        e = w.start(Engine, warmup=1, skin=skin)
        e.checkpoint_load(state)
    else:
        e = w.start(Engine, warmup=steps, skin=skin)
        if state:
            e.checkpoint_save(state)
    g = e.download()
    s = dict(w.systems)
    for k in ("pos", "vel", "acc"):
        s[k] = g[k]
    s["prop"] = g["prop"].astype(s["prop"].dtype)
    w.systems = s
    return e


# relaxation steps run after the relaxed state is saved and before the
# warm-up: the host's copy of the state leaves the GPU idle for a while, and
# the steps that follow an idle spell run slower (~5% over the first ~40,
# measured) until the device is busy again
RESUME_STEPS = 200


def config_gen_bench(args, Engine):
    """Initial configuration (config_maker's flat mode): --chains H1T3
    amphiphiles as a bilayer in the 100x100x40 membrane box with Maxwell-
    Boltzmann velocities, generated on the device straight into an engine
    context (mdpd_make_config), K times after W warm-up calls; the oracle's
    sequential restatement of config_maker times the same spec on one host
    core.  Per bead the device writes pos, vel, vel_buf (3 x 24 B), the four
    ids (16 B), delta_sumr and acc (24 + 12 B) and chain/slot maps (8 B), and
    reads the velocities twice for the drift and temperature sums (48 B)."""
    from implicit_dpd_amd import config_maker as cm
    from implicit_dpd_amd import configs
    box = (100.0, 100.0, 40.0)
    spec = cm.spec("flat", box, args.chains, seed=21)
    n = cm.config_count(spec)
    p, _ = configs.membrane(n_chain=1, box=box)
    e = Engine(p, capacity=n)
    for _ in range(max(args.warmup, 1)):
        e.make_config(spec)
    steps = max(args.steps, 1)
    t = time.perf_counter()
    for _ in range(steps):
        kt = e.make_config(spec)
    wall = (time.perf_counter() - t) / steps
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        t = time.perf_counter()
        rc, _ = po.make_config(spec)
        dt = time.perf_counter() - t
        assert rc == 0
        cpu = {"value": round(n / dt, 1), "unit": "beads/s", "cores": 1, "kind": "port",
               "sample": f"oracle restatement of config_maker (sequential), the same {n}-bead spec, "
                         f"{dt:.2f} s"}
    bytes_per_bead = 3 * 24 + 16 + 36 + 8 + 48
    print(json.dumps({
        "metric": "initial-configuration beads/s (config_maker flat bilayer, generated on the device "
                  "into the engine context)",
        "value": round(n / wall, 1), "unit": "beads/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_call": round(wall * 1e3, 4), "higher_is_better": True,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"flat bilayer, {args.chains} H1T3 amphiphiles ({n} beads), box "
                               "100x100x40, T=1, seed 21", "kin_temp": kt},
        "algorithmic_bytes_per_bead": bytes_per_bead,
        "achieved_gbs_wall": round(n * bytes_per_bead / wall / 1e9, 1),
        "cpu_baseline": cpu}), flush=True)
    e.close()


def traj_io_bench(args, Engine):
    """Trajectory output (io_util::WriteXYZForm, the reference's traject.xyz
    frame every step_mic steps): the membrane after its warm-up steps, one
    frame written by mdpd_write_xyz_async (device gather, copy to pinned
    memory, parallel %.15g formatting on host threads, file write) timed
    to completion; then K steps alone and K steps with a frame in flight.
    The single-threaded baseline formats the same records with the Python
    writer of the same format on a 65,536-bead sample."""
    from implicit_dpd_amd.fpdpd import write_xyz
    w = workload("membrane", args.chains)
    e = w.start(Engine, warmup=args.warmup)
    n = w.n
    d = tempfile.mkdtemp(prefix="traj_")  # host scratch for the frames, removed below
    path = os.path.join(d, "traject.xyz")
    e.sync()
    walls = []
    for k in range(3):
        t = time.perf_counter()
        e.write_xyz_async(path, k, append=False)
        e.io_wait()
        walls.append(time.perf_counter() - t)
    size = os.path.getsize(path)
    frame = min(walls)
    steps = max(args.steps, 1)
    e.sync()
    t = time.perf_counter()
    e.run(steps)
    e.sync()
    alone = time.perf_counter() - t
    t = time.perf_counter()
    e.write_xyz_async(path, 0, append=False)
    e.run(steps)
    e.sync()
    with_io = time.perf_counter() - t
    e.io_wait()
    cpu = None
    if not args.no_cpu:
        rec = e.download_fpdpd()[:65536]
        t = time.perf_counter()
        write_xyz(os.path.join(d, "sample.xyz"), rec, 0)
        dt = time.perf_counter() - t
        cpu = {"value": round(len(rec) / dt, 1), "unit": "beads/s", "cores": 1, "kind": "port",
               "sample": f"the same %.15g xyz lines by implicit_dpd_amd.fpdpd.write_xyz, one thread, "
                         f"{len(rec)} beads, {dt:.2f} s"}
    for f in os.listdir(d):
        os.remove(os.path.join(d, f))
    os.rmdir(d)
    print(json.dumps({
        "metric": "trajectory frame output, beads/s (xyz text of io_util::WriteXYZForm, asynchronous)",
        "value": round(n / frame, 1), "unit": "beads/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_frame": round(frame * 1e3, 2), "frame_bytes": size,
        "frame_mb_per_s": round(size / frame / 1e6, 1), "higher_is_better": True, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"membrane after {args.warmup} steps, {n} beads"},
        # the steps do not wait for the formatting and the file: a frame
        # costs them only its device gather and copy to pinned memory
        "overlap": {"ms_per_step_alone": round(alone * 1e3 / steps, 4),
                    "ms_per_step_with_frame_in_flight": round(with_io * 1e3 / steps, 4),
                    "frame_cost_to_stepping_ms": round((with_io - alone) * 1e3, 2)},
        "cpu_baseline": cpu}), flush=True)
    e.close()


def driver_bench(args, Engine):
    """The drop-in program (implicit_dpd_amd/bin/mdpd_main, main.cpp's control
    flow over the C++ facade) on the headline state: the lamellar membrane
    written as init_config.xyz with a run_param.txt of the reference's own
    multibody model (CalcInterCoef of chi 30, kappa 50, rho_co 25), all_time =
    K + 1 steps, step_mac 100 (observations), no trajectory frames; three
    bilayers instead of five, so that the amphiphiles pack as densely per
    membrane area (~8.7 per unit area) as the reference's own inputs (9,600
    in 31x30, test/mpi_bonded/result_mpi/31x31/run_param.txt).  Its
    main loop replays mdpd_run between observation points and prints its
    own wall time; `value` is beads * steps / that time.  --chem adds the
    RandomChemEvent branch from the first main-loop step on (--chem-mt: with
    the reference's MT19937 draws)."""
    import subprocess
    from implicit_dpd_amd import configs
    from implicit_dpd_amd.fpdpd import FPDPD_DTYPE, write_xyz
    p, st = configs.membrane_stack(n_chain=args.chains, n_layers=3)
    n = len(st["pos"])
    d = tempfile.mkdtemp(prefix="mdpd_driver_")  # a 250 MB input: host scratch, not results
    box = tuple(p.box[:])
    steps = max(args.steps, 1)
    configs.write_run_param(d, box, args.chains, sol_num=0, dt=0.005, all_time=steps + 1,
                            step_mic=10 ** 9, step_mac=100, p_thresld=2e-3, eps=0.2,
                            max_amp_num=args.chains + 8192, influ_rad=3.0, influ_hei=1.0,
                            influ_dep=3.0,
                            influ_grd=0.0, core=(0, 1, 2, 3))
    rec = np.zeros(n, dtype=FPDPD_DTYPE)
    rec["pos"], rec["vel"], rec["vel_buf"] = st["pos"], st["vel"], st["vel"]
    rec["id"], rec["prop"], rec["amp_id"], rec["unit"] = st["ids"], st["prop"], st["amp_id"], st["unit"]
    log(f"[bench] writing init_config.xyz ({n} beads) to {d}")
    write_xyz(os.path.join(d, "init_config.xyz"), rec, 0)
    exe = os.path.join(ROOT, "implicit_dpd_amd", "bin", "mdpd_main")
    cmd = [exe, d, "--seed", "77", "--nlist-width", "192"] + (
        ["--chem", "--beg-chem", "0"] if args.chem else []) + (
        ["--chem-mt"] if args.chem and args.chem_mt else [])
    log("[bench] " + " ".join(cmd))
    try:
        t = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t
    finally:
        for f in os.listdir(d):
            os.remove(os.path.join(d, f))
        os.rmdir(d)
    if r.returncode != 0:
        raise SystemExit(f"driver failed: {r.stderr[-2000:]}")
    line = [x for x in r.stdout.splitlines() if x.startswith("main loop:")][-1]
    loop_steps, loop_s = int(line.split()[2]), float(line.split()[5])
    ms = loop_s * 1e3 / loop_steps
    print(json.dumps({
        "metric": "drop-in driver bead-steps/s (mdpd_main main loop, main.cpp control flow)",
        "value": round(n / (ms * 1e-3), 1), "unit": "bead-steps/s", "n_gpus": 1,
        "steps": loop_steps, "warmup": 1, "ms_per_step": round(ms, 5), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"lamellar membrane (3 bilayers), {args.chains} H1T3 chains ({n} beads), box "
                               "100x100x40, the reference's multibody model from run_param.txt "
                               "(chi 30, kappa 50, rho_co 25, cf_r 3, dt 0.005), step_mac 100, "
                               "pair-list width 192"
                               + (", RandomChemEvent every step (4 cores, p_thresld 2e-3%s)"
                                  % (", the reference's MT19937 stream" if args.chem_mt else "")
                                  if args.chem else ""),
                   "process_wall_s": round(wall, 2)}}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="membrane",
                    choices=["membrane", "membrane_gas", "slab", "groot_warren", "ensemble"],
                    help="membrane (stacked bilayers) is the headline system; membrane_gas the "
                         "same model on random chains; the others are the smaller "
                         "configurations of the parity tests")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--state", default=None,
                    help="checkpoint of the relaxed state: loaded if the file exists, else "
                         "written after the relaxation (profiling runs share one state)")
    ap.add_argument("--relax", type=int, default=-1,
                    help="steps run from the constructed state before the warm-up "
                         "(default: 2000 for the membrane, 0 otherwise)")
    ap.add_argument("--no-chem", action="store_true", help="skip the chemistry-on steps")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--chains", type=int, default=262144)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour-list skin (0: cell list rebuilt every step)")
    ap.add_argument("--config-gen", action="store_true",
                    help="time the initial-configuration generator instead (flat bilayer of "
                         "--chains H1T3 amphiphiles in the membrane box), next to the oracle")
    ap.add_argument("--traj-io", action="store_true",
                    help="time trajectory output instead: one xyz frame of the membrane "
                         "written asynchronously, and steps with a frame in flight")
    ap.add_argument("--driver", action="store_true",
                    help="time the drop-in program mdpd_main on the headline state instead")
    ap.add_argument("--chem", action="store_true",
                    help="with --driver: the RandomChemEvent branch on from the first step")
    ap.add_argument("--chem-mt", action="store_true",
                    help="with --driver --chem: the chemistry draws from the reference's MT19937 stream")
    ap.add_argument("--calib-bytes", type=float, default=0.0,
                    help="only run the copy-kernel probe of this many bytes (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(spatial sharding with halo exchange is not built)", file=sys.stderr)
        sys.exit(2)

    from implicit_dpd_amd.engine import Engine, bandwidth_probe, device_name

    if args.calib_bytes:
        g = bandwidth_probe(int(args.calib_bytes), 10)
        print(json.dumps({"copy_probe_bytes": int(args.calib_bytes), "reps": 11,
                          "gbps": g}), flush=True)
        return

    if args.config_gen:
        config_gen_bench(args, Engine)
        return
    if args.traj_io:
        traj_io_bench(args, Engine)
        return
    if args.driver:
        driver_bench(args, Engine)
        return

    w = workload(args.workload, args.chains)
    n = w.n
    rho_mean = n / w.volume
    log(f"[bench] device {device_name()}; workload {w.name}: {n} beads")
    relax_steps = RELAX_STEPS.get(w.name, 0) if args.relax < 0 else args.relax
    ckpt = None
    if relax_steps:
        # the relaxed run's exact state goes to a checkpoint, from which the
        # replays below start again; the timed steps continue the run in the
        # same engine after RESUME_STEPS more steps
        t = time.perf_counter()
        e = relax(w, Engine, relax_steps, skin=args.skin, state=args.state)
        w.text += (f"; the state after {relax_steps + RESUME_STEPS} steps from the lamellar "
                   "construction")
        ckpt = os.path.join(tempfile.gettempdir(), f"mdpd_bench_{os.getpid()}.ckpt")
        atexit.register(lambda: os.path.exists(ckpt) and os.remove(ckpt))
        e.checkpoint_save(ckpt)
        e.run(RESUME_STEPS)
        log(f"[bench] relaxed the constructed state for {relax_steps} + {RESUME_STEPS} steps: "
            f"{time.perf_counter() - t:.2f} s")

    def restart(e=None, load=True):
        """The state the timed steps start from: the workload's initial state
        (or the relaxed run's checkpoint and the RESUME_STEPS after it) and
        the warm-up steps after that."""
        if ckpt is None:
            return w.start(Engine, e=e, warmup=args.warmup, skin=args.skin)
        if load:
            e.checkpoint_load(ckpt)
            e.run(RESUME_STEPS)
        e.run(max(args.warmup, 1))
        e.sync()
        return e

    t = time.perf_counter()
    e = restart(e, load=False) if ckpt else restart()
    log(f"[bench] warmup {args.warmup} steps: {time.perf_counter() - t:.2f} s")
    if args.autotune:
        tt = e.autotune(10)
        log(f"[bench] autotune: block_scan {tt.block_scan}, block_force {tt.block_force}")

    def observe():
        if not w.ensemble:
            return e.observe()
        obs = e.observe_replicas()
        return {"kin_temp": [o["kin_temp"][3] for o in obs],
                "n_pairs": sum(o["n_pairs"] for o in obs)}

    # timed region: K steps in production mode (captured graphs), the stream
    # idle on entry and synchronised on exit; HIP events on the engine stream
    # give the device time of the same region
    # the step graphs of the timed run are recorded before the clock starts
    e.prepare_run(args.steps)
    e.sync()
    b0 = e.skin_builds()
    t = time.perf_counter()
    gpu_ms = e.run_timed(args.steps)
    wall = time.perf_counter() - t
    builds, n_skin = e.skin_stats()
    builds -= b0
    ms_step = wall * 1e3 / args.steps
    value = n / (ms_step * 1e-3)
    log(f"[bench] timed {args.steps} steps: {ms_step:.4f} ms/step wall, "
        f"{gpu_ms / args.steps:.4f} ms/step device")

    # the same K steps once more from the same start (the trajectory is
    # deterministic), launched kernel by kernel with HIP events around every
    # phase: the per-kernel times of the roofline
    obs_timed = observe()
    restart(e)
    e.set_profiling(True)
    e.run(args.steps)
    e.sync()
    ph = e.phase_times()
    e.set_profiling(False)
    log("[bench] phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in ph.items()))
    obs = observe()
    if list(obs["kin_temp"]) != list(obs_timed["kin_temp"]):
        log("[bench] note: the replayed steps did not reproduce the timed ones")
    n_pairs = obs["n_pairs"]
    nd = w.nd

    prof = profiled_counters(f"{w.name}_1M" if w.name.startswith("membrane") else w.name)
    st0 = w.systems if not w.ensemble else None
    n_chain_beads = int((st0["unit"] != 0xFFFFFFFF).sum()) if w.bonded else 0
    n_heads = int((st0["unit"] == 0).sum()) if w.bonded else 0
    kern = {}
    dens_algo = (skin_density_bytes_per_launch(n, n_pairs, nd, n_skin) if args.skin > 0
                 else density_bytes_per_launch(n, n_pairs, nd, w.ncell))
    force_algo = force_bytes_per_launch(n, n_pairs, nd, n_chain_beads, n_heads, w.ncell,
                                        keys=args.skin == 0, conv=w.params.conv_p > 0
                                        if not w.ensemble else False)
    for name, phase, algo in (
            ("k_density_skin" if args.skin > 0 else "k_density_q", "density", dens_algo),
            ("k_force", "force", force_algo)):
        ms = ph[phase]
        gbs = algo / (ms * 1e-3) / 1e9
        variant, pc = pick(prof, name)
        k = {"ms": round(ms, 5), "algorithmic_bytes": int(algo), "achieved_gbs": round(gbs, 2),
             "frac": round(gbs / HBM_PEAK_GBS, 4), "profiled_variant": variant}
        if pc:
            rms = pc["avg_ns"] * 1e-6
            lane_ops = pc["valu_insts"] * 64 * pc["valu_util"] / 100.0
            k.update({"ms_rocprof": round(rms, 5),
                      "frac_rocprof": round(algo / (rms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": int(pc["traffic_bytes"]),
                      "traffic_over_algorithmic": round(pc["traffic_bytes"] / algo, 2),
                      "valu_busy_pct": pc["valu_busy"], "valu_lane_util_pct": pc["valu_util"],
                      "valu_lane_ops": lane_ops if math.isfinite(lane_ops) else None,
                      "valu_frac": (round(lane_ops / (VALU_LANE_OPS_PEAK * rms * 1e-3), 4)
                                    if math.isfinite(lane_ops) else None)})
        kern[name] = k
    dom = max(kern, key=lambda k: kern[k]["ms"])
    kd = kern[dom]
    bstep = step_bytes_per_bead(nd, w.bonded, rho_mean)
    roof = {
        "bound": "hbm",
        "achieved": kd["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kd["frac"],
        "traffic": kd.get("traffic"),
        "kernel": dom,
        "kernel_ms": kd["ms"],
        "kernel_ms_source": "HIP events around the kernel in an eager replay of the timed steps "
                            "(the same kernels the graphs launch); ms_rocprof: rocprofv3 "
                            "kernel-trace average of this command (profiles/)",
        "algorithmic_bytes_per_launch": kd["algorithmic_bytes"],
        "pairs_per_bead": round(n_pairs / n, 3),
        "binding_roof": "VALU (FP32/INT32 lane ops: candidate tests, Saru hash, log/sqrt/cos): "
                        "valu_frac = SQ_INSTS_VALU x 64 x VALUUtilization / (256 CUs x 64 lanes x "
                        "2.4 GHz x kernel time), from the committed rocprofv3 counters",
        "valu_frac": kd.get("valu_frac"),
        "valu_busy_pct": kd.get("valu_busy_pct"),
        "valu_lane_util_pct": kd.get("valu_lane_util_pct"),
        "pair_kernels": kern,
    }
    step_hbm = {"bytes_per_bead_step": round(bstep, 1),
                "achieved_gbs": round(value * bstep / 1e9, 1),
                "frac": round(value * bstep / (HBM_PEAK_GBS * 1e9), 4)}

    def chem_leg(mt):
        """RandomChemEvent inside the captured steps from the same start: four
        core amphiphiles, cones of radius 3, depth -1..3 around them; keyed
        draws, or (mt) the reference's MT19937 stream seeded with 0."""
        restart(e)
        core = np.array([4 * a for a in (0, 1, 2, 3)], np.uint32)
        n_amp0 = e.observe()["n_amp"]
        e.set_chemistry(p_thresld=2e-3, eps=0.2, influ_rad=3.0, influ_hei=1.0, influ_dep=3.0,
                        influ_grd=0.0, max_amp_num=10 ** 9, core_ptcl_id=core, beg_chem=0)
        if mt:
            e.set_chem_mt(0)
        e.run(3)
        e.prepare_run(args.steps)  # the chemistry steps' graphs, before the clock starts
        e.sync()
        t = time.perf_counter()
        e.run(args.steps)
        cwall = time.perf_counter() - t
        stopped, events, inserted = e.chem_status()
        draws = e.chem_mt_draws() if mt else None
        # a few more chemistry steps launched kernel by kernel: the event's
        # own time (the "reactions" phase)
        e.set_profiling(True)
        e.run(20)
        e.sync()
        cph = e.phase_times()
        e.set_profiling(False)
        if mt:
            e.set_chem_mt(enable=False)
        e.set_chemistry(0, 0, 0, 0, 0, 0, 0, [], enable=False)
        res = {"ms_per_step": round(cwall * 1e3 / args.steps, 5),
               "ms_per_step_without": round(ms_step, 5),
               "event_ms_eager": round(cph.get("reactions", float("nan")), 5),
               "events": events, "amphiphiles_inserted": inserted,
               "beads_end": e.num_beads(), "n_amp_start": n_amp0,
               "setup": "RandomChemEvent (cone target test, %s) at the end of every "
                        "step inside the captured steps: cores = amphiphiles 0-3, p_thresld 2e-3, "
                        "influ_rad 3, influ_hei 1, influ_dep 3, influ_grd 0, eps 0.2; mdpd_run "
                        "synchronises once at its end to read back the bead count"
                        % ("the reference's MT19937 stream seeded with 0, in its draw order"
                           if mt else "keyed draws")}
        if mt:
            res["mt_draws"] = draws
        log(f"[bench] chemistry on{' (reference MT)' if mt else ''}: {res['ms_per_step']:.4f} "
            f"ms/step, {inserted} amphiphiles inserted in {events} events")
        return res

    chem = chem_mt = None
    if w.bonded and not w.ensemble and args.skin == 0 and not args.no_chem:
        chem = chem_leg(False)
        chem_mt = chem_leg(True)

    cpu = None
    if not args.no_cpu and not w.ensemble:
        log("[bench] CPU baseline (oracle)...")
        s = w.systems
        topo = s["topo"] if w.bonded else []
        cpu = cpu_baseline(w.params, s, topo, args.cpu_seconds, 400)
        log(f"[bench] CPU: {cpu['value']:.4g} bead-steps/s on {cpu['cores']} threads")
        one = cpu_baseline(w.params, s, topo, args.cpu_seconds / 2, 400, threads=1)
        cpu["one_thread"] = {k: one[k] for k in ("value", "cores", "sample")}
        log(f"[bench] CPU: {one['value']:.4g} bead-steps/s on 1 thread")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": w.text,
            "beads": n,
            "relax_steps": relax_steps + (RESUME_STEPS if relax_steps else 0),
            "pairs_per_bead": round(n_pairs / n, 3),
            "precision": "pair arithmetic FP32; integration, wrap, reductions FP64",
            "device_ms_per_step": round(gpu_ms / args.steps, 5),
            "phase_ms_eager": {k: round(v, 5) for k, v in ph.items()},
            "neighbours": ({"mode": "skin list", "skin": args.skin,
                            "builds_in_timed_steps": builds,
                            "skin_entries_per_bead": round(n_skin / n, 2)}
                           if args.skin > 0 else {"mode": "cell list rebuilt every step"}),
        },
        "roofline": roof,
        "step_hbm_roofline": step_hbm,
        "chemistry_on": chem,
        "chemistry_reference_mt": chem_mt,
        "cpu_baseline": cpu,
        "calibration": calibration_record(),
    }
    print(json.dumps(out), flush=True)
    e.close()


if __name__ == "__main__":
    main()
