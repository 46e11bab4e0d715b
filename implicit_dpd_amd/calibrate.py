"""Parameter calibration with the equation-of-state surrogate vs a plain sweep.

Problem: Warren MDPD (F_C = A (1-r) + B (rho_i + rho_j)(1 - r/rd)) with
B = 25, rd = 0.75, kBT = 1; find A such that a liquid slab coexisting with
its vapour has bulk density rho = 6.0 (within 1%) and normal pressure
P_zz = 0 (the coexistence condition).

- A *plain run* is one slab system of ~200k beads (box L x L x 2L, liquid
  started in the middle half at density 6), run on the GPU; it measures the
  slab's bulk density from the z density profile and P_zz, averaged over the
  second half of the run.
- The *sweep* finds A by regula falsi on plain runs alone; the *grid
  sweep* (``--grid``) steps A by 1 from the dilute end of the range until a
  run is within 1% or passes the target, then bisects the last interval.
- The *surrogate* is trained on ensemble runs (64 replicas of U[4k, 64k]
  beads per launch with A ~ U[-60,-20], B ~ U[10,40], rd ~ U[0.65,0.85],
  kBT ~ U[0.8,1.2]) and proposes A by scanning the network at kBT = 1; each proposal is verified
  by a plain run, and a miss is followed by a Newton step on the measured
  residual with the network's slope (the secant once two runs exist).

The headline of `main()` is the number of plain runs each method needs.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import configs
from .engine import Engine
from .surrogate import Scaler, Surrogate

TARGET_RHO = 6.0
TOL = 0.01
A_RANGE = (-60.0, -20.0)
B_RANGE = (10.0, 40.0)
RD_RANGE = (0.65, 0.85)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def slab_rho(profile, bin_vol, vapour_frac=0.25):
    """Bulk density of a slab from a z bead-count profile: bins above half
    the peak form the slab (periodic); the two outermost bins on each side
    are interface and dropped.  Without a dilute vapour (no bin below
    vapour_frac of the peak: the liquid does not coexist with a vapour, or
    has filled the box) there is no slab and the density of the homogeneous
    fluid, the mean over all bins, is returned; so the estimate falls
    steadily to N / V as the attraction weakens instead of tracking noise."""
    rho = np.asarray(profile, np.float64) / bin_vol
    inside = rho > 0.5 * rho.max()
    if inside.all() or rho.min() > vapour_frac * rho.max():
        return float(rho.mean())
    # rotate so that the slab does not wrap
    start = int(np.argmin(inside))
    rho, inside = np.roll(rho, -start), np.roll(inside, -start)
    idx = np.nonzero(inside)[0]
    core = idx[2:-2] if len(idx) > 6 else idx
    return float(rho[core].mean()) if len(core) else float(rho.max())


def measure(e, boxes, steps, every, nbins):
    """Runs `steps` steps; over the second half, every `every` steps, sums
    the z bead-count profile and P_zz of every replica.  Returns per replica
    (rho_bulk, P_zz)."""
    half = steps // 2
    e.run(half)
    nrep = len(boxes)
    prof = np.zeros((nrep, nbins))
    pzz = np.zeros(nrep)
    k = 0
    done = half
    while done < steps:
        m = min(every, steps - done)
        e.run(m)
        done += m
        prof += e.density_profile(2, nbins)
        pzz += np.array([o["pressure"][2] for o in e.observe_replicas()])
        k += 1
    out = []
    for r, b in enumerate(boxes):
        bin_vol = b[0] * b[1] * b[2] / nbins
        out.append((slab_rho(prof[r] / k, bin_vol), pzz[r] / k))
    return out


def slab(A, B, rd, n, seed, rho0=6.0, kbt=1.0):
    lx = max(4.0, math.ceil((n / rho0) ** (1.0 / 3.0)))
    box = (lx, lx, 2.0 * lx)
    n = int(rho0 * lx * lx * lx)
    p = configs.warren_params(box, A=A, B=B, rd=rd, kbt=kbt, dt=0.005, pair_seed=seed)
    s = configs.random_fluid(n, box, seed, zlim=(box[2] / 4, 3 * box[2] / 4))
    return p, s, box


def plain_run(A, B=25.0, rd=0.75, n=200000, steps=3000, every=50, seed=11, kbt=1.0):
    p, s, box = slab(A, B, rd, n, seed, kbt=kbt)
    e = Engine(p, capacity=len(s["pos"]))
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"])
    e.set_clock(0, p.pair_seed)
    (rho, pzz), = measure(e, [box], steps, every, int(box[2] / 0.5))
    e.close()
    return rho, pzz


KBT_RANGE = (0.8, 1.2)


def ensemble_samples(n_launch, n_rep=64, nmin=4096, nmax=65536, steps=3000, every=100, seed=100):
    """(A, B, rd, kBT) -> (rho, P_zz) samples: n_launch ensemble launches of
    n_rep slab replicas of U[nmin, nmax] beads with the parameters uniform
    over their ranges, kBT over KBT_RANGE."""
    X, Y = [], []
    for b in range(n_launch):
        rng = np.random.default_rng(seed + b)
        params, systems, boxes, xs = [], [], [], []
        for r in range(n_rep):
            A = rng.uniform(*A_RANGE)
            B = rng.uniform(*B_RANGE)
            rd = rng.uniform(*RD_RANGE)
            kbt = rng.uniform(*KBT_RANGE)
            n = int(rng.integers(nmin, nmax + 1))
            p, s, box = slab(A, B, rd, n, seed=(seed + b) * 1000 + r + 1, kbt=kbt)
            params.append(p)
            systems.append(s)
            boxes.append(box)
            xs.append((A, B, rd, kbt))
        allb = configs.concat_systems(systems)
        e = Engine(params, n_beads=[len(s["pos"]) for s in systems])
        # strongly attractive, weakly repulsive corners of the sampled range
        # condense to rho > 20: room for up to 256 neighbours per bead
        e.set_tuning(nlist_width=256)
        e.upload(allb["pos"], allb["vel"], allb["ids"], allb["prop"])
        e.set_clock(0, 0)
        t0 = time.perf_counter()
        res = measure(e, boxes, steps, every, 64)
        e.close()
        X += xs
        Y += res
        log(f"[calib] ensemble launch {b + 1}/{n_launch}: {sum(len(s['pos']) for s in systems)} "
            f"beads, {time.perf_counter() - t0:.1f} s")
    return np.array(X), np.array(Y)


def sweep(plain, target=TARGET_RHO, lo=A_RANGE[0], hi=A_RANGE[1], max_runs=12):
    """Regula falsi (Illinois) on A with plain runs; rho decreases as A
    rises.  Returns (A, runs); A is None when [lo, hi] does not bracket the
    target."""
    runs = []
    f_lo = plain(lo)[0] - target
    runs.append((lo, f_lo + target))
    f_hi = plain(hi)[0] - target
    runs.append((hi, f_hi + target))
    for a, rho in runs:
        if abs(rho - target) <= TOL * target:
            return a, runs
    if (f_lo > 0) == (f_hi > 0):
        return None, runs
    a = None
    while len(runs) < max_runs:
        a = hi - f_hi * (hi - lo) / (f_hi - f_lo)
        rho, _ = plain(a)
        runs.append((a, rho))
        f = rho - target
        if abs(f) <= TOL * target:
            return a, runs
        if (f > 0) == (f_lo > 0):
            lo, f_lo = a, f
            f_hi *= 0.5  # Illinois step keeps both ends moving
        else:
            hi, f_hi = a, f
            f_lo *= 0.5
    return None, runs


def grid_sweep(plain, target=TARGET_RHO, lo=A_RANGE[0], hi=A_RANGE[1], step=1.0,
               max_refine=8):
    """The plain parameter sweep: A on a grid of `step` from the dilute end
    (hi) towards lo until a run is within TOL of the target or passes it,
    then bisection between the last two grid points.  Returns (A, runs); A
    is None when the grid never reaches the target."""
    runs = []
    prev = None
    n = int(round((hi - lo) / step))
    for k in range(n + 1):
        a = hi - k * step
        rho, _ = plain(a)
        runs.append((a, rho))
        if abs(rho - target) <= TOL * target:
            return a, runs
        if rho > target:
            break
        prev = (a, rho)
    else:
        return None, runs
    if prev is None:
        return None, runs  # already past the target at the dilute end
    (a_hi, _), (a_lo, _) = prev, runs[-1]
    for _ in range(max_refine):
        a = 0.5 * (a_hi + a_lo)
        rho, _ = plain(a)
        runs.append((a, rho))
        if abs(rho - target) <= TOL * target:
            return a, runs
        if rho > target:
            a_lo = a
        else:
            a_hi = a
    return None, runs


def train_surrogate(X, Y, epochs=300, seed=3):
    xs, ys = Scaler(X), Scaler(Y)
    net = Surrogate(max_batch=4096, seed=seed)
    hist = net.train(xs.fwd(X), ys.fwd(Y), epochs=epochs, batch=256, lr=2e-3, seed=seed)
    return net, xs, ys, hist


def surrogate_calibrate(net, xs, ys, plain, target=TARGET_RHO, B=25.0, rd=0.75, max_runs=12,
                        reach=0.05):
    """Surrogate-guided root finding on A at fixed B, rd, kBT = 1.

    The network is scanned over A for the first proposal.  A target the
    network puts more than `reach` (relative) away from every A of the range
    is checked once at the closest A and reported unreachable.  After a miss
    the next A is a Newton step on the measured residual, with the slope of
    the network at the current A (and the secant of the last two plain runs
    once there are two, when it has the network's sign).  Once plain runs lie
    on both sides of the target (rho falls as A rises), a step that leaves
    that bracket is replaced by its midpoint (safeguarded Newton), so noisy
    runs cannot make it oscillate.  The run budget is the bracketing
    sweep's."""
    grid = np.linspace(A_RANGE[0], A_RANGE[1], 4001)
    cand = np.stack([grid, np.full_like(grid, B), np.full_like(grid, rd), np.ones_like(grid)], 1)
    pred = ys.inv(net.forward(xs.fwd(cand)))[:, 0]
    i = int(np.argmin(np.abs(pred - target)))
    a = float(grid[i])
    runs = []

    def slope_at(x):
        j = int(np.clip(np.searchsorted(grid, x), 40, len(grid) - 41))
        return (pred[j + 40] - pred[j - 40]) / (grid[j + 40] - grid[j - 40])

    if abs(pred[i] - target) > reach * target:
        rho, _ = plain(a)
        runs.append((a, rho, float(pred[i])))
        return None, runs
    below = above = None  # the nearest A with rho above / below the target
    for _ in range(max_runs):
        rho, _ = plain(a)
        runs.append((a, rho, float(np.interp(a, grid, pred))))
        if abs(rho - target) <= TOL * target:
            return a, runs
        if rho > target:
            below = a if below is None else max(below, a)
        else:
            above = a if above is None else min(above, a)
        sl = slope_at(a)
        if len(runs) >= 2 and abs(runs[-1][0] - runs[-2][0]) > 1e-6:
            sec = (runs[-1][1] - runs[-2][1]) / (runs[-1][0] - runs[-2][0])
            if sec * sl > 0:
                sl = sec
        if sl == 0 or not np.isfinite(sl):
            break
        a = float(np.clip(a + (target - rho) / sl, A_RANGE[0], A_RANGE[1]))
        if below is not None and above is not None and below < above and not below < a < above:
            a = 0.5 * (below + above)
    return None, runs


# (target bulk density, B, rd): the headline case first, then further
# targets inside the sampled ranges that show how the sample set amortises
MORE_TARGETS = [(5.5, 25.0, 0.75), (6.5, 25.0, 0.75), (6.0, 20.0, 0.70), (6.0, 35.0, 0.80),
                (5.0, 30.0, 0.70), (7.0, 30.0, 0.80)]


def main(argv=None):
    ap = argparse.ArgumentParser(description="surrogate vs plain-sweep calibration")
    ap.add_argument("--launches", type=int, default=64, help="ensemble launches of 64 replicas")
    ap.add_argument("--nmin", type=int, default=4096)
    ap.add_argument("--nmax", type=int, default=65536)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--plain-beads", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--more-targets", action="store_true",
                    help="also calibrate six further (rho, B, rd) targets")
    ap.add_argument("--grid", action="store_true",
                    help="also run the plain grid sweep (A in steps of 1 from -20, then "
                         "bisection)")
    ap.add_argument("--save-samples", default=None, help="write the ensemble samples (.npz)")
    ap.add_argument("--load-samples", default=None,
                    help="reuse ensemble samples written by --save-samples")
    args = ap.parse_args(argv)

    secs = {"samples": 0.0, "train": 0.0, "surrogate_runs": 0.0, "sweep_runs": 0.0,
            "grid_runs": 0.0}

    def plain(key, B, rd):
        def run(a):
            t0 = time.perf_counter()
            r = plain_run(a, B=B, rd=rd, n=args.plain_beads, steps=args.steps)
            secs[key] += time.perf_counter() - t0
            log(f"[calib] plain run A={a:.4f} B={B} rd={rd}: rho={r[0]:.4f} P_zz={r[1]:.4f}")
            return r
        return run

    if args.load_samples:
        z = np.load(args.load_samples)
        X, Y = z["X"], z["Y"]
        secs["samples"] = float(z["seconds"])  # GPU time of the run that made them
    else:
        t0 = time.perf_counter()
        X, Y = ensemble_samples(args.launches, nmin=args.nmin, nmax=args.nmax, steps=args.steps)
        secs["samples"] = time.perf_counter() - t0
    if args.save_samples:
        np.savez(args.save_samples, X=X, Y=Y, seconds=secs["samples"])
    t0 = time.perf_counter()
    net, xs, ys, hist = train_surrogate(X, Y, epochs=args.epochs)
    secs["train"] = time.perf_counter() - t0
    targets = [(TARGET_RHO, 25.0, 0.75)] + (MORE_TARGETS if args.more_targets else [])
    results = []
    for rho_t, B, rd in targets:
        a_s, runs_s = surrogate_calibrate(net, xs, ys, plain("surrogate_runs", B, rd),
                                          target=rho_t, B=B, rd=rd)
        a_p, runs_p = sweep(plain("sweep_runs", B, rd), target=rho_t)
        a_g, runs_g = (grid_sweep(plain("grid_runs", B, rd), target=rho_t) if args.grid
                       else (None, []))
        results.append({
            "target": {"rho": rho_t, "tol": TOL, "B": B, "rd": rd, "kBT": 1.0},
            "surrogate": {"A": a_s, "plain_runs": len(runs_s),
                          "runs": [{"A": a, "rho": r, "rho_pred": q} for a, r, q in runs_s]},
            "sweep": {"A": a_p, "plain_runs": len(runs_p),
                      "runs": [{"A": a, "rho": r} for a, r in runs_p]},
            "grid": {"A": a_g, "plain_runs": len(runs_g),
                     "runs": [{"A": a, "rho": r} for a, r in runs_g]}})
    net.close()
    n_s = sum(r["surrogate"]["plain_runs"] for r in results)
    n_p = sum(r["sweep"]["plain_runs"] for r in results)
    n_g = sum(r["grid"]["plain_runs"] for r in results)
    out = {
        "samples": int(len(X)),
        "train_loss_first_last": [float(hist[0]), float(hist[-1])],
        "targets": results,
        "plain_runs": {"surrogate": n_s, "sweep": n_p, "ratio": round(n_p / max(n_s, 1), 3),
                       "grid": n_g, "grid_ratio": round(n_g / max(n_s, 1), 3)},
        "hits": {"surrogate": sum(r["surrogate"]["A"] is not None for r in results),
                 "sweep": sum(r["sweep"]["A"] is not None for r in results),
                 "grid": sum(r["grid"]["A"] is not None for r in results)},
        "seconds": {k: round(v, 2) for k, v in secs.items()},
    }
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
