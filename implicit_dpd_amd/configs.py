"""Model parameters and seeded initial configurations of the benchmark systems.

Parameter builders fill the engine's mdpd_params for the three force laws
the engine carries: the reference's multibody model (CalcInterCoef,
parameter.hpp:138-186, read from run_param.txt by the C ABI), Warren MDPD
and Groot-Warren DPD.  Configuration builders follow the roles of the
reference's config_maker (src/config_maker.cpp:60-72 velocities with the
centre-of-mass drift removed, :264-279 random chains), seeded with numpy.
"""
import math
import os
import tempfile

import numpy as np

from .capi import HYPHIL, HYPHOB, SOLVENT, MdpdParams


def _set33(dst, v):
    a = np.broadcast_to(np.asarray(v, dtype=np.float64), (3, 3))
    for i in range(3):
        for j in range(3):
            dst[i][j] = float(a[i, j])


def _base(box, dt, gamma, kbt, envelope, n_species, pair_seed):
    p = MdpdParams()
    for a in range(3):
        p.box[a] = float(box[a])
    p.dt = float(dt)
    p.n_species = int(n_species)
    p.envelope = int(envelope)
    sigma = math.sqrt(2.0 * gamma * kbt) / math.sqrt(dt)
    _set33(p.sigma, sigma)
    _set33(p.gamma, gamma)
    p.bond_leng = 0.5
    p.pair_seed = int(pair_seed) & 0xFFFFFFFF
    p.chem_seed = (int(pair_seed) * 2654435761 + 12345) & 0xFFFFFFFF
    p.cm_corr = 100
    p.arc = 0.0
    return p


def warren_params(box, A=-40.0, B=25.0, rd=0.75, gamma=4.5, kbt=1.0, dt=0.005, envelope=1,
                  n_species=1, pair_seed=1):
    """Warren MDPD: F_C = A_ij (1-r) + B (rho_i + rho_j)(1 - r/rd), rho with
    weight 15/(2 pi rd^3) (1 - r/rd)^2."""
    p = _base(box, dt, gamma, kbt, envelope, n_species, pair_seed)
    p.n_dens = 1
    p.rd = float(rd)
    p.rdm = float(rd)
    p.crho = 15.0 / (2.0 * math.pi * rd ** 3)
    _set33(p.a0, A)
    _set33(p.c1, 0.0)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                p.mb[i][j][k] = float(B) if k == 0 else 0.0
    return p


def groot_warren_params(box, a=25.0, gamma=4.5, kbt=1.0, dt=0.01, envelope=1, pair_seed=1):
    """Groot-Warren DPD (the reference's PAIRWISE_DPD build with cf_pair = a,
    nbdf_calculator.hpp:114-117)."""
    p = _base(box, dt, gamma, kbt, envelope, 1, pair_seed)
    p.n_dens = 1
    p.rd = 1.0
    p.rdm = 1.0
    p.crho = 1.0
    _set33(p.a0, a)
    _set33(p.c1, 0.0)
    return p


RUN_PARAM_TEMPLATE = """box_leng {bx} {by} {bz}
init_amp_num {init_amp_num}
sol_num {sol_num}
dt {dt}
cf_r {cf_r}
cf_s {cf_s}
cf_b {cf_b}
cf_b_rigid {cf_b_rigid}
chi {chi}
kappa {kappa}
rho_co {rho_co}
all_time {all_time}
step_mic {step_mic}
step_mac {step_mac}
p_thresld {p_thresld}
eps {eps}
max_amp_num {max_amp_num}
influ_rad {influ_rad}
influ_hei {influ_hei}
influ_dep {influ_dep}
influ_grd {influ_grd}
core_amp_id {core}
"""


def write_run_param(dirname, box, init_amp_num, sol_num=0, dt=0.005, cf_r=(3.0,) * 6,
                    cf_s=100.0, cf_b=10.0, cf_b_rigid=40.0, chi=30.0, kappa=50.0, rho_co=25.0,
                    all_time=1000, step_mic=1000, step_mac=100, p_thresld=0.1, eps=0.2,
                    max_amp_num=None, influ_rad=5.0, influ_hei=0.8, influ_dep=4.0,
                    influ_grd=0.0, core=(0,)):
    os.makedirs(dirname, exist_ok=True)
    txt = RUN_PARAM_TEMPLATE.format(
        bx=box[0], by=box[1], bz=box[2], init_amp_num=init_amp_num, sol_num=sol_num, dt=dt,
        cf_r=" ".join(str(x) for x in cf_r), cf_s=cf_s, cf_b=cf_b, cf_b_rigid=cf_b_rigid,
        chi=chi, kappa=kappa, rho_co=rho_co, all_time=all_time, step_mic=step_mic,
        step_mac=step_mac, p_thresld=p_thresld, eps=eps,
        max_amp_num=max_amp_num if max_amp_num is not None else init_amp_num * 2,
        influ_rad=influ_rad, influ_hei=influ_hei, influ_dep=influ_dep, influ_grd=influ_grd,
        core=" ".join(str(c) for c in core))
    path = os.path.join(dirname, "run_param.txt")
    with open(path, "w") as f:
        f.write(txt)
    return path


def reference_params(box, pair_seed=1, **kw):
    """The reference's multibody model: run_param.txt -> CalcInterCoef through
    the engine's own parser (mdpd_read_run_param)."""
    from .engine import read_run_param
    with tempfile.TemporaryDirectory() as d:
        write_run_param(d, box, kw.pop("init_amp_num", 1), **kw)
        p, ri = read_run_param(d)
    p.pair_seed = int(pair_seed) & 0xFFFFFFFF
    p.chem_seed = (int(pair_seed) * 2654435761 + 12345) & 0xFFFFFFFF
    return p, ri


# ---------------------------------------------------------------- systems
def mb_velocities(rng, n, kbt=1.0):
    """Maxwell-Boltzmann velocities with the centre-of-mass drift removed
    (config_maker.cpp:60-72)."""
    v = rng.normal(0.0, math.sqrt(kbt), size=(n, 3))
    v -= v.mean(axis=0)
    return v


def random_fluid(n, box, seed, n_species=1, zlim=None):
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    pos = rng.random((n, 3)) * box
    if zlim is not None:
        pos[:, 2] = zlim[0] + rng.random(n) * (zlim[1] - zlim[0])
    vel = mb_velocities(rng, n)
    prop = rng.integers(0, n_species, n).astype(np.uint32) if n_species > 1 else np.zeros(n, np.uint32)
    return dict(pos=pos, vel=vel, ids=np.arange(n, dtype=np.uint32), prop=prop,
                amp_id=np.full(n, 0xFFFFFFFF, np.uint32), unit=np.full(n, 0xFFFFFFFF, np.uint32),
                topo=np.zeros(0, np.uint32))


def random_chains(n_chain, box, seed, bond=0.5, n_solvent=0):
    """H1T3 amphiphiles (head type HYPHIL, three HYPHOB tail beads) at random
    positions with random orientation, bead u at head + u*bond*dir, wrapped;
    amphiphile a occupies bead slots 4a..4a+3 and solvent beads follow."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    head = rng.random((n_chain, 3)) * box
    d = rng.normal(size=(n_chain, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u = np.arange(4, dtype=np.float64)
    pos = (head[:, None, :] + bond * u[None, :, None] * d[:, None, :]).reshape(-1, 3)
    n = 4 * n_chain + n_solvent
    if n_solvent:
        pos = np.concatenate([pos, rng.random((n_solvent, 3)) * box])
    pos -= np.floor(pos / box) * box
    pos = np.minimum(pos, np.nextafter(box, 0))
    vel = mb_velocities(rng, n)
    prop = np.concatenate([np.tile(np.array([HYPHIL, HYPHOB, HYPHOB, HYPHOB], np.uint32), n_chain),
                           np.full(n_solvent, SOLVENT, np.uint32)])
    amp_id = np.concatenate([np.repeat(np.arange(n_chain, dtype=np.uint32), 4),
                             np.full(n_solvent, 0xFFFFFFFF, np.uint32)])
    unit = np.concatenate([np.tile(np.arange(4, dtype=np.uint32), n_chain),
                           np.full(n_solvent, 0xFFFFFFFF, np.uint32)])
    topo = np.arange(4 * n_chain, dtype=np.uint32)
    return dict(pos=pos, vel=vel, ids=np.arange(n, dtype=np.uint32), prop=prop, amp_id=amp_id,
                unit=unit, topo=topo)


# ---------------------------------------------------------------- workloads
def groot_warren_fluid(seed=1):
    """Plain DPD fluid, L = 20, rho = 3 (24,000 beads), a = 25, gamma = 4.5,
    kBT = 1, w_R = 1 - r, dt = 0.01."""
    box = (20.0, 20.0, 20.0)
    return groot_warren_params(box, pair_seed=100), random_fluid(24000, box, seed)


def mdpd_slab(seed=2, n=198246, box=(32.0, 32.0, 64.0), A=-40.0, B=25.0, rd=0.75):
    """Single-component MDPD liquid slab: beads uniform in the middle half of
    z, A = -40, B = 25, rd = 0.75, gamma = 4.5, kBT = 1, dt = 0.005."""
    p = warren_params(box, A=A, B=B, rd=rd, dt=0.005, pair_seed=seed + 100)
    return p, random_fluid(n, box, seed, zlim=(box[2] / 4, 3 * box[2] / 4))


# A_ij by bead type for the membrane: head H(0), tail T(1), converted head
# C(2, a tail-like head), all within [-40, -20]
MEMBRANE_A = np.array([[-20.0, -30.0, -30.0],
                       [-30.0, -40.0, -40.0],
                       [-30.0, -40.0, -40.0]])


def membrane(seed=3, n_chain=262144, box=(100.0, 100.0, 40.0), conv_p=1e-4):
    """Implicit-solvent amphiphile system: H1T3 chains, Warren MDPD with
    A_ij in [-40, -20] by type, B = 25, rd = 0.75, harmonic bonds
    k = 100, l0 = 0.5 (cf_s = k l0 = 50 in F = cf_s (1/r - 1/l0) dr),
    bending k_theta = 10 (cf_b = 10 in U = cf_b (1 - cos)), head conversion
    H -> C with probability conv_p per step, dt = 0.005."""
    p = warren_params(box, A=MEMBRANE_A, B=25.0, rd=0.75, dt=0.005, n_species=3,
                      pair_seed=seed + 100)
    p.cf_s = 100.0 * 0.5
    p.cf_b = 10.0
    p.cf_b_rigid = 10.0
    p.conv_p = conv_p
    p.conv_from = HYPHIL
    p.conv_to = SOLVENT
    return p, random_chains(n_chain, box, seed)


def lamellar_chains(n_chain, box, seed, n_layers=5, bond=0.5, jitter=0.05):
    """H1T3 amphiphiles as a stack of n_layers flat bilayers along z, each
    the flat sheet of config_maker (src/config_maker.cpp:103-137): in-plane
    positions uniform, leaflets alternating, the head at 3 bond + 1e-5 from
    the sheet's mid-plane and the chain straight towards it, here with a
    small Gaussian jitter on every bead; layer k is centred at
    (k + 1/2) Lz / n_layers.  Amphiphile a occupies beads 4a..4a+3."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    layer = np.arange(n_chain) % n_layers
    zc = (layer + 0.5) * box[2] / n_layers
    sign = np.where((np.arange(n_chain) // n_layers) % 2 == 0, 1.0, -1.0)
    head = np.column_stack([rng.random(n_chain) * box[0], rng.random(n_chain) * box[1],
                            zc + (3 * bond + 1e-5) * sign])
    u = np.arange(4, dtype=np.float64)
    pos = head[:, None, :].repeat(4, axis=1)
    pos[:, :, 2] -= sign[:, None] * bond * u[None, :]
    pos += rng.normal(scale=jitter, size=pos.shape)
    pos = pos.reshape(-1, 3)
    pos -= np.floor(pos / box) * box
    pos = np.minimum(pos, np.nextafter(box, 0))
    n = 4 * n_chain
    return dict(pos=pos, vel=mb_velocities(rng, n), ids=np.arange(n, dtype=np.uint32),
                prop=np.tile(np.array([HYPHIL, HYPHOB, HYPHOB, HYPHOB], np.uint32), n_chain),
                amp_id=np.repeat(np.arange(n_chain, dtype=np.uint32), 4),
                unit=np.tile(np.arange(4, dtype=np.uint32), n_chain),
                topo=np.arange(n, dtype=np.uint32))


def membrane_stack(seed=5, n_chain=262144, box=(100.0, 100.0, 40.0), n_layers=5, conv_p=1e-4):
    """The membrane model of membrane() on an actual membrane state: the
    262,144 H1T3 chains as n_layers stacked bilayers (lamellar phase) of
    ~52k amphiphiles each, so the beads sit at the MDPD liquid density
    (~6 per rc^3) inside the bilayers instead of a uniform 2.6."""
    p, _ = membrane(seed=seed, n_chain=1, box=box, conv_p=conv_p)
    return p, lamellar_chains(n_chain, box, seed, n_layers=n_layers)


def ensemble(seed=4, n_rep=64, nmin=4096, nmax=65536, rho=6.0):
    """Replicas of uneven size (U[nmin, nmax] beads), A ~ U[-60,-20],
    B ~ U[10,40], rd ~ U[0.65,0.85], kBT = 1, independent seeds; each a
    liquid slab at initial density rho in the middle half of an Lx=Ly,
    Lz = 2 Lx box."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, n_rep)
    params, systems = [], []
    for r in range(n_rep):
        A = rng.uniform(-60.0, -20.0)
        B = rng.uniform(10.0, 40.0)
        rd = rng.uniform(0.65, 0.85)
        n = int(sizes[r])
        lx = max(4.0, math.ceil((n / (rho * 1.0)) ** (1.0 / 3.0)))
        box = (lx, lx, 2.0 * lx)
        # liquid of density rho in z in [Lz/4, 3Lz/4): n / (lx*lx*lx) = rho
        p = warren_params(box, A=A, B=B, rd=rd, dt=0.005, pair_seed=seed * 1000 + r + 1)
        params.append(p)
        systems.append(random_fluid(n, box, seed * 7919 + r, zlim=(box[2] / 4, 3 * box[2] / 4)))
    return params, systems


def concat_systems(systems):
    out = {}
    off = 0
    parts = {k: [] for k in ("pos", "vel", "ids", "prop", "amp_id", "unit", "topo")}
    for s in systems:
        for k in ("pos", "vel", "prop", "amp_id", "unit"):
            parts[k].append(s[k])
        parts["ids"].append(s["ids"])
        parts["topo"].append(s["topo"] + off)
        off += len(s["pos"])
    for k, v in parts.items():
        out[k] = np.concatenate(v) if v else v
    return out
