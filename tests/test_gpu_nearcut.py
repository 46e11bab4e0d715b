"""Isolated pairs next to the cutoff: the case where the FP32 value of 1 - r
decides whether a bead's force meets 1e-5 of its pair-force sum.

Most beads have exactly one neighbour (a few see the next pair too), at r
within 0.025 below the cutoff (within 2e-3 for the reference's
sqrt(1 - r) envelope, which magnifies the error of r most), with keyed
noise xi of either sign.  The force pass has to see that the FP32 error of r (about 1e-7,
times sigma' |xi| of up to ~170 at dt = 0.005, and magnified by
1 / (2 sqrt(1 - r)) under the square root) is not small against so small a
force, and evaluate these pairs' 1 - r from the FP64 master positions; the
forces then agree with the FP64 oracle (CalcForceEpEpDPD,
nbdf_calculator.hpp:73-143) to 1e-5 of the pair-force sum even for forces
of 1e-3.
"""
import numpy as np
import pytest

from implicit_dpd_amd import configs

pytestmark = pytest.mark.gpu


def _pairs(m, spacing, r_lo, r_hi, seed, types=None):
    """m^3 pairs on a grid of centres `spacing` apart (no two pairs within
    the cutoff of each other), separation uniform in [r_lo, r_hi)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
    c = (g + 0.5) * spacing + rng.uniform(-0.2, 0.2, g.shape)
    u = rng.normal(size=g.shape)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = rng.uniform(r_lo, r_hi, len(g))[:, None]
    pos = np.concatenate([c + 0.5 * r * u, c - 0.5 * r * u])
    n = len(pos)
    box = np.full(3, m * spacing)
    pos %= box
    vel = configs.mb_velocities(rng, n)
    prop = (np.zeros(n, np.uint32) if types is None
            else rng.integers(0, types, n).astype(np.uint32))
    return box, dict(pos=pos, vel=vel, ids=rng.permutation(n).astype(np.uint32), prop=prop,
                     amp_id=np.full(n, 0xFFFFFFFF, np.uint32),
                     unit=np.full(n, 0xFFFFFFFF, np.uint32))


CASES = {
    # linear envelope (Warren MDPD)
    "warren_linear": (0.976, 0.99999, None),
    # the reference's multibody model, sqrt(1 - r)
    "reference_sqrt": (0.9981, 0.999999, 3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_isolated_near_cutoff_pairs_match_oracle(engine_mod, oracle, name):
    r_lo, r_hi, types = CASES[name]
    box, s = _pairs(13, 2.3, r_lo, r_hi, seed=17, types=types)
    if name == "warren_linear":
        p = configs.warren_params(tuple(box), A=-40.0, B=25.0, rd=0.75, pair_seed=9)
    else:
        p, _ = configs.reference_params(tuple(box), pair_seed=9, init_amp_num=1, dt=0.005)
    time = 31
    e = engine_mod.Engine(p, capacity=len(s["pos"]))
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_clock(time, p.pair_seed)
    e.rebuild()
    e.density()
    e.force()
    g = e.download()
    e.close()
    dens = oracle.density(p, s["pos"], s["prop"])
    acc, _ = oracle.force(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + time)
    fa = oracle.force_abs(p, s["pos"], s["vel"], s["ids"], s["prop"], dens, p.pair_seed + time)
    # every bead has its one partner inside the cutoff
    assert np.all(fa > 0)
    err = np.abs(g["acc"] - acc).max(axis=1)
    # the bar of the dense systems (1e-5 of the pair-force sum, floored at
    # 1e-3) plus the noise floor of the FP32 normal: where xi is near zero
    # the force is tiny and xi's own absolute error (about 1e-7, from the
    # FP32 uniform and cosine) times sigma' wR (below 7 in the band) is
    # what remains.  With 1 - r from FP32 the error would be sigma' |xi|
    # 1e-7 (times 1 / (2 sqrt(1 - r)) under the square root): up to 1e-5
    # against forces of 0.1 under the linear envelope, and far more under
    # the square root.
    bar = 1e-5 * np.maximum(fa, 1e-3) + 7e-7
    worst = int((err / bar).argmax())
    assert np.all(err < bar), (err[worst], fa[worst], worst)
