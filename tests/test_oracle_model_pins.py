"""Pins of the oracle's reference-model density pass (per-species density and
the neighbour centre-of-mass sums) and of its reference-model pair force,
against brute force and against closed forms on 2- and 3-bead systems whose
answers follow from the reference's formulas by hand (no GPU).

- Neighbour CM (nbdf_calculator.hpp:18-58, write-back ptcl_class.hpp:53-57):
  the hydrophilic sum starts with the bead itself (count 1) and adds the
  Hyphil neighbours with r^2 < 1; the hydrophobic sum adds the Hyphob
  neighbours with r^2 < rn_c^2 = 1.44; a neighbour enters at its periodic
  image nearest to the bead.
- Per-species density (nbdf_calculator.hpp:38-40): rho_i[k] sums (1 - r)^2
  over neighbours j of type k = prop_j.
- Reference pair force (nbdf_calculator.hpp:99-135) with the CalcInterCoef
  tables (parameter.hpp:138-186), whose cf_m is zero on the pure-Hyphil and
  pure-solvent diagonals.
"""
import math

import numpy as np
import pytest

from implicit_dpd_amd import configs

HYPHIL, HYPHOB, SOLVENT = configs.HYPHIL, configs.HYPHOB, configs.SOLVENT


def _minimg(d, box):
    return d - box * np.round(d / box)


def cm_brute(pos, prop, box):
    """Neighbour CMs by brute force over all bead pairs (numpy)."""
    box = np.asarray(box, dtype=np.float64)
    d = _minimg(pos[:, None, :] - pos[None, :, :], box)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    img = pos[:, None, :] - d  # image of j nearest to i
    phil = (r2 < 1.0) & (prop[None, :] == HYPHIL)
    phob = (r2 < 1.44) & (prop[None, :] == HYPHOB)
    cl = 1 + phil.sum(1)
    cb = phob.sum(1)
    sl = pos + (img * phil[..., None]).sum(1)
    sb = (img * phob[..., None]).sum(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return sl / cl[:, None], sb / cb[:, None], cl, cb


def _ref_params(box):
    # the current revision's CalcInterCoef (Reo = 1.4), pinned against the
    # reference's dumped tables in test_oracle_pins.py
    from oracle import pyoracle
    p = pyoracle.calc_inter_coef(30.0, 50.0, 25.0, 0.1, [3.0] * 6, 0.005)
    for a in range(3):
        p.box[a] = box[a]
    return p


@pytest.mark.parametrize("box", [(4.5, 5.0, 4.2), (3.0, 3.0, 3.3)])
def test_neighbour_cm_equals_brute_force(oracle, box):
    # (4.5, 5, 4.2): 1.2-wide cells (3+ per axis); (3, 3, 3.3): the oracle's
    # all-pairs path
    rng = np.random.default_rng(11)
    n = int(np.prod(box))  # ~1 bead per unit volume: some beads have no Hyphob neighbour
    pos = rng.random((n, 3)) * np.array(box)
    prop = rng.integers(0, 3, n).astype(np.uint32)
    p = _ref_params(box)
    _, cm, cnt = oracle.density(p, pos, prop, with_cm=True)
    hl, hb, cl, cb = cm_brute(pos, prop, box)
    assert np.array_equal(cnt[:, 0], cl) and np.array_equal(cnt[:, 1], cb)
    assert np.allclose(cm[:, :3], hl, rtol=0, atol=1e-12)
    assert np.allclose(cm[:, 3:], hb, rtol=0, atol=1e-12, equal_nan=True)
    assert (cb == 0).any() and (cb > 1).any() and (cl > 1).any()


def test_neighbour_cm_constructed_case(oracle):
    """Bead 0 (Hyphil) near the x face of a 4^3 box:
    - Hyphob at +1.1 x (inside 1.2, outside 1): hydrophobic only;
    - Hyphil at +1.1 y (outside 1): not counted;
    - Hyphil at +0.5 z: hydrophilic;
    - Hyphob at -0.6 x, across the face (stored at x = 3.7): hydrophobic, at
      its image x = -0.3;
    - solvent at -0.7 y: in neither sum.
    Hydrophilic CM of bead 0 = c + (0, 0, 0.25) with count 2; hydrophobic
    CM = c + (0.25, 0, 0) with count 2.  Bead 3 sees bead 0 (hydrophilic)
    and only the -0.6 x Hyphob (the +1.1 x one is 1.208 away)."""
    box = (4.0, 4.0, 4.0)
    c = np.array([0.3, 2.0, 2.0])
    off = np.array([[0, 0, 0], [1.1, 0, 0], [0, 1.1, 0], [0, 0, 0.5], [-0.6, 0, 0], [0, -0.7, 0]])
    pos = c + off
    pos -= np.floor(pos / 4.0) * 4.0
    assert pos[4, 0] > 3.0
    prop = np.array([HYPHIL, HYPHOB, HYPHIL, HYPHIL, HYPHOB, SOLVENT], np.uint32)
    p = _ref_params(box)
    _, cm, cnt = oracle.density(p, pos, prop, with_cm=True)
    assert tuple(cnt[0]) == (2, 2)
    assert np.allclose(cm[0, :3], c + [0, 0, 0.25], atol=1e-14)
    assert np.allclose(cm[0, 3:], c + [0.25, 0, 0], atol=1e-14)
    assert tuple(cnt[3]) == (2, 1)
    assert np.allclose(cm[3, :3], c + [0, 0, 0.25], atol=1e-14)
    assert np.allclose(cm[3, 3:], c + [-0.6, 0, 0], atol=1e-14)


def test_reference_density_per_species_closed_form(oracle):
    """Three beads of types Hyphil, Hyphob, solvent at r01 = 0.3, r02 = 0.6,
    r12 = sqrt(0.45): each bead's density of species k is (1 - r)^2 of its
    neighbour of type k, and zero for its own type (no self term)."""
    box = (5.0, 5.0, 5.0)
    c = np.array([2.0, 2.0, 2.0])
    pos = np.array([c, c + [0.3, 0, 0], c + [0, 0.6, 0]])
    prop = np.array([HYPHIL, HYPHOB, SOLVENT], np.uint32)
    p = _ref_params(box)
    dens = oracle.density(p, pos, prop)
    w01, w02, w12 = (1 - 0.3) ** 2, (1 - 0.6) ** 2, (1 - math.sqrt(0.45)) ** 2
    exp = np.array([[0.0, w01, w02], [w01, 0.0, w12], [w02, w12, 0.0]])
    assert np.allclose(dens, exp, rtol=1e-14, atol=0)


def _pair_force_case(oracle, p, ti, tj, r, vi, vj, seed3=321):
    """Two beads 0, 1 at distance r along a fixed unit vector; returns the
    oracle's acc and press of bead 0 and of bead 1."""
    e = np.array([2.0, -1.0, 2.0]) / 3.0
    c = np.array([3.0, 3.0, 3.0])
    pos = np.array([c, c - r * e])       # r_0 - r_1 = r e
    vel = np.array([vi, vj], dtype=np.float64)
    ids = np.array([7, 3], np.uint32)
    prop = np.array([ti, tj], np.uint32)
    dens = oracle.density(p, pos, prop)
    acc, press = oracle.force(p, pos, vel, ids, prop, dens, seed3)
    return e, pos, dens, acc, press


@pytest.mark.parametrize("ti,tj,r", [
    (HYPHIL, HYPHIL, 0.95),   # cf_m[0][0][0] = 0: no many-body term; cf_c past arc
    (SOLVENT, SOLVENT, 0.93), # cf_m[2][2][2] = 0
    (HYPHIL, HYPHOB, 0.5),    # inside arc: many-body only, 2 m (1-r)^2
    (HYPHOB, HYPHOB, 0.7),    # the Hyphob diagonal of cf_m is not zero
    (HYPHOB, SOLVENT, 0.97),
])
def test_reference_pair_force_closed_form(oracle, ti, tj, r):
    """F_0 = [(cf_c (r - arc) H(r - arc) + M) (1 - r) + sigma sqrt(1 - r) xi
    - gamma (1 - r) (e . v_01)] e with M = cf_m[ti][tj][ti] w + cf_m[ti][tj][tj] w
    (w = (1 - r)^2: each bead's density of the other's species), xi the
    pair's Saru normal at (min id, max id, seed); F_1 = -F_0; press_0 = F_0 * r e."""
    p = _ref_params((6.0, 6.0, 6.0))
    vi, vj = np.array([0.3, -1.1, 0.4]), np.array([-0.5, 0.2, 0.9])
    e, pos, dens, acc, press = _pair_force_case(oracle, p, ti, tj, r, vi, vj)
    r = float(np.linalg.norm(pos[0] - pos[1]))  # the stored positions' distance
    w = (1 - r) ** 2
    # each bead holds w in the other's species slot
    assert math.isclose(dens[0, tj], w, rel_tol=1e-14) and math.isclose(dens[1, ti], w, rel_tol=1e-14)
    m_off = p.mb[0][1][0]
    mb = lambda a, b, k: 0.0 if (a == b == k and a in (HYPHIL, SOLVENT)) else m_off
    assert p.mb[HYPHIL][HYPHIL][HYPHIL] == 0.0 and p.mb[SOLVENT][SOLVENT][SOLVENT] == 0.0
    M = mb(ti, tj, ti) * w + mb(ti, tj, tj) * w if ti != tj else mb(ti, tj, ti) * 2 * w
    cc = p.c1[ti][tj] * (r - 0.9) if r >= 0.9 else 0.0
    xi = oracle.pair_nrml(3, 7, 321)
    ev = float(e @ (vi - vj))
    f = (cc + M) * (1 - r) + p.sigma[ti][tj] * math.sqrt(1 - r) * xi - p.gamma[ti][tj] * (1 - r) * ev
    F0 = f * e
    assert np.allclose(acc[0], F0, rtol=1e-12, atol=1e-12)
    assert np.allclose(acc[1], -F0, rtol=1e-12, atol=1e-12)
    assert np.allclose(press[0], F0 * r * e, rtol=1e-12, atol=1e-12)
    assert np.allclose(press[1], press[0], rtol=1e-12, atol=1e-12)


def test_reference_pair_force_conservative_part_signs(oracle):
    """Without velocities and noise (sigma = 0): a Hyphil-Hyphil pair past
    arc feels only cf_c (r - arc)(1 - r), with the sign of cf_c[0][0] (the vbb
    term times the w2 normalisation, which is positive at arc = 0.9); a
    Hyphob pair inside arc feels only the many-body term, 2 m (1 - r)^3 > 0
    (repulsive: F_0 along r_0 - r_1)."""
    p = _ref_params((6.0, 6.0, 6.0))
    for i in range(3):
        for j in range(3):
            p.sigma[i][j] = 0.0
    z = np.zeros(3)
    e, _, _, acc, _ = _pair_force_case(oracle, p, HYPHIL, HYPHIL, 0.96, z, z)
    f = float(acc[0] @ e)
    arc = 0.9
    w2 = -45.0 / ((arc ** 5 * (2 * arc - 3) + (3 * arc - 2)) * math.pi)
    assert w2 > 0 and math.isclose(p.c1[0][0], 0.1 * 1.4 ** 3 / 16 * w2, rel_tol=1e-12)
    assert math.isclose(f, p.c1[0][0] * 0.06 * 0.04, rel_tol=1e-12) and f > 0
    e, _, _, acc, _ = _pair_force_case(oracle, p, HYPHOB, HYPHOB, 0.6, z, z)
    f = float(acc[0] @ e)
    assert math.isclose(f, 2 * p.mb[1][1][1] * 0.4 ** 3, rel_tol=1e-12) and f > 0


def test_reference_force_three_beads_unequal_densities(oracle):
    """Beads 0, 1 Hyphil at r01 = 0.6 and bead 2 Hyphob at r12 = 0.7 on one
    line (r02 = 1.3, outside the cut), so the two beads of a pair carry
    different densities: rho_0 = (w01, 0, 0), rho_1 = (w01, w12, 0),
    rho_2 = (w12, 0, 0).  Pair (0, 1): M = cf_m[0][0][1] (0 + w12) (the
    cf_m[0][0][0] term vanishes).  Pair (2, 1): M = m (w12 + w01) + m (0 + w12).
    No velocities, noise switched off (sigma = 0): the forces are
    conservative, along x, with F_1 = -(F_0 + F_2)."""
    p = _ref_params((6.0, 6.0, 6.0))
    for i in range(3):
        for j in range(3):
            p.sigma[i][j] = 0.0
    c = np.array([2.0, 3.0, 3.0])
    pos = np.array([c, c + [0.6, 0, 0], c + [1.3, 0, 0]])
    prop = np.array([HYPHIL, HYPHIL, HYPHOB], np.uint32)
    dens = oracle.density(p, pos, prop)
    r01 = pos[1, 0] - pos[0, 0]
    r12 = pos[2, 0] - pos[1, 0]
    w01, w12 = (1 - r01) ** 2, (1 - r12) ** 2
    assert np.allclose(dens, [[w01, 0, 0], [w01, w12, 0], [w12, 0, 0]], rtol=1e-13, atol=0)
    acc, _ = oracle.force(p, pos, np.zeros((3, 3)), np.arange(3), prop, dens, 1)
    m = p.mb[0][0][1]
    f0 = m * w12 * (1 - r01)                  # repulsive: bead 0 pushed to -x
    f2 = m * (w01 + 2 * w12) * (1 - r12)      # bead 2 pushed to +x
    assert np.allclose(acc[0], [-f0, 0, 0], rtol=1e-12, atol=1e-14)
    assert np.allclose(acc[2], [f2, 0, 0], rtol=1e-12, atol=1e-14)
    assert np.allclose(acc[1], -(acc[0] + acc[2]), rtol=1e-12, atol=1e-14)
