"""numpy references of the observables of k_obs.hip: the same FP64
expressions (minimum image by half-box comparisons, squared lengths summed
x, y, z in order, bins against the squared edges (k w)^2, cosine bins
decided without a square root), so counts must agree exactly."""
import numpy as np


def min_image(v, box):
    v = v.copy()
    for a in range(3):
        hi = v[:, a] > 0.5 * box[a]
        lo = v[:, a] < -0.5 * box[a]
        v[hi, a] -= box[a]
        v[lo, a] += box[a]
    return v


def sq(v):
    return v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]


def edges(nb, width):
    x = np.arange(nb + 1) * width
    return x * x


def hist_sq(l2, e2):
    m = (l2 >= e2[0]) & (l2 < e2[-1])
    k = np.searchsorted(e2, l2[m], side="right") - 1
    return np.bincount(k, minlength=len(e2) - 1).astype(float)


def rdf_ref(pos, box, rmax, nb):
    n = len(pos)
    e2 = edges(nb, rmax / nb)
    h = np.zeros(nb)
    for i in range(n):
        v = min_image(pos[i] - pos, box)
        l2 = sq(v)
        l2 = np.delete(l2, i)
        h += hist_sq(l2, e2)
    return h


def cos_bins(dot, l, nb):
    k = np.arange(nb)
    ck = -1.0 + 2.0 * k / nb
    d = dot[:, None]
    ll = l[:, None]
    pos_edge = (d >= 0.0) & (d * d >= ck * ck * ll)
    neg_edge = (d >= 0.0) | (d * d <= ck * ck * ll)
    ge = np.where(ck >= 0.0, pos_edge, neg_edge)
    return np.maximum(ge.sum(axis=1) - 1, 0)


def chain_ref(pos, topo, box, nb, bmax):
    x = pos[topo.reshape(-1, 4)]
    eb2 = edges(nb, bmax / nb)
    ee2 = edges(nb, 3.0 * bmax / nb)
    out = np.zeros((4, nb))
    b = [min_image(x[:, k + 1] - x[:, k], box) for k in range(3)]
    l2 = [sq(v) for v in b]
    out[0] += hist_sq(l2[0], eb2)
    out[1] += hist_sq(l2[1], eb2) + hist_sq(l2[2], eb2)
    for k in range(2):
        dot = b[k][:, 0] * b[k + 1][:, 0] + b[k][:, 1] * b[k + 1][:, 1] + b[k][:, 2] * b[k + 1][:, 2]
        ll = l2[k] * l2[k + 1]
        ok = ll != 0.0
        out[2] += np.bincount(cos_bins(dot[ok], ll[ok], nb), minlength=nb)
    e = min_image(x[:, 3] - x[:, 0], box)
    out[3] += hist_sq(sq(e), ee2)
    return out


def count_ref(pos, prop, box, g, mask):
    sel = ((mask >> prop.astype(np.int64)) & 1) == 1
    p = pos[sel]
    idx = []
    for a in range(3):
        edge = box[a] / g[a]
        k = (p[:, a] / edge).astype(np.int64)
        idx.append(np.clip(k, 0, g[a] - 1))
    flat = (idx[2] * g[1] + idx[1]) * g[0] + idx[0]
    return np.bincount(flat, minlength=g[0] * g[1] * g[2]).reshape(g[2], g[1], g[0]).astype(float)


def rdf_from_pairs(pos, box, i, j, rmax, nb):
    """rdf_ref over a given ordered pair list (a superset of the pairs
    inside rmax), for systems too large for the all-pairs loop."""
    e2 = edges(nb, rmax / nb)
    return hist_sq(sq(min_image(pos[i] - pos[j], box)), e2)
