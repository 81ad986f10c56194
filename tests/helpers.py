"""Shared test inputs and independent checkers.

``chain_halo`` restates the constructive generator of the reference's tests
(test_fof.py:361-458): a chain in which each particle sits 0.5-0.95 b from
the previous one (so the chain must link), plus noise particles placed beyond
range + 5 b from the chain centre and more than b from each other, shuffled.
The expected membership is exact by construction.
"""
import numpy as np


def chain_halo(n_halo, b, n_noise, rng):
    xh = np.zeros((3, n_halo))
    for i in range(1, n_halo):
        r = rng.uniform(0.5 * b, 0.95 * b)
        th = rng.uniform(0, 2 * np.pi)
        ph = np.arccos(rng.uniform(-1, 1))
        xh[:, i] = xh[:, i - 1] + r * np.array(
            [np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
    lo, hi = xh.min(axis=1), xh.max(axis=1)
    centre = 0.5 * (lo + hi)
    dmin = (hi - lo).max() + 5 * b
    xn = np.zeros((3, n_noise))
    for i in range(n_noise):
        while True:
            th = rng.uniform(0, 2 * np.pi)
            ph = np.arccos(rng.uniform(-1, 1))
            r = rng.uniform(dmin, 2 * dmin)
            p = centre + r * np.array([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
            if i == 0 or np.all(np.sqrt(((xn[:, :i] - p[:, None]) ** 2).sum(0)) > b):
                xn[:, i] = p
                break
    xp = np.hstack([xh, xn])
    expect = np.zeros(n_halo + n_noise, bool)
    expect[:n_halo] = True
    perm = rng.permutation(n_halo + n_noise)
    return np.asfortranarray(xp[:, perm]), expect[perm]


def modulo(a, p):
    """gfortran real MODULO, elementwise (fmod then shift into sign of p)."""
    r = np.fmod(a, p)
    fix = (r != 0) & ((r < 0) != (p < 0))
    return np.where(fix, r + p, r)


def brute_force_labels(xp, L, b):
    """O(N^2) connected components of one region under the reference's
    predicate applied to EVERY pair (no cells), root = smallest input index.

    Valid as a checker when the reference's cells are at least b (+ f32
    rounding) wide, so that every pair within b sits in adjacent cells.
    Returns labels[i] = smallest input index in i's component.
    """
    n = xp.shape[1]
    Lh = 0.5 * L
    xr = modulo(xp - xp[:, :1] + Lh, L) - Lh
    rng_ = xr.max(axis=1) - xr.min(axis=1)
    periodic = bool(np.any(rng_ + 2 * b >= L))
    xc = xr.astype(np.float32)
    parent = np.arange(n)

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for i in range(n):
        d32 = xc[:, i + 1:] - xc[:, i:i + 1]  # float32 differences
        if periodic:
            d = np.abs(d32.astype(np.float64))
            d = np.where(d > Lh, L - d, d)
            r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        else:
            s = d32[0] * d32[0]
            s = s + d32[1] * d32[1]
            s = s + d32[2] * d32[2]
            r2 = s.astype(np.float64)
        for j in np.nonzero(r2 <= b * b)[0] + i + 1:
            ri, rj = find(i), find(j)
            if ri != rj:
                parent[max(ri, rj)] = min(ri, rj)
    return np.array([find(i) for i in range(n)])


def labels_to_canonical(labels):
    """Map any labelling to 'smallest member index' labels."""
    labels = np.asarray(labels)
    n = len(labels)
    order = np.argsort(labels, kind="stable")
    sl = labels[order]
    starts = np.concatenate([[True], sl[1:] != sl[:-1]])
    first = order[starts]
    grp = np.cumsum(starts) - 1
    out = np.empty(n, dtype=np.int64)
    mins = np.minimum.reduceat(order, np.nonzero(starts)[0])
    out[order] = mins[grp]
    del first
    return out
