"""Pins the C oracle (FoF part) before anything is compared against it.

Checks: the reference's constructive known-answer tests (test_fof.py:653-729,
804-866), brute-force O(N^2) components with the reference predicate,
scipy cKDTree + connected_components on inputs with no near-threshold pair,
the cell-geometry rule (fof_fortran.f90:518-570) in closed form, halo
numbering order (fof_fortran.f90:866-872, 209-223) and the angular-momentum
closed form (fof_fortran.f90:224-248).
"""
import numpy as np
import pytest

from helpers import (TINY_GRID_CASES, brute_force_labels, chain_halo, labels_to_canonical, modulo,
                     wrapped_clumps)
from oracle import oracle_fof, oracle_geometry

B = 2.0
L = 512.0


def test_chain_halo_known_answer():
    rng = np.random.default_rng(999)
    xp, expect = chain_halo(200, B, 400, rng)
    in_halo, hid, am = oracle_fof(xp, [xp.shape[1]], L, B, 10,
                                  vp=np.zeros_like(xp), mp=np.ones(xp.shape[1]))
    assert np.array_equal(in_halo.astype(bool), expect)
    assert np.all(am == 0.0)
    assert set(np.unique(hid)) == {0, 1}


def test_multiple_regions_known_answer():
    shifts = [(-50.0, -50.0, 0.0), (50.0, -50.0, 0.0), (0.0, 50.0, 30.0)]
    xs, es = [], []
    for k, s in enumerate(shifts):
        xp, e = chain_halo(100, B, 200, np.random.default_rng(777 + k))
        xs.append(xp + np.array(s)[:, None])
        es.append(e)
    xp = np.asfortranarray(np.hstack(xs))
    in_halo, hid, _ = oracle_fof(xp, [300, 300, 300], L, B, 10)
    assert np.array_equal(in_halo.astype(bool), np.concatenate(es))
    # one halo per region, numbered in region order
    for k in range(3):
        ids = np.unique(hid[300 * k:300 * (k + 1)])
        assert list(ids) == [0, k + 1]


def _partition_equal(a, b):
    return np.array_equal(labels_to_canonical(a), labels_to_canonical(b))


@pytest.mark.parametrize("seed,box,n,Lb", [
    (1, 6.0, 1500, 512.0),    # compact, non-periodic
    (2, 64.0, 1200, 64.0),    # uniform over the whole box: periodic
    (3, 3.0, 2000, 512.0),    # dense: cells clamp at L_mesh = b
    (4, 20.0, 800, 20.5),     # periodic in every dim, few cells
])
def test_partition_vs_brute_force(seed, box, n, Lb):
    rng = np.random.default_rng(seed)
    b = 0.35 if box > 10 else 0.2
    xp = np.asfortranarray(rng.uniform(0, box, (3, n)) + 5.0)
    _, _, _, lab = oracle_fof(xp, [n], Lb, b, 1, with_labels=True)
    ref = brute_force_labels(xp, Lb, b)
    assert _partition_equal(lab, ref)


def test_partition_periodic_one_dim():
    # a slab: wraps in x only, so the periodic predicate is used for all dims
    rng = np.random.default_rng(5)
    n = 1500
    xp = np.stack([rng.uniform(0, 30.0, n), rng.uniform(0, 4.0, n), rng.uniform(0, 4.0, n)])
    xp = np.asfortranarray(xp)
    g = oracle_geometry(xp, 30.0, 0.3)
    assert g["periodic"] == [True, False, False]
    _, _, _, lab = oracle_fof(xp, [n], 30.0, 0.3, 1, with_labels=True)
    assert _partition_equal(lab, brute_force_labels(xp, 30.0, 0.3))


@pytest.mark.parametrize("case", range(4))
def test_partition_tiny_periodic_grids_vs_brute_force(case):
    # 1 or 2 cells per axis: each neighbour cell is visited once
    # (fof_fortran.f90:729-755); every pair lies in adjacent cells, so the
    # all-pairs brute force is the reference partition
    L, b, pc, cs, nbg, sig, nd = TINY_GRID_CASES[case]
    xp = wrapped_clumps(L, pc, cs, nbg, seed=case + 1, sigma=sig)
    n = xp.shape[1]
    g = oracle_geometry(xp, L, b)
    assert g["n"] == [nd] * 3 and g["periodic"] == [True] * 3
    _, _, _, lab = oracle_fof(xp, [n], L, b, 1, with_labels=True)
    assert _partition_equal(lab, brute_force_labels(xp, L, b))
    assert len(np.unique(lab)) >= 2


def test_partition_vs_ckdtree():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(7)
    Lb, b, n = 32.0, 0.3, 40000
    xp = rng.uniform(0, Lb, (3, n))
    # clumps so that large groups exist
    cen = rng.uniform(0, Lb, (3, 40))
    own = rng.integers(0, 40, 15000)
    xp[:, :15000] = np.mod(cen[:, own] + 0.4 * rng.standard_normal((3, 15000)), Lb)
    xp = np.asfortranarray(xp)
    t = cKDTree(xp.T, boxsize=Lb)
    near = t.query_pairs(b * (1 + 1e-5), output_type="ndarray")
    far = t.query_pairs(b * (1 - 1e-5), output_type="ndarray")
    # pairs within 1e-5 of b are decided by the f32 arithmetic of the predicate;
    # there are few, decide them with that arithmetic and keep the rest textbook
    if len(near) != len(far):
        fs = set(map(tuple, far))
        band = np.array([p for p in map(tuple, near) if p not in fs])
        Lh = 0.5 * Lb
        xr = (modulo(xp - xp[:, :1] + Lh, Lb) - Lh).astype(np.float32)
        d = np.abs((xr[:, band[:, 0]] - xr[:, band[:, 1]]).astype(np.float64))
        d = np.where(d > Lh, Lb - d, d)
        ok = ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) <= b * b
        far = np.concatenate([far, band[ok]])
    g = coo_matrix((np.ones(len(far)), (far[:, 0], far[:, 1])), shape=(n, n))
    _, comp = connected_components(g, directed=False)
    _, _, _, lab = oracle_fof(xp, [n], Lb, b, 1, with_labels=True)
    assert _partition_equal(lab, comp)


def test_numbering_follows_crs_root_order():
    rng = np.random.default_rng(11)
    Lb, b, n = 16.0, 0.3, 6000
    xp = rng.uniform(0, Lb, (3, n))
    cen = rng.uniform(0, Lb, (3, 30))
    own = rng.integers(0, 30, 3000)
    xp[:, :3000] = np.mod(cen[:, own] + 0.5 * rng.standard_normal((3, 3000)), Lb)
    xp = np.asfortranarray(xp)
    in_halo, hid, _, lab = oracle_fof(xp, [n], Lb, b, 5, with_labels=True)
    g = oracle_geometry(xp, Lb, b)
    Lh = 0.5 * Lb
    xr = modulo(xp - xp[:, :1] + Lh, Lb) - Lh
    ic = []
    for d in range(3):
        i = np.floor((xr[d] - g["origin"][d]) / g["cell"][d]).astype(np.int64) + 1
        ic.append(np.clip(i, 1, g["n"][d]))
    icell = ic[0] + g["n"][0] * (ic[1] - 1) + g["n"][0] * g["n"][1] * (ic[2] - 1)
    key = icell * n + np.arange(n)
    ids = np.unique(hid[hid > 0])
    assert list(ids) == list(range(1, len(ids) + 1))
    # each halo's root (label) is its min-key member, ids ascend with that key
    mins = []
    for h in ids:
        m = np.nonzero(hid == h)[0]
        kmin = m[np.argmin(key[m])]
        assert np.all(lab[m] == kmin)
        mins.append(key[kmin])
    assert np.all(np.diff(mins) > 0)
    # size threshold
    sizes = np.bincount(labels_to_canonical(lab))
    assert in_halo.sum() == sizes[sizes >= 5].sum()


def test_geometry_rule_closed_form():
    # uniform filling of the periodic box: cells = mean spacing
    rng = np.random.default_rng(0)
    xp = np.asfortranarray(rng.uniform(0, 64.0, (3, 262144)))
    g = oracle_geometry(xp, 64.0, 0.2)
    assert g["n"] == [64, 64, 64] and g["cell"] == [1.0, 1.0, 1.0]
    assert g["origin"] == [-32.0] * 3 and g["is_periodic"]
    # dense compact region: L_mesh clamps to b, non-periodic, origin = bbox min
    xp = np.asfortranarray(rng.uniform(0, 12.0, (3, 10000)) + 50.0)
    g = oracle_geometry(xp, 512.0, 2.0)
    assert g["cell"] == [2.0, 2.0, 2.0] and not g["is_periodic"]
    rngs = np.array(g["bmax"]) - np.array(g["bmin"])
    assert g["n"] == [int(np.ceil(r / 2.0)) for r in rngs]
    assert g["origin"] == g["bmin"]
    # sparse region triggers the refinement rule (N^3 > max(262144, 4 np))
    xp = np.asfortranarray(rng.uniform(0, 400.0, (3, 100)))
    g = oracle_geometry(xp, 512.0, 0.2)
    rngs = np.array(g["bmax"]) - np.array(g["bmin"])
    vol = np.prod(np.maximum(rngs, 0.2))
    lm1 = max(0.2, (vol / 100) ** (1 / 3))
    assert np.prod(np.ceil(rngs / lm1)) <= 262144 or g["cell"][0] == max(0.2, (vol / 200) ** (1 / 3))


def test_ang_mom_closed_form():
    rng = np.random.default_rng(3)
    xp, _ = chain_halo(120, B, 30, rng)
    xp = xp + 200.0
    vp = np.asfortranarray(rng.standard_normal(xp.shape))
    mp = rng.uniform(0.5, 2.0, xp.shape[1])
    in_halo, hid, am = oracle_fof(xp, [xp.shape[1]], L, B, 10, vp=vp, mp=mp)
    m = np.nonzero(hid == 1)[0]
    w = mp[m]
    x = xp[:, m] - xp[:, m].mean(axis=1, keepdims=True)
    x = x - (w * x).sum(1, keepdims=True) / w.sum()
    v = vp[:, m] - (w * vp[:, m]).sum(1, keepdims=True) / w.sum()
    expect = (w * np.cross(x.T, v.T).T).sum(axis=1)
    assert np.allclose(am[:, 0], expect, rtol=1e-10, atol=1e-10)


def test_empty_and_small_regions():
    xp = np.asfortranarray(np.random.default_rng(0).uniform(0, 5, (3, 9)))
    in_halo, hid, am = oracle_fof(xp, [9], L, B, 10)
    assert in_halo.sum() == 0 and hid.sum() == 0
    in_halo, hid, am = oracle_fof(xp, [4, 5], L, 10.0, 3)
    assert list(hid) == [1] * 4 + [2] * 5
    # regions below the threshold are skipped entirely
    in_halo, hid, am = oracle_fof(xp, [2, 7], L, 10.0, 3)
    assert list(hid) == [0] * 2 + [1] * 7
    in_halo, hid, am = oracle_fof(np.zeros((3, 0), order="F"), [0], L, B, 1)
    assert len(hid) == 0
