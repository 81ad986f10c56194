"""Pins the oracle's incremental potential updates and the outer-pass cap.

After the first pass the reference never recomputes a potential: each
refresh subtracts the contribution of the members deleted since the last one.
Three paths do that:
  * direct summation adds the deleted sources back (bound_fortran.f90:2630-2641);
  * the C2 residual subtracts deleted neighbours over 27 cells
    (bound_fortran.f90:3490-3552, 3316-3386);
  * the two-level short-range sums do the same with per-target cutoffs
    (bound_fortran.f90:1345-1498).
Each is checked here against a recomputation from scratch over the members
still kept (the quantity the update must equal), one removal batch at a time,
and end to end: the whole removal loop run with every refresh recomputed
gives the same statuses and members as the incremental loop. A constructed
halo reaches the 20-pass cap (status 5, bound_fortran.f90:2857-2864) with a
retained count fixed by construction.
"""
import numpy as np
import pytest

from helpers import max_iter_halo
from oracle import (oracle_c2_check, oracle_direct_check, oracle_fof_bound, oracle_set_force_full,
                    oracle_tl_pp_check)

L = 512.0


def _clump(rng, n, scale):
    """(3, n) f32 positions: a Gaussian clump with a denser core."""
    q = rng.standard_normal((3, n)) * scale
    q[:, : n // 4] *= 0.25
    return q.astype(np.float32)


def _batch(rng, n, keep_frac=0.9, rem_frac=0.05):
    keep0 = rng.uniform(size=n) < keep_frac
    kept = np.nonzero(keep0)[0]
    rem = rng.choice(kept, max(1, int(rem_frac * len(kept))), replace=False)
    keep1 = keep0.copy()
    keep1[rem] = False
    return keep0, rem, keep1


def _assert_update_equals_full(inc, full, before, keep1, rtol):
    """inc equals the recomputation on the kept members to rounding, and
    the batch moved the potentials by far more than that tolerance (so an
    update that does nothing, or subtracts the wrong terms, fails)."""
    k = keep1
    # every term of a sum has one sign, so |before| bounds the sum of the
    # magnitudes the rounding of either path scales with
    scale = np.maximum(np.abs(before[k]), np.abs(full[k])).astype(np.float64)
    tol = rtol * scale + 1e-30
    err = np.abs(inc[k].astype(np.float64) - full[k])
    assert np.all(err <= tol), (float(np.max(err / tol)), int(np.sum(err > tol)))
    moved = np.abs(before[k].astype(np.float64) - full[k]) > 100.0 * tol
    assert moved.mean() > 0.2, moved.mean()


@pytest.mark.parametrize("seed", [1, 2])
def test_c2_incremental_equals_full_recompute(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    q = _clump(rng, n, 0.4)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    sp, app = 0.05, 0.02
    keep0, rem, keep1 = _batch(rng, n)
    inc, full = oracle_c2_check(q, w, keep0, rem, sp, app)
    before, same = oracle_c2_check(q, w, keep0, [], sp, app)
    assert np.array_equal(before, same)
    assert np.all(full[keep1] <= 0)  # every residual term is negative here
    _assert_update_equals_full(inc, full, before, keep1, rtol=2e-6)


@pytest.mark.parametrize("seed", [3, 4])
def test_two_level_pp_update_equals_full_recompute(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    q = _clump(rng, n, 0.6)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    sp, app = 0.1, 0.02  # 3.5 sp / 4 > app: ratio 4 is not demoted
    ratios = rng.choice([1, 2, 4], n)
    keep0, rem, keep1 = _batch(rng, n)
    inc, full = oracle_tl_pp_check(q, w, keep0, rem, sp, app, ratios)
    before, _ = oracle_tl_pp_check(q, w, keep0, [], sp, app, ratios)
    _assert_update_equals_full(inc, full, before, keep1, rtol=2e-6)
    # the per-target cutoff matters: all ratios 1 gives other sums
    inc1, _ = oracle_tl_pp_check(q, w, keep0, rem, sp, app, np.ones(n, np.int32))
    assert np.mean(np.abs(inc1[keep1] - inc[keep1]) > 1e-3 * np.abs(inc[keep1])) > 0.2


def test_direct_addback_equals_fresh_sum():
    rng = np.random.default_rng(5)
    n = 700
    q = _clump(rng, n, 0.5)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    keep0, rem, keep1 = _batch(rng, n, keep_frac=1.0, rem_frac=0.1)
    inc, full = oracle_direct_check(q, w, keep0, rem, 0.05)
    before, _ = oracle_direct_check(q, w, keep0, [], 0.05)
    _assert_update_equals_full(inc, full, before, keep1, rtol=1e-12)
    # a second batch on top of the first (the reference adds back every batch)
    kept = np.nonzero(keep1)[0]
    rem2 = rng.choice(kept, 50, replace=False)
    inc2, full2 = oracle_direct_check(q, w, keep1, rem2, 0.05)
    keep2 = keep1.copy()
    keep2[rem2] = False
    _assert_update_equals_full(inc2, full2, full, keep2, rtol=1e-12)


def test_removal_with_invalid_batch_is_refused():
    q = np.zeros((3, 4), np.float32)
    w = np.ones(4, np.float32)
    with pytest.raises(ValueError):
        oracle_c2_check(q, w, np.array([1, 1, 0, 1]), [2], 0.1, 0.02)  # not kept


def _isolated(sizes, seed):
    from py_fof_amd.synth import isolated_halos
    xs, vs = isolated_halos(seed=seed, sizes=np.asarray(sizes), L=L)
    xp = np.asfortranarray(np.hstack(xs))
    vp = np.asfortranarray(np.hstack(vs))
    return xp, vp, [x.shape[1] for x in xs]


@pytest.mark.parametrize("sizes,seed", [([1500, 2600, 4000, 7000], 31), ([60, 300, 900, 1000], 12)])
def test_removal_loop_incremental_equals_recomputed(sizes, seed):
    """End to end: the reference's incremental refreshes and refreshes
    recomputed from scratch give the same bound members. The f32 sums are
    formed in another order, so a member whose energy sits within rounding of
    zero may differ; none does on these inputs."""
    xp, vp, nps = _isolated(sizes, seed)
    s15 = 15.0 * 0.67 * L / 300.0
    try:
        oracle_set_force_full(False)
        inc = oracle_fof_bound(xp, nps, L, 0.25, 10, vp, s15=s15)
        oracle_set_force_full(True)
        full = oracle_fof_bound(xp, nps, L, 0.25, 10, vp, s15=s15)
    finally:
        oracle_set_force_full(False)
    assert inc["ierr"] == 0 and full["ierr"] == 0
    assert np.array_equal(inc["gid"], full["gid"])
    assert np.array_equal(inc["status"], full["status"])
    assert np.array_equal(inc["np_bound"], full["np_bound"])
    assert np.array_equal(inc["in_halo"], full["in_halo"])
    # the loops really ran several refreshes: members were removed
    assert np.sum(inc["np_bound"]) < np.sum(nps)


def test_max_iter_status_and_retained_count():
    # 24 shells: each pass removes the outermost shell; after 20 passes 4
    # shells and the core (121) remain -> status 5, not written as a halo
    x, v, retained = max_iter_halo(n_shells=24)
    for force in (False, True):
        try:
            oracle_set_force_full(force)
            r = oracle_fof_bound(x, [x.shape[1]], L, 0.2, 10, v, app=0.1)
        finally:
            oracle_set_force_full(False)
        assert r["ierr"] == 0
        assert list(r["gid"]) == [0] and list(r["status"]) == [5]
        assert list(r["np_bound"]) == [retained] == [145]
        assert r["in_halo"].sum() == 0 and r["halo_id"].sum() == 0
    # 16 shells: 16 passes remove them all, pass 17 finds none -> OK, the core
    x, v, _ = max_iter_halo(n_shells=16)
    r = oracle_fof_bound(x, [x.shape[1]], L, 0.2, 10, v, app=0.1)
    assert list(r["status"]) == [0] and list(r["np_bound"]) == [121]
    assert r["in_halo"].sum() == 121
