"""Pins the oracle's unbinding restatement (bound_fortran.f90) before any GPU
result is compared against it.

Known answers of the reference's own tests (test_fof.py:734-799, 871-969,
1022-1169); closed-form properties of the softened depth kernel
(bound_fortran.f90:3080-3102); the Green function's real-space response
(-1/r between the split softening and the truncation radius,
bound_fortran.f90:2300-2382); direct-sum potentials against an f64 numpy
brute force; PM+C2 potentials against that brute force within method error;
the grid/level selection table at s_15 = 17.152 (bound_fortran.f90:1854-2040).
"""
import numpy as np
import pytest

from helpers import chain_halo
from oracle import oracle_depth, oracle_fof_bound, oracle_green, oracle_select, oracle_tl_info

B = 2.0
L = 512.0


def _with_highv(seed, n_halo, n_noise, frac=0.1):
    rng = np.random.default_rng(seed)
    xp, expect = chain_halo(n_halo, B, n_noise, rng)
    vp = rng.standard_normal(xp.shape) * 0.1
    hv = rng.choice(np.nonzero(expect)[0], max(1, int(n_halo * frac)), replace=False)
    vp[:, hv] *= 100.0
    return xp, np.asfortranarray(vp), expect, hv


def test_known_answer_high_velocity_removed():
    xp, vp, expect, hv = _with_highv(888, 200, 400)
    r = oracle_fof_bound(xp, [xp.shape[1]], L, B, 10, vp, app=0.1)
    assert r["ierr"] == 0
    assert r["in_halo"][hv].sum() == 0
    assert r["in_halo"].sum() == 180
    assert np.all(r["in_halo"][~expect] == 0)
    assert list(r["status"]) == [0] and list(r["gid"]) == [0]


def test_known_answer_three_regions():
    xs, vs, hvs, start = [], [], [], 0
    shifts = [(-50.0, -50.0, 0.0), (50.0, -50.0, 0.0), (0.0, 50.0, 30.0)]
    for k, s in enumerate(shifts):
        xp, vp, expect, hv = _with_highv(666 + k, 100, 200)
        xs.append(xp + np.array(s)[:, None])
        vs.append(vp)
        hvs.append(hv + start)
        start += xp.shape[1]
    xp = np.asfortranarray(np.hstack(xs))
    vp = np.asfortranarray(np.hstack(vs))
    r = oracle_fof_bound(xp, [300] * 3, L, B, 10, vp)
    assert r["ierr"] == 0 and r["in_halo"].sum() == 270
    assert r["in_halo"][np.concatenate(hvs)].sum() == 0
    assert sorted(np.unique(r["halo_id"])) == [0, 1, 2, 3]


def test_masses_scalar_equals_array_and_heavy_centre_binds_more():
    rng = np.random.default_rng(202)
    xp, _ = chain_halo(100, B, 0, rng)
    vp = np.asfortranarray(rng.standard_normal((3, 100)) * 0.25)
    light = oracle_fof_bound(xp, [100], L, B, 10, vp, mp=np.ones(100))
    same = oracle_fof_bound(xp, [100], L, B, 10, vp)
    assert np.array_equal(light["in_halo"], same["in_halo"])
    mh = np.ones(100)
    mh[0] = 200.0
    heavy = oracle_fof_bound(xp, [100], L, B, 10, vp, mp=mh)
    assert heavy["in_halo"].sum() >= light["in_halo"].sum()


def test_span_over_s15_raises_ierr2_and_small_ng_ierr1():
    xline = np.vstack([np.linspace(10.0, 35.0, 1050), np.zeros(1050), np.zeros(1050)])
    vz = np.zeros((3, 1050))
    r = oracle_fof_bound(np.asfortranarray(xline), [1050], L, B, 10, vz, s15=17.152)
    assert r["ierr"] == 2 and r["in_halo"].sum() == 0
    xp, vp, _, _ = _with_highv(1, 50, 10)
    assert oracle_fof_bound(xp, [60], L, B, 10, vp, ng=128)["ierr"] == 1
    # no candidate at all: the batch is never entered, no error
    far = np.asfortranarray(np.random.default_rng(0).uniform(0, 500, (3, 5)))
    assert oracle_fof_bound(far, [5], L, B, 10, np.zeros((3, 5)), ng=128)["ierr"] == 0


def test_depth_kernel_closed_form():
    h = 0.37
    D = np.vectorize(lambda r: oracle_depth(r, h))
    assert D(0.0) == pytest.approx(208.0 / (70.0 * h), rel=1e-15)
    # Newtonian outside h, continuous (value and slope) at e = 1 and at r = h
    r_out = np.array([h, 1.3 * h, 5 * h])
    assert np.allclose(D(r_out), 1.0 / r_out, rtol=1e-14)
    for rb in (0.5 * h, h):
        lo, hi = D(rb * (1 - 1e-9)), D(rb * (1 + 1e-9))
        assert abs(lo - hi) < 1e-7 / h
        eps = 1e-6 * h
        s_lo = (D(rb - eps) - D(rb - 2 * eps)) / eps
        s_hi = (D(rb + 2 * eps) - D(rb + eps)) / eps
        assert abs(s_lo - s_hi) < 1e-4 / h ** 2
    # the smoothing density rho = lap(D)/(4 pi) is non-negative and has unit mass
    r = np.linspace(1e-4 * h, h, 4001)
    dr = r[1] - r[0]
    phi = -D(r)
    g = np.gradient(phi, dr)
    lap = np.gradient(r * r * g, dr) / (r * r)
    rho = lap / (4 * np.pi)
    assert rho[5:-5].min() > -1e-3 / h ** 3
    # Gauss: the mass inside r is r^2 dphi/dr; it reaches 1 at r = h
    eps = 1e-7 * h
    for rr, m_expect in ((h, 1.0), (0.5 * h, None)):
        m = rr * rr * (D(rr - eps) - D(rr + eps)) / (2 * eps)
        if m_expect is not None:
            assert m == pytest.approx(m_expect, rel=1e-6)
        else:
            assert 0.0 < m < 1.0
    # and the density integrates to the same mass
    m_int = np.trapezoid(rho[5:-5] * 4 * np.pi * r[5:-5] ** 2, r[5:-5])
    assert m_int == pytest.approx(1.0, rel=5e-3)


@pytest.mark.parametrize("n", [48, 64])
def test_green_real_space_response(n):
    G, self_ = oracle_green(n)
    nh = n // 2 + 1
    # Hermitian extension of the half spectrum, inverse transform (numpy pocketfft)
    full = np.zeros((n, n, n))
    full[:, :, :nh] = G
    for x in range(nh, n):
        full[:, :, x] = G[(-np.arange(n)) % n][:, (-np.arange(n)) % n][:, :, n - x]
    g = np.real(np.fft.ifftn(full))
    # response of a unit mass at the origin along x: the Newtonian -1/r shifted
    # to vanish at the truncation radius R = n/2 (-1/r + 1/R), between the
    # split softening (3.5 cells) and R; the constant is what the reference's
    # zero-mode term (M_keep - w_i)/(n/2) restores
    R = n / 2
    r = np.arange(9, n // 2 - 3)
    assert np.allclose(g[0, 0, r], -1.0 / r + 1.0 / R, rtol=0.03, atol=2e-4)
    # truncated: no response beyond 3/4 of n/2 along the diagonal wrap region
    assert np.abs(g[0, 0, n // 2]) < 0.2 / (n / 2)
    # the 5^3 self kernel is the centre of that response
    c = np.array([(-2 + i) % n for i in range(5)])
    assert np.allclose(self_, g[np.ix_(c, c, c)], rtol=1e-5, atol=1e-6)
    assert self_[2, 2, 2] < 0 and np.argmin(self_) == 62


def _direct_phi_f64(q, w, a, app):
    """Textbook O(N^2) softened potential in f64 (independent of the oracle)."""
    d = q[:, :, None] - q[:, None, :]
    r = np.sqrt((d * d).sum(0))
    h = app
    e = 2 * r / h
    with np.errstate(divide="ignore", invalid="ignore"):
        inner = (208 - e ** 2 * (112 + e ** 2 * (-56 + e * (14 + e * (8 - 3 * e))))) / (70 * h)
        outer = (128 - (e ** 2 * (448 + e * (-280 + e * (56 + e * (14 + e * (-8 + e))))) - 224 * e - 12 / e)) / (70 * h)
        D = np.where(r >= h, 1 / r, np.where(e <= 1, inner, outer))
    np.fill_diagonal(D, 0.0)
    return -(a / (6 * np.pi)) * (D * w[None, :]).sum(1)


def _nfw_halo(rng, n, rvir, c=8.0):
    from py_fof_amd.synth import isotropic, nfw_radii
    return nfw_radii(rng, n, c, rvir) * isotropic(rng, n)


def test_direct_potential_vs_numpy():
    rng = np.random.default_rng(5)
    x = _nfw_halo(rng, 600, 1.0) + 100.0
    v = np.zeros_like(x)
    r = oracle_fof_bound(np.asfortranarray(x), [600], L, 0.4, 10, np.asfortranarray(v),
                         app=0.05, want_phi=True)
    assert list(r["gid"]) == [0]
    mem = np.nonzero(r["in_halo"] | (r["halo_id"] >= 0))[0]
    # members of the (single) candidate, pid order = index order
    q = (x - x[:, :1]).astype(np.float32).astype(np.float64)
    ref = _direct_phi_f64(q, np.ones(600), 1.0, 0.05)
    got = r["phi"][:600].astype(np.float64)
    assert np.allclose(got, ref, rtol=2e-6)
    del mem


def test_pm_potential_vs_direct_within_method_error():
    rng = np.random.default_rng(9)
    n = 3000
    x = _nfw_halo(rng, n, 2.0) + 200.0
    r = oracle_fof_bound(np.asfortranarray(x), [n], L, 0.5, 10, np.zeros((3, n)), app=0.1,
                         want_phi=True)
    assert r["ierr"] == 0 and r["gid"][0] >= 1
    q = (x - x[:, :1]).astype(np.float32).astype(np.float64)
    ref = _direct_phi_f64(q, np.ones(n), 1.0, 0.1)
    got = r["phi"][:n].astype(np.float64)
    rel = np.abs(got - ref) / np.abs(ref)
    assert np.median(rel) < 5e-3 and rel.max() < 3e-2


def test_selection_table():
    # default ladders at s_15 = 17.152 (fixed bands first, then the global ladder)
    assert oracle_select(900, 50.0)[:2] == (0, 0)            # direct summation
    assert oracle_select(2000, 2.0)[1:] == (1, 3, 0.125)     # 48 grid, 1/8 (cover 2.625)
    assert oracle_select(2000, 3.0)[1:] == (2, 3, 0.125)     # 64 grid, 1/8 (cover 3.625)
    assert oracle_select(8000, 3.0)[1:] == (3, 2, 0.0625)    # 128 grid, 1/16 (cover 3.8125)
    assert oracle_select(8000, 9.0)[1:] == (4, 3, 0.125)     # 256 grid, 1/8 (cover 15.625)
    st, g, lev, sp = oracle_select(30000, 17.0)
    assert (st, g) == (0, 4) and sp == pytest.approx(17.152 / 125)
    assert oracle_select(30000, 17.2)[0] == 2


def test_two_level_block_ratio_rule():
    # bound_fortran.f90:551-591: blocks of 16^3 coarse cells; < 3926 members
    # ratio 1, < 14688 ratio 2, else 4; demoted while 3.5 sp / r < softening
    sp = 0.125
    origin = np.zeros(3, np.float32)
    counts = {0: 3925, 1: 3926, 2: 14687, 3: 14688}
    pts = []
    for b, c in counts.items():
        # block b along x: cells [16 b, 16 b + 16); members at the block centre
        centre = np.array([(16 * b + 8) * sp, 8 * sp, 8 * sp])
        pts.append(np.repeat(centre[:, None], c, axis=1))
    q = np.hstack(pts).astype(np.float32)
    na, cnt, rat = oracle_tl_info(q, sp, origin, 0.1)
    assert list(cnt[:4]) == [3925, 3926, 14687, 14688] and cnt[4:].sum() == 0
    assert list(rat[:4]) == [1, 2, 2, 4] and na == 3
    # 3.5 * 0.125 / 4 = 0.109: ratio 4 survives app 0.1, falls to 2 at app 0.11,
    # and 3.5 * 0.125 / 2 = 0.219 < 0.22 demotes ratio 2 to 1
    assert list(oracle_tl_info(q, sp, origin, 0.11)[2][:4]) == [1, 2, 2, 2]
    na2, _, rat2 = oracle_tl_info(q, sp, origin, 0.22)
    assert list(rat2[:4]) == [1, 1, 1, 1] and na2 == 0


def _direct_phi_targets(q, tgt, w, a, app, chunk=500):
    """FP64 softened potential at the members tgt from every member of q."""
    out = np.empty(len(tgt))
    for s0 in range(0, len(tgt), chunk):
        ti = tgt[s0:s0 + chunk]
        d = q[:, ti, None] - q[:, None, :]
        r = np.sqrt((d * d).sum(0))
        h = app
        e = 2 * r / h
        with np.errstate(divide="ignore", invalid="ignore"):
            inner = (208 - e ** 2 * (112 + e ** 2 * (-56 + e * (14 + e * (8 - 3 * e))))) / (70 * h)
            outer = (128 - (e ** 2 * (448 + e * (-280 + e * (56 + e * (14 + e * (-8 + e)))))
                            - 224 * e - 12 / e)) / (70 * h)
            D = np.where(r >= h, 1 / r, np.where(e <= 1, inner, outer))
        D[np.arange(len(ti)), ti] = 0.0
        out[s0:s0 + chunk] = -(a / (6 * np.pi)) * (D * w[None, :]).sum(1)
    return out


@pytest.mark.slow
def test_two_level_potential_vs_direct_within_method_error():
    # a concentrated 60,000-member halo on the 256 grid whose central block is
    # refined: ratio 2 at softening 0.1, ratio 4 at 0.05 (3.5 sp / 4 >= 0.05)
    rng = np.random.default_rng(31)
    n = 60000
    x = _nfw_halo(rng, n, 2.5, c=30.0) + 256.0
    tgt = np.sort(rng.choice(n, 3000, replace=False))
    for app, ratio in ((0.1, 2), (0.05, 4)):
        r = oracle_fof_bound(np.asfortranarray(x), [n], L, 0.6, 10, np.zeros((3, n)), app=app,
                             want_phi=True)
        assert r["ierr"] == 0 and list(r["gid"]) == [4]
        q = (x - x[:, :1]).astype(np.float32)
        lo = np.minimum(q.min(1), 0).astype(np.float32)
        sp = np.float32(r["spacing"][0])
        na, cnt, rat = oracle_tl_info(q, sp, (lo - sp).astype(np.float32), app)
        assert na >= 1 and (rat[cnt > 0] == ratio).any()
        # members are in index order (one region, one candidate of all members)
        ref = _direct_phi_targets(q.astype(np.float64), tgt, np.ones(n), 1.0, app)
        got = r["phi"][tgt].astype(np.float64)
        rel = np.abs(got - ref) / np.abs(ref)
        assert np.median(rel) < 5e-3 and rel.max() < 3e-2, (app, np.median(rel), rel.max())
