"""GPU parity of the FoF path against the oracle (bit-exact partition).

Every case compares, for the same seeded input, the HIP library (through
the C-ABI, via py_fof_amd._fof_core) with the C oracle:
  * roots: per particle, the input index of the first CRS member of its
    component -- the canonical label of the whole partition;
  * in_halo and halo_id (numbering order included), exactly;
  * the fof_with_info catalog against the host construction of the reference
    (py_fof/__init__.py:481-507) applied to the oracle's halo_id;
  * per-region angular momentum to rtol 1e-9 (f64, summation order differs).
"""
import sys

import numpy as np
import pytest

from helpers import TINY_GRID_CASES, chain_halo, wrapped_clumps
from oracle import oracle_fof

import py_fof_amd
from py_fof_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_both(xps, L, b, np_min, vps=None, mps=None, am=False):
    xp = np.asfortranarray(np.hstack(xps))
    nh = [x.shape[1] for x in xps]
    vp = np.asfortranarray(np.hstack(vps)) if vps is not None else None
    mp = np.concatenate(mps) if mps is not None else None
    _say(f"case: N={xp.shape[1]} regions={len(nh)} L={L} b={b} np_min={np_min}")
    in_h, hid, ang, ex = py_fof_amd._fof_core(
        xps, vp_list=vps, L_Box=L, linking_length=b, np_halo_min=np_min,
        mp_list=mp, return_ang_mom=am, _want_roots=True, _want_catalog=True)
    o_in, o_hid, o_am, o_lab = oracle_fof(xp, nh, L, b, np_min, vp=vp, mp=mp,
                                          n_threads=2, with_labels=True)
    return (in_h, hid, ang, ex), (o_in, o_hid, o_am, o_lab)


def _dense_cells(xps, L, b):
    """Dense cells (with sub-cells) the library used for this input."""
    from py_fof_amd import _native
    ctx = _native.Context(0)
    try:
        py_fof_amd._fof_core(xps, L_Box=L, linking_length=b, np_halo_min=10, _ctx=ctx)
        return ctx.stats()["ndense"]
    finally:
        ctx.close()


def _check(gpu, ref, am=False):
    (in_h, hid, ang, ex), (o_in, o_hid, o_am, o_lab) = gpu, ref
    assert np.array_equal(ex["roots"], o_lab), "partition roots differ"
    assert np.array_equal(in_h, o_in.astype(bool)), "in_halo differs"
    assert np.array_equal(hid, o_hid), "halo_id differs"
    # catalog == host construction from halo_id (py_fof/__init__.py:489-505)
    off, idx = ex["catalog"]
    valid = np.nonzero(o_hid > 0)[0]
    order = np.argsort(o_hid[valid], kind="stable")
    assert np.array_equal(idx, valid[order])
    counts = np.bincount(o_hid[valid], minlength=o_hid.max() + 1)[1:] if len(valid) else []
    assert np.array_equal(np.diff(off), counts)
    if am:
        scale = np.abs(o_am).max() + 1e-300
        assert np.allclose(ang, o_am, rtol=1e-9, atol=1e-12 * scale)


def test_library_loads():
    lib = py_fof_amd._native.load()
    assert b"gfx950" in lib.pyfof_version()


def test_chain_halo_known_answer():
    rng = np.random.default_rng(999)
    xp, expect = chain_halo(200, 2.0, 400, rng)
    in_h, am = py_fof_amd.fof(xp, L_Box=512.0, linking_length=2.0, np_halo_min=10,
                              return_ang_mom=True)
    assert np.array_equal(in_h, expect)
    assert np.allclose(am, 0.0)


def test_multiple_regions_known_answer():
    shifts = [(-50.0, -50.0, 0.0), (50.0, -50.0, 0.0), (0.0, 50.0, 30.0)]
    xs, es = [], []
    for k, s in enumerate(shifts):
        xp, e = chain_halo(100, 2.0, 200, np.random.default_rng(777 + k))
        xs.append(xp + np.array(s)[:, None])
        es.append(e)
    in_h = py_fof_amd.fof(xs, L_Box=512.0, linking_length=2.0, np_halo_min=10)
    assert np.array_equal(in_h, np.concatenate(es))
    g, r = _run_both(xs, 512.0, 2.0, 10)
    _check(g, r)


def test_clumpy_small_periodic():
    xp = synth.clumpy_plummer(n=32768, L=32.0, n_clumps=60, seed=5)
    g, r = _run_both([xp], 32.0, 0.2, 10)
    _check(g, r)


@pytest.mark.parametrize("np_min", [1, 2, 3])
def test_clumpy_small_halo_minimum(np_min):
    # the candidate test counts a root through the root bitmap for np_min 1
    # and through size >= np_min - 1 from 2 on; isolated particles and
    # components spanning many count chunks both occur here
    xp = synth.clumpy_plummer(n=65536, L=40.0, n_clumps=80, seed=9)
    g, r = _run_both([xp], 40.0, 0.2, np_min)
    _check(g, r)


def test_clumpy_plummer_262k_bit_exact():
    xp = synth.clumpy_plummer()  # 262,144 particles, L = 64, seed 0
    g, r = _run_both([xp], 64.0, 0.2, 10)
    _check(g, r)


def test_clumpy_nfw_2m_bit_exact():
    # the benchmark's generator at 2,097,152 particles and the same density
    # (L = 128): sparse cells, dense cells of one and of several components,
    # and the sparse-dense interface between them
    xp = synth.clumpy_nfw(n=2_097_152, L=128.0, seed=2)
    g, r = _run_both([xp], 128.0, 0.2, 10)
    _check(g, r)
    assert _dense_cells([xp], 128.0, 0.2) > 100


def test_dense_clump_subcells_and_angmom():
    # one very dense clump: dense cells, clique sub-cells; velocities + masses
    rng = np.random.default_rng(21)
    n = 30000
    x = 100.0 + 0.3 * rng.standard_normal((3, n))
    x = np.concatenate([x, rng.uniform(90, 110, (3, 5000))], axis=1)
    v = rng.standard_normal(x.shape)
    m = rng.uniform(0.5, 2.0, x.shape[1])
    g, r = _run_both([np.asfortranarray(x)], 512.0, 0.05, 10, vps=[np.asfortranarray(v)],
                     mps=[m], am=True)
    _check(g, r, am=True)
    assert _dense_cells([np.asfortranarray(x)], 512.0, 0.05) > 0


def test_dense_clique_subcells_nonperiodic_and_coarse_k():
    rng = np.random.default_rng(31)
    # isolated dense blob: L_mesh clamps to b, two clique sub-cells per axis
    x = 200.0 + 0.28 * rng.standard_normal((3, 20000))
    g, r = _run_both([np.asfortranarray(x)], 512.0, 0.2, 10)
    _check(g, r)
    assert _dense_cells([np.asfortranarray(x)], 512.0, 0.2) > 0
    # periodic clumps with b = 0.35 in unit cells (k = 5 sub-cells per axis)
    xp = synth.clumpy_plummer(n=65536, L=40.0, n_clumps=50, a=0.5, seed=12)
    g, r = _run_both([xp], 40.0, 0.35, 10)
    _check(g, r)
    assert _dense_cells([xp], 40.0, 0.35) > 0


def test_nonperiodic_cells_equal_to_b():
    # the reference's perf-test geometry: L_mesh clamps to b (cells == b)
    rng = np.random.default_rng(3)
    xs = [np.asfortranarray(rng.uniform(0, 12.0, (3, 10000)) + 50.0 + 150.0 * k) for k in range(3)]
    g, r = _run_both(xs, 512.0, 2.0, 10)
    _check(g, r)


def test_many_regions_mixed():
    rng = np.random.default_rng(8)
    xs = []
    for k in range(300):
        n = int(rng.integers(1, 400))
        c = rng.uniform(0, 512.0, 3)
        xs.append(np.asfortranarray(np.mod(c[:, None] + 0.6 * rng.standard_normal((3, n)), 512.0)))
    vs = [np.asfortranarray(rng.standard_normal(x.shape)) for x in xs]
    g, r = _run_both(xs, 512.0, 0.2, 10, vps=vs, am=True)
    _check(g, r, am=True)


def test_slab_periodic_one_axis_and_tiny_grid():
    rng = np.random.default_rng(5)
    n = 6000
    xp = np.stack([rng.uniform(0, 30.0, n), rng.uniform(0, 4.0, n), rng.uniform(0, 4.0, n)])
    g, r = _run_both([np.asfortranarray(xp)], 30.0, 0.3, 5)
    _check(g, r)
    # a periodic box with only 3 cells per axis (no pruning path)
    xp = rng.uniform(0, 3.0, (3, 4000))
    g, r = _run_both([np.asfortranarray(xp)], 3.0, 0.9, 3)
    _check(g, r)


@pytest.mark.parametrize("case", range(len(TINY_GRID_CASES)))
def test_periodic_one_and_two_cells_per_axis(case):
    # the -1 and +1 neighbour of a cell are one cell (fof_fortran.f90:729-755);
    # the larger clumps put 150-450 members in a cell (mid and dense cells)
    L, b, pc, cs, nbg, sig, nd = TINY_GRID_CASES[case]
    xp = wrapped_clumps(L, pc, cs, nbg, seed=case + 1, sigma=sig)
    g, r = _run_both([xp], L, b, 5)
    _check(g, r)
    assert len(np.unique(r[3])) >= 2
    # the same clumps inside a larger region: one cell per axis only in the
    # sub-box of two regions
    xq = wrapped_clumps(L, pc, cs, nbg, seed=case + 11, sigma=sig)
    g, r = _run_both([xp, xq], L, b, 5)
    _check(g, r)


def test_edges_inclusive_wrap_and_thresholds():
    # pair at exactly b (exactly representable): the predicate is inclusive
    xp = np.asfortranarray(np.array([[1.0, 1.5, 40.0], [2.0, 2.0, 40.0], [3.0, 3.0, 40.0]]))
    in_h = py_fof_amd.fof(xp, L_Box=512.0, linking_length=0.5, np_halo_min=2)
    assert list(in_h) == [True, True, False]
    # periodic wrap across the box edge
    rng = np.random.default_rng(1)
    bg = rng.uniform(0, 16.0, (3, 3000))
    pair = np.array([[0.05, 15.95], [8.0, 8.0], [8.0, 8.0]])
    xp = np.asfortranarray(np.concatenate([pair, bg], axis=1))
    g, r = _run_both([xp], 16.0, 0.2, 2)
    _check(g, r)
    assert r[3][0] == r[3][1]
    # np_halo_min <= 1: every particle is a halo; np_min larger than a region
    xs = [np.asfortranarray(rng.uniform(0, 5, (3, 50))), np.asfortranarray(rng.uniform(20, 25, (3, 7)))]
    for npmin in (0, 1, 8):
        g, r = _run_both(xs, 512.0, 0.5, npmin)
        _check(g, r)
    # identical positions
    xp = np.asfortranarray(np.repeat(np.array([[3.0], [4.0], [5.0]]), 100, axis=1))
    g, r = _run_both([xp], 512.0, 0.2, 10)
    _check(g, r)


def test_empty_inputs():
    res = py_fof_amd.fof_with_info(np.zeros((3, 0)), L_Box=512.0)
    assert res["nhalo_found"] == 0 and len(res["in_halo"]) == 0
    xp = np.asfortranarray(np.random.default_rng(0).uniform(0, 50, (3, 20)))
    res = py_fof_amd.fof_with_info(xp, L_Box=512.0, linking_length=0.01)
    assert res["nhalo_found"] == 0 and res["in_halo"].sum() == 0


def test_fof_with_info_consistency():
    xp = synth.clumpy_plummer(n=20000, L=20.0, n_clumps=30, seed=9)
    res = py_fof_amd.fof_with_info(xp, L_Box=20.0, linking_length=0.2)
    in_h = py_fof_amd.fof(xp, L_Box=20.0, linking_length=0.2)
    assert np.array_equal(res["in_halo"], in_h)
    allidx = np.concatenate(res["halo_indices"])
    assert len(allidx) == len(np.unique(allidx)) == in_h.sum()
    assert np.sum(res["np_per_halo"]) == in_h.sum()
    assert res["nhalo_found"] == len(res["halo_indices"]) == len(res["bound_status"])


def test_repeat_calls_identical():
    xp = synth.clumpy_plummer(n=50000, L=36.0, n_clumps=80, seed=4)
    a = py_fof_amd._fof_core(xp, L_Box=36.0, linking_length=0.2, _want_roots=True)
    b = py_fof_amd._fof_core(xp, L_Box=36.0, linking_length=0.2, _want_roots=True)
    assert np.array_equal(a[3]["roots"], b[3]["roots"])
    assert np.array_equal(a[1], b[1])


def _near_threshold_pairs(rng, b, npairs, eps, origin, spacing):
    """Pairs of particles at separations b (1 + u), u uniform in [-eps, eps],
    in random directions, the pairs spaced far apart on a lattice."""
    k = int(np.ceil(npairs ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), 0).reshape(3, -1)[:, :npairs]
    p = origin + spacing * g.astype(np.float64)
    d = rng.normal(size=(3, npairs))
    d /= np.linalg.norm(d, axis=0)
    q = p + d * (b * (1 + rng.uniform(-eps, eps, npairs)))
    return np.asfortranarray(np.concatenate([p, q], axis=1))


def test_near_threshold_pairs_bit_exact():
    # pairs placed at b (1 +/- eps): both outcomes of the inclusive f64
    # comparison occur, down to separations a few ulps of f32 from b
    rng = np.random.default_rng(2024)
    for b, eps in ((0.2, 3e-6), (0.2, 3e-7), (1.7, 3e-6), (0.013, 2e-6)):
        xp = _near_threshold_pairs(rng, b, 20000, eps, 3 * b, 3 * b)
        L = float(np.ceil(xp.max() + 3 * b))
        g, r = _run_both([xp], L, b, 2)
        _check(g, r)
        n = xp.shape[1] // 2
        linked = int(np.sum(r[3][:n] == r[3][n:]))
        assert 0 < linked < n, linked  # both outcomes occur
    # a small periodic box where most pairs are near the minimum-image wrap
    xp = np.asfortranarray(rng.uniform(0.0, 2.0, (3, 4000)))
    g, r = _run_both([xp], 2.0, 0.3, 2)
    _check(g, r)
