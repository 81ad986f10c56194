"""Solver forms of the unbinding on the GPU (context options).

* plain (one candidate at a time) and batched runs give identical results,
  in FP64 and in f32: every kernel treats a candidate on its own;
* direct summation above the reference's 1000 members (global-memory
  kernels) in FP64 matches an independent numpy FP64 brute force;
* the f32 direct-summation form (f32 pair terms, partial sums over 256
  partners in f32 added in f64) stays within the stated potential error of
  the FP64 form: relative error of every member's first-pass potential
  <= 1e-5, RMS <= 1e-6; bound status differs for at most 0.1 % of the
  members;
* the FFT convolutions' lines per workgroup (4, 8, 16: a launch shape) give
  bit-identical potentials and members on every mesh size, two-level inner
  grids included.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 512.0


def _run(xs, vs, b=0.25, np_min=10, softening=0.1, direct_max=1000, f32=False, batched=True,
         record=False, fft_lines=None):
    import py_fof_amd as pf
    from py_fof_amd import _native
    ctx = _native.Context(0)
    for n in (48, 64, 128, 216, 256):
        if fft_lines:
            ctx.set_option(_native.OPT_FFT_LINES, 1000 * n + fft_lines)
    ctx.set_option(_native.OPT_DIRECT_MAX, direct_max)
    ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
    ctx.set_option(_native.OPT_BATCHED, int(batched))
    ctx.set_option(_native.OPT_RECORD_PHI, int(record))
    inh, hid, am = pf._fof_core(xs, vs, L_Box=L, linking_length=b, np_halo_min=np_min,
                                bound_check=True, return_ang_mom=True, softening=softening,
                                ng_global=4096.0, _ctx=ctx)
    out = {"in_halo": inh, "halo_id": hid, "ang_mom": am, "stats": ctx.stats()}
    if record:
        out["phi"], out["index"] = ctx.potential()
    ctx.close()
    return out


def _halos(sizes, seed):
    from py_fof_amd.synth import isolated_halos
    return isolated_halos(seed=seed, sizes=np.asarray(sizes), L=L)


def _same(a, b):
    assert np.array_equal(a["in_halo"], b["in_halo"])
    assert np.array_equal(a["halo_id"], b["halo_id"])
    assert np.array_equal(a["ang_mom"], b["ang_mom"])


@pytest.mark.parametrize("f32", [False, True])
def test_plain_equals_batched(f32):
    rng = np.random.default_rng(41)
    sizes = np.concatenate([rng.integers(20, 900, 40), [1500, 2600, 4000, 7000]])
    xs, vs = _halos(sizes, 41)
    a = _run(xs, vs, f32=f32, batched=True)
    b = _run(xs, vs, f32=f32, batched=False)
    _same(a, b)
    assert a["stats"]["n_pm48"] + a["stats"]["n_pm64"] + a["stats"]["n_pm128"] > 0
    # with large direct halos as well
    c = _run(xs, vs, f32=f32, batched=True, direct_max=8000)
    d = _run(xs, vs, f32=f32, batched=False, direct_max=8000)
    _same(c, d)
    assert c["stats"]["n_direct_large"] > 0


def _direct_phi_f64(q, w, a, app):
    """Textbook O(N^2) softened potential in FP64."""
    d = q[:, :, None] - q[:, None, :]
    r = np.sqrt((d * d).sum(0))
    h = app
    e = 2 * r / h
    with np.errstate(divide="ignore", invalid="ignore"):
        inner = (208 - e ** 2 * (112 + e ** 2 * (-56 + e * (14 + e * (8 - 3 * e))))) / (70 * h)
        outer = (128 - (e ** 2 * (448 + e * (-280 + e * (56 + e * (14 + e * (-8 + e))))) - 224 * e
                        - 12 / e)) / (70 * h)
        D = np.where(r >= h, 1 / r, np.where(e <= 1, inner, outer))
    np.fill_diagonal(D, 0.0)
    return -(a / (6 * np.pi)) * (D * w[None, :]).sum(1)


def test_large_direct_fp64_matches_numpy():
    from py_fof_amd.synth import isotropic, nfw_radii
    rng = np.random.default_rng(7)
    n = 3000
    x = nfw_radii(rng, n, 8.0, 2.0) * isotropic(rng, n) + 200.0
    x = np.asfortranarray(x)
    v = np.zeros((3, n))
    r = _run([x], [v], b=1.0, direct_max=4000, record=True)
    assert r["stats"]["n_direct_large"] == 1
    idx = r["index"]
    assert len(idx) == n and np.array_equal(np.sort(idx), np.arange(n))
    # positions as the solver holds them: f32 offsets from the candidate's anchor
    # (its first member in cell order); the potential is translation invariant
    q = (x - x[:, :1]).astype(np.float32).astype(np.float64)
    ref = _direct_phi_f64(q[:, idx], np.ones(n), 1.0, 0.1)
    assert np.allclose(r["phi"].astype(np.float64), ref, rtol=2e-6)


def test_f32_potential_error_and_status():
    rng = np.random.default_rng(12)
    sizes = np.concatenate([rng.integers(20, 1000, 150), [1200, 2500, 5000, 9000, 15000, 40000]])
    xs, vs = _halos(sizes, 12)
    r64 = _run(xs, vs, direct_max=65536, f32=False, record=True)
    r32 = _run(xs, vs, direct_max=65536, f32=True, record=True)
    assert np.array_equal(r64["index"], r32["index"])
    p64 = r64["phi"].astype(np.float64)
    p32 = r32["phi"].astype(np.float64)
    rel = np.abs(p32 - p64) / np.abs(p64)
    assert rel.max() <= 1e-5, rel.max()
    assert np.sqrt(np.mean(rel ** 2)) <= 1e-6
    diff = int(np.sum(r64["in_halo"] != r32["in_halo"]))
    assert diff <= 1e-3 * max(1, int(r64["in_halo"].sum())), diff


def test_no_reads_of_unwritten_device_memory():
    # buffers are reused and never cleared between calls, so a kernel that
    # read memory no pass had written (say the zero padding of a convolution
    # grid) would depend on what earlier calls left there. New contexts whose
    # allocations are filled with 0x00 and with 0xFF bytes must agree exactly,
    # for the unbinding (all four grid classes, two-level off and on) and for
    # the FoF linking (dense and sparse cells).
    from py_fof_amd import _native, synth
    import py_fof_amd as pf
    xs, vs = _halos(np.array([300, 1500, 5000, 12000, 40000, 90000]), 23)
    xc = synth.clumpy_plummer(n=200_000, L=48.0, n_clumps=120, seed=3)

    def run(fill):
        probe = _native.Context(0)
        probe.set_option(_native.OPT_POISON_ALLOC, 1 + fill)
        probe.close()
        ctx = _native.Context(0)
        try:
            ctx.set_option(_native.OPT_RECORD_PHI, 1)
            r = pf._fof_core(xs, vs, L_Box=L, linking_length=0.25, np_halo_min=10, bound_check=True,
                             return_ang_mom=True, softening=0.1, ng_global=4096.0, _ctx=ctx)
            phi = ctx.potential()
            r2 = pf._fof_core(xs[-1:], vs[-1:], L_Box=L, linking_length=0.6, np_halo_min=10,
                              bound_check=True, return_ang_mom=True, softening=0.05, ng_global=4096.0,
                              _ctx=ctx)
            f = pf._fof_core([xc], L_Box=48.0, linking_length=0.2, np_halo_min=10, _ctx=ctx)
            return [a for a in list(r) + list(phi) + list(r2) + list(f) if a is not None]
        finally:
            ctx.close()

    try:
        zero = run(0x00)
        ones = run(0xFF)
    finally:
        _native.Context(0).set_option(_native.OPT_POISON_ALLOC, 0)
    for a, b in zip(zero, ones):
        assert np.array_equal(a, b)


def test_fft_lines_per_workgroup_do_not_change_results():
    sizes = [1500, 3000, 6000, 12000, 30000, 90000]
    xs, vs = _halos(sizes, 43)
    runs = {nl: _run(xs, vs, softening=0.05, record=True, fft_lines=nl) for nl in (4, 8, 16)}
    st = runs[8]["stats"]
    assert st["n_pm256"] >= 1 and st["n_pm48"] + st["n_pm64"] + st["n_pm128"] >= 1
    for nl in (4, 16):
        _same(runs[nl], runs[8])
        assert np.array_equal(runs[nl]["phi"], runs[8]["phi"])
    with pytest.raises(RuntimeError):
        from py_fof_amd import _native
        c = _native.Context(0)
        try:
            c.set_option(_native.OPT_FFT_LINES, 100 * 1000 + 8)
        finally:
            c.close()
