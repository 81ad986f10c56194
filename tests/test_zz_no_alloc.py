"""pyfof_fof_noalloc, the allocation-free variant of pyfof_fof, against the
allocating entry point on the same device inputs, through the C-ABI.

Two calls in a row, the second on a smaller input, so that the second runs
on the scratch the first left behind; both must equal pyfof_fof_ctx bit for
bit (in_halo, halo_id) and to rounding (ang_mom). The device's free memory
does not move across a call, oversize inputs and bound_check are refused with
their ierr and zeroed outputs, and an input whose cell grid exceeds the
scratch fails cleanly with ierr 3.
"""
import ctypes
import sys

import numpy as np
import pytest
# imported before the library loads, so that torch's HIP runtime starts first
import torch

pytestmark = pytest.mark.gpu

MAX_N = 524_288


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    from py_fof_amd import _native
    lib = _native.load()
    d, i32, i64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    lib.pyfof_fof_noalloc.restype = i32
    lib.pyfof_fof_noalloc.argtypes = [vp, vp, vp, vp, i32, d, d, i32, i32, i32, d, d, d, d, i32,
                                      i32, i64, d, vp, vp, vp]
    return lib


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda:0")


def _call_noalloc(lib, x_d, v_d, nh, L, b, np_min=10, bound=0, nt=0, fill=-7):
    n = int(nh.sum())
    in_d = torch.full((max(n, 1),), fill, dtype=torch.int32, device="cuda:0")
    id_d = torch.full((max(n, 1),), fill, dtype=torch.int32, device="cuda:0")
    am_d = torch.full((len(nh), 3), float(fill), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    ierr = lib.pyfof_fof_noalloc(x_d.data_ptr(), v_d.data_ptr() if v_d is not None else None,
                                 None, nh.ctypes.data, len(nh), L, b, np_min, bound, 0, 0.1, 1.0,
                                 0.0, 17.152, nt, n, 256, 512.0, in_d.data_ptr(),
                                 id_d.data_ptr(), am_d.data_ptr())
    return ierr, in_d[:n].cpu().numpy(), id_d[:n].cpu().numpy(), am_d.cpu().numpy().T


def _call_alloc(x_d, v_d, nh, L, b, np_min=10):
    from py_fof_amd import _native
    n = int(nh.sum())
    ctx = _native.Context(0)
    try:
        in_d = torch.empty(n, dtype=torch.int32, device="cuda:0")
        id_d = torch.empty(n, dtype=torch.int32, device="cuda:0")
        am_d = torch.empty((len(nh), 3), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        ierr = ctx.fof(x_d, v_d, None, nh, L, b, np_min, 0, 0, 0.1, 1.0, 0.0, 17.152, 0, n, 256,
                       512.0, in_d, id_d, am_d, None, _native.MEM_DEVICE)
        assert ierr == 0, _native.last_error()
        return in_d.cpu().numpy(), id_d.cpu().numpy(), am_d.cpu().numpy().T
    finally:
        ctx.close()


def _inputs(n, L, seed, regions=1):
    from py_fof_amd.synth import clumpy_plummer
    xs = [clumpy_plummer(n=n // regions, L=L, n_clumps=max(4, n // 600), seed=seed + k)
          for k in range(regions)]
    xp = np.asfortranarray(np.hstack(xs))
    vp = np.asfortranarray(np.random.default_rng(seed).standard_normal(xp.shape))
    return xp, vp, np.array([x.shape[1] for x in xs], np.int32)


def _same(a, b):
    (i1, h1, m1), (i2, h2, m2) = a, b
    assert np.array_equal(i1, i2), "in_halo differs"
    assert np.array_equal(h1, h2), "halo_id differs"
    assert np.allclose(m1, m2, rtol=1e-12, atol=1e-12 * (np.abs(m2).max() + 1e-300))


def test_noalloc_matches_allocating_entry_two_calls(lib):
    # largest accepted size first, then a smaller multi-region input
    cases = [(MAX_N, 64.0 * (MAX_N / 262_144) ** (1 / 3), 1, 21),
             (150_000, 40.0, 3, 22)]
    free_before = None
    for n, L, regions, seed in cases:
        xp, vp, nh = _inputs(n, L, seed, regions)
        _say(f"case: noalloc N={int(nh.sum())} regions={regions} L={L:.2f} seed={seed}")
        x_d, v_d = _dev(xp), _dev(vp)
        ref = _call_alloc(x_d, v_d, nh, L, 0.2)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        ierr, i1, h1, m1 = _call_noalloc(lib, x_d, v_d, nh, L, 0.2)
        assert ierr == 0, lib.pyfof_last_error()
        _same((i1, h1, m1), ref)
        assert h1.max() > 10
        if free_before is not None:
            # outputs were allocated by torch before free0; the call itself
            # takes no device memory
            assert torch.cuda.mem_get_info()[0] >= free0 - (2 << 20)
        free_before = free0
        # without velocities: no angular momentum, same partition
        ierr, i2, h2, m2 = _call_noalloc(lib, x_d, None, nh, L, 0.2)
        assert ierr == 0 and np.array_equal(i2, i1) and np.array_equal(h2, h1) and not m2.any()


def test_noalloc_refusals_zero_outputs(lib):
    xp, vp, nh = _inputs(20_000, 16.0, 31)
    x_d, v_d = _dev(xp), _dev(vp)
    ierr, i, h, m = _call_noalloc(lib, x_d, v_d, nh, 16.0, 0.2, bound=1)
    assert ierr == 1 and not i.any() and not h.any() and not m.any()
    ierr, i, h, m = _call_noalloc(lib, x_d, v_d, nh, 16.0, 0.2, nt=300)
    assert ierr == 10 and not i.any() and not h.any() and not m.any()
    # one particle over the maximum
    big = np.array([MAX_N + 1], np.int32)
    xb = torch.zeros((MAX_N + 1, 3), dtype=torch.float64, device="cuda:0")
    ierr, i, h, m = _call_noalloc(lib, xb, None, big, 16.0, 0.2)
    assert ierr == 21 and not i.any() and not h.any()
    assert b"at most" in lib.pyfof_last_error()
    # as many one-particle regions (each one halo) as particles allowed: the
    # per-region tables (geometry, boxes, chunks, cells) and the per-halo
    # angular-momentum sums exceed the scratch -> ierr 3
    rng = np.random.default_rng(3)
    xs = np.asfortranarray(rng.uniform(0.0, 100.0, (3, MAX_N)))
    nh = np.ones(MAX_N, np.int32)
    ierr, i, h, m = _call_noalloc(lib, _dev(xs), _dev(xs), nh, 512.0, 0.2, np_min=1)
    assert ierr == 3 and not i.any() and not h.any(), lib.pyfof_last_error()
    assert b"scratch" in lib.pyfof_last_error()
    # and the next call works again on the same scratch
    ierr, i, h, m = _call_noalloc(lib, x_d, v_d, _inputs(20_000, 16.0, 31)[2], 16.0, 0.2)
    assert ierr == 0
    _same((i, h, m), _call_alloc(x_d, v_d, _inputs(20_000, 16.0, 31)[2], 16.0, 0.2))
