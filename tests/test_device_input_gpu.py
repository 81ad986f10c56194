"""fof() / fof_with_info() / the server on torch tensors in HBM.

Every device-tensor call is checked against the oracle (FoF bit-exact,
unbinding on direct-summation halos bit-exact, the physical-unit conversion
restated from the reference in tests/helpers.physical_to_code) and must give
exactly what the same call on numpy arrays gives: the library computes on
the same float64 values, so in_halo, halo_id, the catalog and ang_mom are
equal bit for bit. Inputs cover the layouts a
caller can hold: a (3, N) row-major tensor (packed by one device copy), the
transposed view of an (N, 3) tensor (read in place), a list of regions,
float32 positions, and physical units with device masses.
"""
import sys

import numpy as np
import pytest
# imported before any test loads the library: the library then lets torch's
# own HIP runtime initialise first (see py_fof_amd._native._torch_first)
import torch  # noqa: F401

from helpers import physical_to_code
from oracle import oracle_fof, oracle_fof_bound

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch
    return torch


def _dev(a, layout="rows"):
    torch = _torch()
    if layout == "rows":  # (3, N) row-major: positions of one axis contiguous
        return torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")
    # (N, 3) row-major seen as (3, N): the reference's Fortran (3, N) layout
    return torch.as_tensor(np.ascontiguousarray(a.T), device="cuda:0").t()


def _host(t):
    return t.cpu().numpy()


def _box(n, seed):
    from py_fof_amd.synth import clumpy_nfw
    L = float(n) ** (1.0 / 3.0)
    xp = clumpy_nfw(n=n, L=L, seed=seed)
    vp = np.asfortranarray(np.random.default_rng(seed).normal(0.0, 0.5, xp.shape))
    return xp, vp, L


@pytest.mark.parametrize("layout", ["rows", "aos"])
def test_fof_device_tensor_equals_host(layout):
    import py_fof_amd as pf
    xp, vp, L = _box(300_000, 11)
    _say(f"case: device fof N={xp.shape[1]} layout={layout}")
    kw = dict(L_Box=L, linking_length=0.2, np_halo_min=10, return_ang_mom=True)
    h_in, h_id, h_am = pf._fof_core(xp, vp, **kw)
    xt, vt = _dev(xp, layout), _dev(vp, layout)
    if layout == "aos":
        assert xt.stride() == (1, 3)
    d_in, d_id, d_am = pf._fof_core(xt, vt, **kw)
    assert d_in.is_cuda and d_in.dtype == _torch().bool and d_id.is_cuda and d_am.is_cuda
    assert np.array_equal(_host(d_in), h_in)
    assert np.array_equal(_host(d_id), h_id)
    assert np.array_equal(_host(d_am), h_am)
    assert h_in.any() and h_id.max() > 10
    o_in, o_id, o_am = oracle_fof(xp, [xp.shape[1]], L, 0.2, 10, vp=vp, n_threads=4)
    assert np.array_equal(_host(d_in), o_in.astype(bool))
    assert np.array_equal(_host(d_id), o_id)
    assert np.allclose(_host(d_am), o_am, rtol=1e-9, atol=1e-12 * np.abs(o_am).max())
    # fof() returns the same in_halo, and no input tensor was modified
    out = pf.fof(xt, L_Box=L, linking_length=0.2, np_halo_min=10)
    assert np.array_equal(_host(out), h_in)
    assert np.array_equal(_host(xt), xp)


def test_regions_float32_and_catalog():
    import py_fof_amd as pf
    torch = _torch()
    rng = np.random.default_rng(5)
    parts = [_box(n, s)[0] for n, s in ((40_000, 1), (5_000, 2), (120_000, 3))]
    L = float(120_000) ** (1.0 / 3.0)
    parts = [np.mod(p, L) for p in parts]
    _say("case: three regions, float32 device tensors, fof_with_info")
    p32 = [p.astype(np.float32) for p in parts]
    host = pf.fof_with_info([p.astype(np.float64) for p in p32], L_Box=L, linking_length=0.2,
                            np_halo_min=8)
    dev = pf.fof_with_info([torch.as_tensor(p, device="cuda:0") for p in p32], L_Box=L,
                           linking_length=0.2, np_halo_min=8)
    assert np.array_equal(_host(dev["in_halo"]), host["in_halo"])
    x64 = np.asfortranarray(np.hstack([p.astype(np.float64) for p in p32]))
    o_in, o_id, _ = oracle_fof(x64, [p.shape[1] for p in p32], L, 0.2, 8, n_threads=4)
    assert np.array_equal(_host(dev["in_halo"]), o_in.astype(bool))
    assert dev["nhalo_found"] == int(o_id.max())
    assert dev["nhalo_found"] == host["nhalo_found"] > 0
    assert np.array_equal(dev["np_per_halo"], host["np_per_halo"])
    assert len(dev["halo_indices"]) == len(host["halo_indices"])
    for a, b in zip(dev["halo_indices"], host["halo_indices"]):
        assert a.is_cuda and np.array_equal(_host(a), b)
    del rng


def test_unbinding_physical_units_device_masses():
    import py_fof_amd as pf
    from py_fof_amd.synth import isolated_halos
    torch = _torch()
    xs, vs = isolated_halos(n_halos=30, seed=9, nmax=20000)
    # physical units: positions in Mpc/h on a 300 Mpc/h box, masses per region
    box, ng = 300.0, 512.0
    dx = box / ng
    xs_p = [np.asfortranarray(x * dx) for x in xs]
    mps = [np.full(x.shape[1], 1.0e10 * (1.0 + 0.01 * i)) for i, x in enumerate(xs)]
    kw = dict(L_Box=box, linking_length=0.2 * dx, np_halo_min=10, bound_check=True,
              return_ang_mom=True, units="physical", softening=0.1 * dx, ng_global=ng)
    _say(f"case: unbinding, physical units, {len(xs)} regions, device masses")
    h_in, h_id, h_am = pf._fof_core(xs_p, vs, mp_list=mps, **kw)
    d_in, d_id, d_am = pf._fof_core([_dev(x, "aos") for x in xs_p], [_dev(v, "aos") for v in vs],
                                    mp_list=[torch.as_tensor(m, device="cuda:0") for m in mps], **kw)
    assert np.array_equal(_host(d_in), h_in)
    assert np.array_equal(_host(d_id), h_id)
    assert np.array_equal(_host(d_am), h_am)
    assert h_id.max() > 0
    # the oracle on the code-unit inputs of the reference's conversion: every
    # region here is summed directly, so the bound sets are exact
    xp_c, vp_c, mp_c, ok = physical_to_code(xs_p, vs, np.concatenate(mps), box, 0.2 * dx,
                                            0.1 * dx, box=box, ng_global=ng)
    r = oracle_fof_bound(xp_c, [x.shape[1] for x in xs], ok["L"], ok["b"], 10, vp_c, mp=mp_c,
                         app=ok["app"], a_val=ok["a_val"], b_val=ok["b_val"], s15=ok["s15"])
    assert r["ierr"] == 0
    ref = r["in_halo"].astype(bool)
    assert int(np.sum(_host(d_in) != ref)) <= 1e-4 * len(ref)
    if np.all(r["gid"] == 0):
        assert np.array_equal(_host(d_in), ref) and np.array_equal(_host(d_id), r["halo_id"])


def test_errors_and_empty_input():
    import py_fof_amd as pf
    torch = _torch()
    xt = torch.zeros((3, 10), dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError, match="vp_list is required when bound_check=True"):
        pf.fof(xt, bound_check=True)
    with pytest.raises(ValueError, match=r"must have shape \(3, N\)"):
        pf.fof([torch.zeros((4, 10), dtype=torch.float64, device="cuda:0")])
    with pytest.raises(ValueError, match="host array among device tensors"):
        pf.fof([xt, np.zeros((3, 4))])
    with pytest.raises(ValueError, match="must be on the device of xp_list"):
        pf.fof(xt, vp_list=np.zeros((3, 10)))
    with pytest.raises(ValueError, match="All particle masses in mp_list must be positive and finite"):
        pf.fof(xt, mp_list=torch.full((10,), -1.0, dtype=torch.float64, device="cuda:0"))
    e = torch.zeros((3, 0), dtype=torch.float64, device="cuda:0")
    out, am = pf.fof(e, return_ang_mom=True)
    assert out.is_cuda and out.numel() == 0 and tuple(am.shape) == (3, 1)
    info = pf.fof_with_info(e)
    assert info["nhalo_found"] == 0 and info["halo_indices"] == []


def test_server_takes_device_tensors():
    import py_fof_amd as pf
    from py_fof_amd.serve import Server
    xp, vp, L = _box(200_000, 21)
    kw = dict(L_Box=L, linking_length=0.2, np_halo_min=10, return_ang_mom=True)
    _say("case: server jobs on device tensors")
    ref = pf._fof_core(xp, vp, **kw)
    with Server(slots=2) as srv:
        jobs = [srv.submit(_dev(xp, "aos"), _dev(vp, "aos"), **kw),
                srv.submit(_dev(xp, "rows"), _dev(vp, "rows"), **kw)]
        for j in jobs:
            inh, hid, am = j.result(with_halo_id=True)
            assert inh.is_cuda
            assert np.array_equal(_host(inh), ref[0])
            assert np.array_equal(_host(hid), ref[1])
            assert np.array_equal(_host(am), ref[2])
